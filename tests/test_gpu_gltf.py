"""The reference's own city-block scene (assets/scenes/city-block.gltf, from
the committed fixture tests/golden/city_block_gltf.npz) through the GPU path:
transport bit-exact against the oracle's accumulators and counters, steady
solve with the fixed 145-patch sky and cotangent conduction against an FP64
Newton of the same operator, sparse layout identical to the dense one."""
import json
import os

import numpy as np
import pytest

from oracle import physics, solver

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "city_block_gltf.npz")


def _scene():
    from irtt import scenes
    fx = dict(np.load(FIXTURE, allow_pickle=False))
    sd = scenes._uniform_scene("city-block", fx["xyz"], fx["idx"])
    for k in ("rho_d", "rho_s", "thickness", "density", "heat_capacity", "conductivity",
              "flags", "kelvin", "fixed"):
        setattr(sd, k, np.array(fx[k]))
    sd.heat_flux[:] = 0.0
    sd.meta["objects"] = json.loads(str(fx["objects"]))
    return sd


def test_reference_scene_transport_bitexact(irtt, oracle):
    sd = _scene()
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        st = s.build_transport(threads_per_tri=16, rays_per_thread=4, batch_seed=1234)
        acc, Q = s.transport_fixed()
    sc = oracle.OracleScene.from_irtt(sd)
    p = oracle.make_params(16, 4, batch_seed=1234, bin_mode=0)
    oacc, _, ost = oracle.transport(sc, p)
    print(f"city-block.gltf: {sd.nv} bins, {st['photons']} photons, {st['rays']} segments, "
          f"absorbed {st['absorbed']}, missed {st['miss']}, backface {st['backface']}")
    for k in ("photons", "rays", "absorbed", "miss", "too_close", "backface", "degenerate",
              "forced_absorb", "node_visits", "tri_tests"):
        assert st[k] == ost[k], (k, st[k], ost[k])
    assert np.array_equal(acc, oacc)


def test_reference_scene_solve_vs_fp64(irtt):
    """Sky patches fixed at their hour-averaged temperatures, conduction by
    the cotangent Laplacian (k x thickness per object), no ambient term."""
    sd = _scene()
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=32, rays_per_thread=8, batch_seed=1234)
        T = s.transport().astype(np.float64)
        sp = irtt.solver_params(ambient_kelvin=-1.0, x_tol=1e-8, f_tol=1e-16, lin_rtol=1e-12)
        st = s.solve(sp)
        x = s.temperatures()
        layout = []
        for lay in (irtt.LAYOUT_DENSE, irtt.LAYOUT_SPARSE):
            s.build_transport(threads_per_tri=32, rays_per_thread=8, batch_seed=1234, layout=lay)
            layout.append(s.transport_sparse())
            s.set_vertex_state(sd.kelvin, sd.fixed)
            s.solve(sp)
            layout.append(s.temperatures())
    L = physics.cotan_laplacian(sd.xyz, sd.idx, sd.conductivity * sd.thickness, sd.nv)
    fixed = sd.fixed.astype(bool)
    k0 = sd.kelvin.astype(np.float64)
    x64, _ = solver.newton(T, np.zeros(sd.nv), fixed, k0, k0, lap=L, tol=1e-12)
    err = np.abs(x - x64).max()
    free = ~fixed
    print(f"max |dT| vs FP64 {err:.2e} K; free bins {x[free].min():.1f}..{x[free].max():.1f} K; "
          f"{st}")
    assert np.allclose(x[fixed], k0[fixed], rtol=0, atol=1e-4)
    assert st["converged"] and err < 1e-6
    (d_rows, xd), (s_rows, xs) = layout[0:2], layout[2:4]
    for a, b in zip(d_rows, s_rows):
        assert np.array_equal(a, b)
    assert np.abs(xd - xs).max() < 1e-9
