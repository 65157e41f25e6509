"""GPU adjoint parity: the transposed solve against FP64 dense solves, the
adjoint photon trace against the CPU oracle's restatement on the same photons,
the reference's NOP test (T + p 1e-8 I, exact), and the full forward +
adjoint gradient against central finite differences with common random
numbers."""
import numpy as np
import pytest

from oracle import physics, solver

pytestmark = pytest.mark.gpu

SIG = 5.670374419e-8


def _restricted_jac_T(T, x, fixed, jac, dt=300.0):
    """(J_FF)^T for the exact (0) or reference (1) steady Jacobian."""
    f = ~fixed
    D = 4.0 * x ** 3
    J = T * D[None, :] if jac == 0 else D[:, None] * T + np.eye(len(x)) / dt
    return J[np.ix_(f, f)].T


@pytest.mark.parametrize("jac", [0, 1])
def test_adjoint_solve_vs_fp64(irtt, jac):
    from irtt import scenes
    sd = scenes.parallel_plates(n=8, emissivity=0.9)
    sd.fixed[:20] = 1
    sd.kelvin[:] = 300.0
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=256, rays_per_thread=4, batch_seed=2)
        sp = irtt.solver_params(jacobian=jac, ambient_kelvin=293.0, x_tol=1e-7, f_tol=1e-14)
        s.solve(sp)
        x = s.temperatures()
        T = s.transport().astype(np.float64)
        rng = np.random.default_rng(0)
        rhs = rng.normal(size=s.n_bins)
        lam, st = s.adjoint_solve(rhs, params=sp)
    fixed = sd.fixed.astype(bool)
    JT = _restricted_jac_T(T, x, fixed, jac)
    lam64 = np.zeros_like(x)
    lam64[~fixed] = np.linalg.solve(JT, rhs[~fixed])
    err = np.abs(lam - lam64).max() / np.abs(lam64).max()
    print(f"jac={jac} rel err {err:.2e}, {st}")
    assert np.all(lam[fixed] == 0.0)
    assert err < 1e-7


def test_nop_parameter_exact(irtt):
    """thermal_params.hpp:102-108 / thermal_renderer.cpp:974-978,1077-1087:
    T <- T + p 1e-8 I; dL/dp = -1e-8 lambda^T x^4 exactly (no Monte Carlo)."""
    from irtt import scenes
    sd = scenes.parallel_plates(n=8, emissivity=0.9)
    with irtt.Scene.from_data(sd, irtt.BIN_TRIANGLE) as s:
        s.build_transport(threads_per_tri=256, rays_per_thread=4, batch_seed=4)
        sp = irtt.solver_params(ambient_kelvin=293.0, x_tol=1e-8, f_tol=1e-16)
        s.solve(sp)
        x = s.temperatures()
        T = s.transport().astype(np.float64)
        b = solver.rhs(physics.bins(sd, 1)[3], s.column_sums(), 293.0)
        w = physics.bins(sd, 1)[0]
        target = x + 1.0
        dLdx = w * (x - target)              # L = 1/2 sum w (x - x*)^2, thermal_objective.cpp:47-57
        lam, _ = s.adjoint_solve(dLdx, params=sp)
    adj = -1e-8 * lam @ x ** 4
    fixed = np.zeros(len(x), bool)

    def L(p):
        xp = solver.steady_linear(T + p * 1e-8 * np.eye(len(x)), b, fixed, np.zeros(len(x)))
        return 0.5 * np.sum(w * (xp - target) ** 2)

    # x moves ~2e3 K per unit p here, so the step must keep the second-order
    # term (relative to the 1 K target offset) small
    h = 1e-6
    fd = (L(h) - L(-h)) / (2 * h)
    print(f"NOP test: adjoint {adj:.8e} FD {fd:.8e}")
    assert abs(adj - fd) < 1e-4 * abs(fd)


@pytest.mark.parametrize("which,flags", [("plates", 7), ("city_small", 7), ("city_small", 23)])
def test_adjoint_trace_matches_oracle(irtt, oracle, which, flags):
    """Same photons, same lambda and x^4: GPU and CPU accumulate the same FP32
    terms (in different orders, in FP64).  flags 23 adds the edge term's
    occluding side (IRTT_ADJ_EDGE_GAIN, an extension)."""
    from irtt import scenes
    sd = {"plates": lambda: scenes.parallel_plates(n=4, emissivity=0.8),
          "city_small": lambda: scenes.city_block(seed=2, nx=2, ny=2, target_tris=1500)}[which]()
    rng = np.random.default_rng(1)
    lam = rng.normal(size=sd.nv)
    x4 = (290.0 + 10 * rng.random(sd.nv)) ** 4
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=32, rays_per_thread=4, batch_seed=9)
        g, st = s.adjoint_trace(lam, x4, flags=flags)
    sc = oracle.OracleScene.from_irtt(sd)
    og, ost = oracle.adjoint(sc, oracle.make_params(32, 4, batch_seed=9, bin_mode=0), lam, x4,
                             flags=flags)
    for k in ("photons", "absorbed", "segments", "continuous", "edge_hits", "occluded", "rays"):
        assert st[k] == ost[k], (k, st[k], ost[k])
    scale = np.abs(og).max()
    err = np.abs(g - og).max() / scale
    print(f"{which}: rel err {err:.2e}; {st}")
    assert err < 1e-9


@pytest.mark.parametrize("which", ["blocker", "blocker_open", "fin", "balcony_sun"])
@pytest.mark.parametrize("flags", [7, 23, 39, 71, 103, 231])
def test_adjoint_trace_edge_scenes_match_oracle(irtt, oracle, which, flags):
    """The scenes that pin the edge term against finite differences
    (tests/test_oracle_edge_term.py): an opaque strip between two plates,
    the same without the receiver (occlusion rays escape), a fin reached at
    grazing incidence, a balcony's shadow on a wall under a collimated sun.
    Every term and extension (16 edge gain, 32 hit-area Jacobian, 64
    projected edge factor, 128 beam transfer of collimated first segments),
    GPU = oracle on the same photons."""
    from irtt import scenes
    sd = {"blocker": lambda: scenes.blocker_plates(),
          "blocker_open": lambda: scenes.blocker_plates(receiver=False),
          "fin": lambda: scenes.fin_plates(),
          "balcony_sun": lambda: scenes.balcony_wall(sun_triangles=128)}[which]()
    rng = np.random.default_rng(5)
    lam = 1.0 + rng.random(sd.nv)
    x4 = rng.random(sd.nv)
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=64, rays_per_thread=8, batch_seed=5)
        g, st = s.adjoint_trace(lam, x4, flags=flags)
    og, ost = oracle.adjoint(oracle.OracleScene.from_irtt(sd),
                             oracle.make_params(64, 8, batch_seed=5, bin_mode=0), lam, x4,
                             flags=flags)
    for k in ("photons", "absorbed", "segments", "continuous", "edge_hits", "occluded", "rays"):
        assert st[k] == ost[k], (k, st[k], ost[k])
    assert ost["occluded"] > 0 or which == "blocker_open"
    err = np.abs(g - og).max() / np.abs(og).max()
    print(f"{which} flags {flags}: rel err {err:.2e}; {st}")
    assert err < 1e-9


@pytest.mark.parametrize("flags", [7, 231])
def test_adjoint_trace_sunlit_facade_matches_oracle(irtt, oracle, flags):
    """The design scene (a facade with balconies and fins, here 2 buildings
    and 25k triangles) under the collimated sun quad: the GPU trace equals
    the oracle's on the same photons, with the reference's terms and with
    every extension for this case (hit area, projected edge, beam transfer);
    lambda and x4 as the design gradient's first row (lambda on the ground,
    x^4 of the scene and the sun)."""
    from irtt import scenes
    sd = scenes.facade(nx=2, ny=1, hmin=10.0, hmax=12.0, target_tris=20000,
                       n_fin_buildings=2, seed=3, sun_model="emitter")
    lam = np.asarray(sd.meta["ground_vertex"], np.float64) + 0.1
    x4 = sd.kelvin.astype(np.float64) ** 4
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=16, rays_per_thread=8, batch_seed=5,
                          layout=irtt.LAYOUT_SPARSE)
        g, st = s.adjoint_trace(lam, x4, flags=flags)
    og, ost = oracle.adjoint(oracle.OracleScene.from_irtt(sd),
                             oracle.make_params(16, 8, batch_seed=5, bin_mode=0), lam, x4,
                             flags=flags)
    for k in ("photons", "absorbed", "segments", "continuous", "edge_hits", "occluded", "rays"):
        assert st[k] == ost[k], (k, st[k], ost[k])
    err = np.abs(g - og).max() / np.abs(og).max()
    print(f"sunlit facade flags {flags}: rel err {err:.2e}; {st}")
    assert err < 1e-9


def _mean_b(irtt, sd, dz, nva, seed, photons):
    s2 = type(sd)(**{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in sd.__dict__.items()})
    s2.xyz[nva:, 2] += dz
    with irtt.Scene.from_data(s2, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=photons, rays_per_thread=8, batch_seed=seed)
        s.solve(irtt.solver_params(ambient_kelvin=293.0, x_tol=1e-8, f_tol=1e-16))
        x = s.temperatures()
        area, _, _ = s.bin_data()
    wb = np.zeros(len(x))
    wb[nva:] = area[nva:] / area[nva:].sum()
    return wb @ x


def test_full_gradient_vs_fd_plate_separation(irtt):
    """L = mean temperature of plate B; parameter = height of plate B.
    Adjoint: lambda from J^T lambda = dL/dx, then the trace with two history
    rows: (lambda, x^4) for T x^4 and (-Ta^4, lambda) for the ambient term
    Ta^4 colsum(T).  FD: central differences of the full forward solve with
    the same photons (common random numbers)."""
    from irtt import scenes
    sd = scenes.parallel_plates(n=8, emissivity=0.9)
    nva = sd.meta["plate_a_verts"]
    photons = 2048
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=photons, rays_per_thread=8, batch_seed=21)
        sp = irtt.solver_params(ambient_kelvin=293.0, x_tol=1e-8, f_tol=1e-16)
        s.solve(sp)
        x = s.temperatures()
        area, _, _ = s.bin_data()
        dLdx = np.zeros(len(x))
        dLdx[nva:] = area[nva:] / area[nva:].sum()
        lam, _ = s.adjoint_solve(dLdx, params=sp)
        Ta4 = 293.0 ** 4
        g, st = s.adjoint_trace(np.stack([lam, np.full(len(x), -Ta4)]), np.stack([x ** 4, lam]))
    adj = g[nva:, 2].sum()
    h = 5e-3
    fd = (_mean_b(irtt, sd, h, nva, 21, photons) - _mean_b(irtt, sd, -h, nva, 21, photons)) / (2 * h)
    print(f"dL/dz: adjoint {adj:.6f} K/m, FD {fd:.6f} K/m, ratio {adj / fd:.4f}; {st}")
    assert np.sign(adj) == np.sign(fd)
    assert abs(adj - fd) < 0.15 * abs(fd)


def test_project_gradient(irtt):
    g = np.arange(12, dtype=np.float64).reshape(4, 3)
    # parameter 0 moves vertices 1 and 3 along z, parameter 1 moves vertex 2 along x
    dp = irtt.project_gradient(g, [1, 3, 2], [0, 0, 1], [[0, 0, 1], [0, 0, 1], [1, 0, 0]], 2)
    assert np.array_equal(dp, [g[1, 2] + g[3, 2], g[2, 0]])


def test_adjoint_with_absorber_record_equals_two_passes(irtt):
    """IRTT_TRANSPORT_KEEP_ABSORBERS: the forward build records each photon's
    absorbing bin and the adjoint skips its first pass; the gradient and every
    counter equal the two-pass replay's (FP64 atomics reorder: 1e-12)."""
    from irtt import scenes
    sd = scenes.city_block(seed=2, nx=2, ny=2, target_tris=1500)
    rng = np.random.default_rng(3)
    lam = rng.normal(size=sd.nv)
    x4 = (290.0 + 10 * rng.random(sd.nv)) ** 4
    out = []
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        for flags in (0, irtt.TRANSPORT_KEEP_ABSORBERS):
            s.build_transport(threads_per_tri=32, rays_per_thread=4, batch_seed=9, flags=flags)
            out.append(s.adjoint_trace(np.stack([lam, -lam]), np.stack([x4, x4 * 0.5]),
                                       flags=23))
    (g0, s0), (g1, s1) = out
    for k in ("photons", "absorbed", "segments", "continuous", "edge_hits", "occluded", "rays"):
        assert s0[k] == s1[k], (k, s0[k], s1[k])
    err = np.abs(g0 - g1).max() / np.abs(g0).max()
    print(f"record vs two passes: rel diff {err:.1e}; {s1}")
    assert err < 1e-12
