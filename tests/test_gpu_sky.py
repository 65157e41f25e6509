"""The Tregenza dome (irtt/sky.py) traced through the transport operator: a
black horizontal plate under the dome absorbs, from each collimated patch of
flux density v_i arriving along -d_i, v_i A cos(theta_i) -- the dome's quads
are large enough that every beam covers the plate -- so the plate's absorbed
power is A sum_i v_i max(0, d_i . z), within the Monte-Carlo error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_dome_irradiance_on_a_plate(irtt):
    from irtt import scenes, sky
    xa, ta = scenes.grid_patch([0, 0, 0], [4, 0, 0], [0, 4, 0], 8, 8)   # 16 m^2 facing +z
    sd0 = scenes._uniform_scene("plate", xa, ta, rho_d=0.0)
    rng = np.random.default_rng(3)
    vals = rng.uniform(50.0, 400.0, 145)
    sd = sky.add_sky(sd0, vals, up="z")
    dirs = sd.meta["sky"]["directions"]
    want = 16.0 * np.sum(vals * np.maximum(0.0, dirs[:, 2]))
    print(f"[sky] plate 4x4 m under the dome, {sd.nt} triangles", flush=True)
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.set_vertex_state(sd.kelvin, sd.fixed)
        st = s.build_transport(threads_per_tri=4096, rays_per_thread=8, batch_seed=9)
        T = s.transport().astype(np.float64)
    sv = sd.meta["sky"]["vertices"]
    pv = np.arange(sd0.nv)
    x4 = sd.kelvin.astype(np.float64) ** 4
    got = -(T[np.ix_(pv, sv)] @ x4[sv]).sum()
    # photons reaching the plate: about A cos / quad area of each patch's photons
    n_hit = st["absorbed"]
    rel = abs(got - want) / want
    print(f"[sky] absorbed {got:.2f} W vs {want:.2f} W (rel {rel:.2e}, {n_hit} absorbed photons)")
    assert rel < 5.0 / np.sqrt(max(n_hit, 1)) + 1e-6
