"""Forward history and the backward adjoint over it (irtt_simulate,
irtt_adjoint_history; thermal_renderer.cpp:958-1133).

Finite differences in a triangle's external heat flux leave the operator bit
for bit unchanged (the flux does not enter the photon transport), so the
discrete adjoint must match them to FD accuracy:
    dL/dq_j = A_j * sum_t lambda_history[t][j]
(r_t = M (x_t - x_{t-1}) + dt (T x_t^4 + L x_t - b_t) and lambda_history
carries dt and the 1/steps objective scale).  Steady: dL/dq_j = A_j lambda_j.
The transient history is also checked step by step against an FP64 Newton of
the same operator."""
import numpy as np
import pytest

from oracle import solver

pytestmark = pytest.mark.gpu


def _plates(flux_delta=None):
    from irtt import scenes
    sd = scenes.parallel_plates(n=6, emissivity=0.9)
    sd.heat_capacity[:] = 2.0e4
    sd.density[:] = 1.0
    sd.thickness[:] = 1.0
    sd.kelvin[:] = 290.0
    if flux_delta is not None:
        j, h = flux_delta
        sd.heat_flux[j] += h
    return sd


def _run(irtt, sd, steps, steady, w, target, grad=False):
    with irtt.Scene.from_data(sd, irtt.BIN_TRIANGLE) as s:
        s.set_bin_state(np.full(sd.nt, 290.0, np.float32), np.zeros(sd.nt, np.uint8))
        s.build_transport(threads_per_tri=256, rays_per_thread=8, batch_seed=11)
        p = irtt.solver_params(steady=steady, ambient_kelvin=293.0, dt=600.0, x_tol=1e-10,
                               f_tol=1e-20, lin_rtol=1e-13)
        H, _ = s.simulate(steps, params=p)
        lam, L, g, st = s.adjoint_history(H, w, target, params=p, gradient=grad)
        T = s.transport().astype(np.float64)
        cs = s.column_sums()
        area, _, mass = s.bin_data()
    return H, lam, L, g, T, cs, area, mass


@pytest.mark.parametrize("steady,steps", [(True, 1), (False, 4)])
def test_adjoint_history_matches_flux_fd(irtt, steady, steps):
    sd = _plates()
    nt = sd.nt
    from irtt import scenes
    area = sd.triangle_areas()
    w = 1e-2 * area                  # objective weight x area
    target = np.full(nt, 300.0)
    H, lam, L, _, T, cs, _, mass = _run(irtt, sd, steps, steady, w, target)
    rng = np.random.default_rng(4)
    js = rng.choice(nt, size=4, replace=False)
    # W/m^2: 0.1 moves a plate-B bin's input by ~3e-4 relative, so the
    # central difference's truncation error (~(3e-4)^2) stays far below 1e-5
    h = 0.1
    for j in js:
        Lp = _run(irtt, _plates((j, h)), steps, steady, w, target)[2]
        Lm = _run(irtt, _plates((j, -h)), steps, steady, w, target)[2]
        fd = (Lp - Lm) / (2 * h)
        adj = area[j] * lam[:, j].sum()
        print(f"steady={steady} steps={steps} tri {j}: adjoint {adj:.9e}, FD {fd:.9e}, "
              f"ratio {adj / fd:.8f}")
        assert abs(adj - fd) <= 1e-5 * abs(fd)
    if not steady:
        # each step against an FP64 Newton of the same operator
        src = sd.heat_flux.astype(np.float64) * area
        b = solver.rhs(src, cs, 293.0)
        x = np.full(nt, 290.0)
        for t in range(steps):
            x, _ = solver.newton(T, b, np.zeros(nt, bool), np.zeros(nt), x, mass=mass, xprev=x,
                                 dt=600.0, steady=False, tol=1e-12)
            err = np.abs(H[t] - x).max()
            print(f"step {t}: max |dT| vs FP64 {err:.2e} K")
            assert err < 1e-6


def test_adjoint_history_gradient_runs_the_trace(irtt):
    """With grad_x requested the trace consumes the history rows; the steady
    single-step gradient equals the trace called by hand with the same rows
    plus the source term lambda^T db/dv (irtt_source_gradient)."""
    sd = _plates()
    area = sd.triangle_areas()
    w = 1e-2 * area
    target = np.full(sd.nt, 300.0)
    with irtt.Scene.from_data(sd, irtt.BIN_TRIANGLE) as s:
        s.build_transport(threads_per_tri=128, rays_per_thread=4, batch_seed=3)
        p = irtt.solver_params(ambient_kelvin=293.0, x_tol=1e-10, f_tol=1e-20)
        H, _ = s.simulate(1, params=p)
        lam, L, g, st = s.adjoint_history(H, w, target, params=p, gradient=True)
        g2, _ = s.adjoint_trace(np.stack([lam[0], np.full(sd.nt, -293.0 ** 4)]),
                                np.stack([H[0] ** 4, lam[0]]))
        g2 = g2 + s.source_gradient(lam[0])
    scale = np.abs(g2).max()
    print(f"|grad| max {scale:.3e}, objective {L:.6e}")
    assert np.abs(g - g2).max() <= 1e-12 * scale
