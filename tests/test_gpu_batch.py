"""Many steady and adjoint solves sharing one operator (irtt_solve_batch,
irtt_adjoint_solve_batch) against FP64 solves of the same FP32 operator.
Steady state without conduction is linear in y = x^4, so the FP64 reference
is exact: y = T^-1 b.  Tolerance: 1e-6 K (FP32 and FP16 forms alike: the
FP16 form's Newton iteration finishes on FP32-operator residuals)."""
import numpy as np
import pytest

from oracle import physics, solver

pytestmark = pytest.mark.gpu

TA = 293.0


def _city(irtt):
    from irtt import scenes
    sd = scenes.city_block(seed=6, nx=2, ny=2, target_tris=2500)
    s = irtt.Scene.from_data(sd, irtt.BIN_VERTEX)
    s.build_transport(threads_per_tri=128, rays_per_thread=4, batch_seed=2)
    return sd, s


def _sources(sd, R, seed):
    src = physics.bins(sd, 0)[3]
    rng = np.random.default_rng(seed)
    return src[None, :] * (0.2 + 1.6 * rng.random((R, 1))) + 20.0 * rng.random((R, src.size))


@pytest.mark.parametrize("op,R", [(0, 3), (2, 40)])
def test_batch_steady_vs_fp64(irtt, op, R):
    sd, s = _city(irtt)
    with s:
        T = s.transport().astype(np.float64)
        cs = s.column_sums()
        s.select_operator(op)
        src = _sources(sd, R, R)
        p = irtt.solver_params(ambient_kelvin=TA, x_tol=1e-9, f_tol=1e-18, lin_rtol=1e-12)
        X, st = s.solve_batch(src, params=p)
        x1, _ = s.solve_batch(src[R - 1:], params=p)
    n = T.shape[0]
    err = 0.0
    for k in range(R):
        b = solver.rhs(src[k], cs, TA)
        x64 = solver.steady_linear(T, b, np.zeros(n, bool), np.zeros(n))
        err = max(err, np.abs(X[k] - x64).max())
    solo = np.abs(x1[0] - X[R - 1]).max()
    print(f"op={op} R={R}: max |dT| vs FP64 {err:.2e} K, column alone vs in the batch "
          f"{solo:.1e} K; {st}")
    assert st["converged"] and err < 1e-6 and solo < 1e-9


@pytest.mark.parametrize("op", [0, 2])
def test_adjoint_batch_vs_fp64(irtt, op):
    sd, s = _city(irtt)
    rng = np.random.default_rng(9)
    with s:
        T = s.transport().astype(np.float64)
        s.select_operator(op)
        n = T.shape[0]
        R = 33
        x = 270.0 + 60.0 * rng.random((R, n))
        rhs = rng.normal(size=(R, n))
        lam, st = s.adjoint_solve_batch(x, rhs, lin_rtol=1e-12)
    rel = 0.0
    for k in range(R):
        J = T * (4.0 * x[k] ** 3)[None, :]
        l64 = np.linalg.solve(J.T, rhs[k])
        rel = max(rel, np.abs(lam[k] - l64).max() / np.abs(l64).max())
    print(f"op={op}: adjoint batch max relative error {rel:.2e}; {st}")
    assert rel < 1e-8
