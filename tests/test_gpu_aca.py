"""Cross approximation on the device against the FP64 restatement
(oracle/aca.py, thermal_transport.cpp:281-597), and the solver on the
compressed and FP16 operator forms against FP64 solves of the same operators.

Factors and pivots: bit-exact (both sides sum in the reference's order).
Temperatures: max |dT| <= 1e-4 K against the FP64 solve of the operator the
device uses (the stated bar for the whole path is 0.05 K)."""
import numpy as np
import pytest

from oracle import aca, physics, solver

pytestmark = pytest.mark.gpu

SIGMA = 5.670374419e-8


def _scene(irtt, kind="plates", mode=0, threads=64, seed=0):
    from irtt import scenes
    if kind == "plates":
        sd = scenes.parallel_plates(emissivity=0.9)
    elif kind == "box":
        sd = scenes.closed_box(4, emissivity=0.85)
    else:
        sd = scenes.city_block(seed=1, nx=2, ny=2, target_tris=2000)
    s = irtt.Scene.from_data(sd, mode)
    s.build_transport(threads_per_tri=threads, rays_per_thread=8, batch_seed=seed)
    return sd, s


def _oracle_factors(Toff, algorithm, k):
    if algorithm == 0:
        return aca.aca_full(Toff, k)
    if algorithm == 1:
        return aca.aca_partial(Toff, k)
    return aca.aca_plus(Toff, k, seed=1)


@pytest.mark.parametrize("kind,algorithm,k", [("plates", 1, 24), ("plates", 0, 12),
                                              ("plates", 2, 16), ("city", 1, 50),
                                              ("city", 2, 20)])
def test_aca_bit_exact_vs_restatement(irtt, kind, algorithm, k):
    sd, s = _scene(irtt, kind)
    with s:
        acc, Q = s.transport_fixed()
        st = s.compress(algorithm=algorithm, rank=k)
        Ut, V, C, pr, pc = s.aca_factors()
        _, ea, _ = s.bin_data()
    Toff = aca.offdiag_operator(acc, Q)
    oUt, oV, opr, opc = _oracle_factors(Toff, algorithm, k)
    print(f"{kind} n={Toff.shape[0]} algorithm={algorithm} k={k}: pivot rows {list(pr[:8])}..., "
          f"err_inf_rel {st['err_inf_rel']:.3e}, err_frob_rel {st['err_frob_rel']:.3e}, "
          f"{st['compress_ms']:.2f} ms")
    assert list(pr) == list(opr[:k]) and list(pc) == list(opc[:k])
    np.testing.assert_array_equal(Ut, oUt)
    np.testing.assert_array_equal(V, oV)
    e = Toff - aca.lowrank(oUt, oV)
    assert st["err_inf"] == np.abs(e).max()
    assert abs(st["err_frob"] / np.linalg.norm(e) - 1) < 1e-12
    assert abs(st["err_frob_rel"] / (np.linalg.norm(e) / np.linalg.norm(Toff)) - 1) < 1e-12
    np.testing.assert_array_equal(C, SIGMA * ea)


def test_column_sum_correction_bit_exact(irtt):
    """A closed box absorbs everything, so columns of T sum to ~0 and the
    reference's correction (thermal_transport.cpp:309-358) raises them."""
    sd, s = _scene(irtt, "box")
    with s:
        acc, Q = s.transport_fixed()
        st = s.compress(rank=10, column_sum_correction=True)
        Ut, V, C, _, _ = s.aca_factors()
        _, ea, _ = s.bin_data()
    Toff = aca.offdiag_operator(acc, Q)
    oUt, oV, _, _ = aca.aca_partial(Toff, 10)
    Cn, mask = aca.column_sum_correction(oUt, oV, SIGMA * ea)
    print(f"corrected {st['n_corrected']} of {len(C)} columns")
    assert st["n_corrected"] == int(mask.sum()) > 0
    np.testing.assert_array_equal(C, Cn)


def _effective_aca(Ut, V, C):
    """The operator the solver multiplies with: FP32 factors, FP64 sums."""
    return aca.lowrank(Ut.astype(np.float32).astype(np.float64),
                       V.astype(np.float32).astype(np.float64)) + np.diag(C)


@pytest.mark.parametrize("jac", [0, 1])
def test_solve_on_aca_operator(irtt, jac):
    sd, s = _scene(irtt, "plates", mode=1, threads=256)
    Ta = 293.0
    with s:
        T = s.transport().astype(np.float64)
        b = solver.rhs(physics.bins(sd, 1)[3], s.column_sums(), Ta)
        p = irtt.solver_params(jacobian=jac, ambient_kelvin=Ta, x_tol=1e-7, f_tol=1e-16)
        s.solve(p)
        x_dense = s.temperatures()
        s.compress(rank=40)
        Ut, V, C, _, _ = s.aca_factors()
        s.set_temperatures(np.zeros(s.n_bins))
        st = s.solve(p)
        x_aca = s.temperatures()
    n = T.shape[0]
    x64, _ = solver.newton(_effective_aca(Ut, V, C), b, np.zeros(n, bool), np.zeros(n),
                           np.full(n, 300.0))
    err = np.abs(x_aca - x64).max()
    print(f"jacobian={jac}: ACA solve vs FP64 on the same operator {err:.2e} K; ACA vs dense "
          f"max |dT| {np.abs(x_aca - x_dense).max():.3f} K, relative "
          f"{np.linalg.norm(x_aca - x_dense) / np.linalg.norm(x_dense):.2e}; {st}")
    assert st["converged"] and err < 1e-4


@pytest.mark.parametrize("jac", [0, 1])
def test_adjoint_solve_on_aca_operator(irtt, jac):
    """J^T lambda = rhs with J = (U V + C) D (exact) or the reference's
    D U V + C + I/dt (thermal_solver.cpp:44-75: C is not scaled by D)."""
    sd, s = _scene(irtt, "plates", mode=1, threads=256)
    rng = np.random.default_rng(3)
    with s:
        s.compress(rank=30)
        Ut, V, C, _, _ = s.aca_factors()
        n = s.n_bins
        x = 280.0 + 40.0 * rng.random(n)
        rhs = rng.normal(size=n)
        p = irtt.solver_params(jacobian=jac, lin_rtol=1e-13)
        lam, st = s.adjoint_solve(rhs, x=x, params=p)
    P = aca.lowrank(Ut.astype(np.float32).astype(np.float64),
                    V.astype(np.float32).astype(np.float64))
    D = np.diag(4.0 * x ** 3)
    if jac == 0:
        J = (P + np.diag(C)) @ D
    else:
        J = D @ P + np.diag(C) + np.eye(n) / 300.0
    lam64 = np.linalg.solve(J.T, rhs)
    rel = np.abs(lam - lam64).max() / np.abs(lam64).max()
    print(f"jacobian={jac}: lambda relative error {rel:.2e}; {st}")
    assert rel < 1e-7


def test_fp16_operator_refines_to_the_fp32_solution(irtt):
    """FP16 off-diagonal + per-row FP32 scale in the Krylov products, FP32
    operator in every Newton residual: the solve lands on the FP32 operator's
    solution; the adjoint solve is refined the same way."""
    sd, s = _scene(irtt, "city", mode=0, threads=128)
    Ta = 293.0
    rng = np.random.default_rng(5)
    with s:
        p = irtt.solver_params(ambient_kelvin=Ta, x_tol=1e-9, f_tol=1e-18, lin_rtol=1e-12)
        s.set_temperatures(np.zeros(s.n_bins))
        st32 = s.solve(p)
        x32 = s.temperatures()
        rhs = 1.0 + rng.random(s.n_bins)
        lam32, _ = s.adjoint_solve(rhs, params=p)
        s.select_operator(irtt.OP_DENSE_F16)
        s.set_temperatures(np.zeros(s.n_bins))
        st16 = s.solve(p)
        x16 = s.temperatures()
        lam16, ast = s.adjoint_solve(rhs, params=p)
        s.select_operator(irtt.OP_DENSE)
        s.set_temperatures(np.zeros(s.n_bins))
        s.solve(p)
        x32b = s.temperatures()
    err = np.abs(x16 - x32).max()
    lerr = np.abs(lam16 - lam32).max() / np.abs(lam32).max()
    print(f"FP16 form: max |dT| vs FP32 {err:.2e} K, adjoint rel {lerr:.2e}; FP32 {st32}; "
          f"FP16 {st16}; adjoint refinement {ast}")
    assert st16["converged"] and err < 1e-6 and lerr < 1e-8
    np.testing.assert_array_equal(x32, x32b)


@pytest.mark.parametrize("precision", [0, 1])
def test_aca_operator_products(irtt, precision):
    """Products with the stored factors (FP32, or FP16 with a row scale):
    within 2^-22 (FP32) / 2^-9 (FP16) of sum_l |U_il||V_lj||x_j| + |C_i x_i|
    from the FP64 factors, forward and transposed."""
    sd, s = _scene(irtt, "city")
    rng = np.random.default_rng(precision)
    with s:
        s.compress(rank=30, precision=precision)
        Ut, V, C, _, _ = s.aca_factors()
        n = s.n_bins
        X = rng.normal(size=(3, n))
        Y = s.apply_operator(X)
        Yt = s.apply_operator(X, transposed=True)
    P = aca.lowrank(Ut, V) + np.diag(C)
    A = aca.lowrank(np.abs(Ut), np.abs(V)) + np.diag(np.abs(C))
    tol = 2.0 ** -22 if precision == 0 else 2.0 ** -9
    for got, want, bound in ((Y, X @ P.T, np.abs(X) @ A.T), (Yt, X @ P, np.abs(X) @ A)):
        ratio = (np.abs(got - want) / bound).max()
        print(f"precision={precision}: max |err| / bound = {ratio:.2e}")
        assert ratio < tol


def test_aca_fp16_factors_solve(irtt):
    """FP16 factors change the temperatures of the ACA solve by far less than
    the 0.05 K bar."""
    sd, s = _scene(irtt, "plates", mode=1, threads=256)
    p = irtt.solver_params(ambient_kelvin=293.0, x_tol=1e-7, f_tol=1e-16)
    with s:
        s.compress(rank=40, precision=0)
        s.set_temperatures(np.zeros(s.n_bins))
        s.solve(p)
        x32 = s.temperatures()
        s.compress(rank=40, precision=1)
        s.set_temperatures(np.zeros(s.n_bins))
        st = s.solve(p)
        x16 = s.temperatures()
    err = np.abs(x16 - x32).max()
    print(f"ACA FP16 factors vs FP32 factors: max |dT| {err:.2e} K; {st}")
    assert st["converged"] and err < 0.05
