"""Pins oracle/sparse_fp64.py (the FP64 reference for large operators)
against the dense-LU restatement oracle/solver.py on the same small
operators, and its record-built operator against the dense accumulators."""
import numpy as np
import pytest

from oracle import physics, solver, sparse_fp64

SIG = 5.670374419e-8


@pytest.fixture(scope="module")
def city(oracle):
    from irtt import scenes
    sd = scenes.city_block(seed=2, nx=2, ny=2, target_tris=1500)
    sc = oracle.OracleScene.from_irtt(sd)
    p = oracle.make_params(16, 4, batch_seed=3, bin_mode=0)
    acc, _, st = oracle.transport(sc, p)
    _, recs, _ = oracle.transport(sc, p, dense=False, records=st["photons"], with_tri=True)
    _, q, Q = oracle.photon_weights(sc, p)
    area, ea, mass, src = physics.bins(sd, 0)
    return sd, acc, recs, q, Q, ea, mass, src


def test_csr_from_records_is_the_stored_operator(city):
    sd, acc, (ra, rs, rt), q, Q, ea, _, _ = city
    rp, col, val, colsum = sparse_fp64.csr_from_records(ra, rs, rt, q, Q, ea, sd.nv)
    T = (np.diag(SIG * ea) - acc.astype(np.float64) * Q)
    mask = acc != 0
    np.fill_diagonal(mask, True)
    rows, cols = np.nonzero(mask)
    assert np.array_equal(col, cols.astype(np.uint32))
    assert np.array_equal(np.diff(rp.astype(np.int64)), np.bincount(rows, minlength=sd.nv))
    assert np.array_equal(val, T[rows, cols].astype(np.float32))
    assert np.array_equal(colsum, SIG * ea - acc.sum(0).astype(np.float64) * Q)


def test_steady_linear_matches_dense_lu(city):
    sd, acc, (ra, rs, rt), q, Q, ea, _, src = city
    rp, col, val, colsum = sparse_fp64.csr_from_records(ra, rs, rt, q, Q, ea, sd.nv)
    T = sparse_fp64.to_scipy(rp, col, val, sd.nv)
    b = sparse_fp64.rhs(src, colsum, 293.0)
    x, res = sparse_fp64.steady_linear(T, b)
    x_lu = solver.steady_linear(T.toarray(), b, np.zeros(sd.nv, bool), np.zeros(sd.nv))
    fixed = np.zeros(sd.nv, bool)
    fixed[:10] = True
    kfix = np.full(sd.nv, 300.0)
    xf, _ = sparse_fp64.steady_linear(T, b, fixed, kfix)
    xf_lu = solver.steady_linear(T.toarray(), b, fixed, kfix)
    print(f"residual {res:.1e}, max |dT| vs LU {np.abs(x - x_lu).max():.1e} K, "
          f"with fixed bins {np.abs(xf - xf_lu).max():.1e} K")
    assert res < 1e-13
    assert np.abs(x - x_lu).max() < 1e-8
    assert np.abs(xf - xf_lu).max() < 1e-8 and np.all(xf[:10] == 300.0)


def test_implicit_euler_matches_dense_newton(city):
    sd, acc, (ra, rs, rt), q, Q, ea, mass, src = city
    rp, col, val, colsum = sparse_fp64.csr_from_records(ra, rs, rt, q, Q, ea, sd.nv)
    T = sparse_fp64.to_scipy(rp, col, val, sd.nv)
    b = sparse_fp64.rhs(src, colsum, 293.0)
    x0 = np.full(sd.nv, 293.0)
    x1, it = sparse_fp64.implicit_euler_step(T, b, mass, x0, 300.0)
    want, _ = solver.newton(T.toarray(), b, np.zeros(sd.nv, bool), x0, x0, mass=mass, xprev=x0,
                            dt=300.0, steady=False, tol=1e-12)
    print(f"{it} Newton iterations, max |dT| vs dense Newton {np.abs(x1 - want).max():.1e} K")
    assert np.abs(x1 - want).max() < 1e-8
