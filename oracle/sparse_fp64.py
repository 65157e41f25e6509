"""FP64 solves on a stored sparse operator (test infrastructure only).

The same discrete system as ``oracle/solver.py`` (steady and implicit-Euler
forms of ``src/implementations/thermal_renderer/thermal_solver.cpp:13-229``
with the library's source and ambient terms), for operators too large for
dense LU: the operator is taken as stored (compressed rows, FP32 entries
widened exactly), every product and sum is FP64 (scipy.sparse), and the
linear systems are solved by Jacobi-preconditioned GMRES to a relative
residual of 1e-14 -- an independent FP64 reference for the GPU solves of the
same operator.  ``csr_from_records`` builds that operator from the oracle
tracer's per-photon records exactly as the library stores it.

Pinned by: ``oracle/solver.py``'s dense LU solves on the same small operators
(tests/test_oracle_sparse_fp64.py), the residual each solve returns, and the
exact linear form of the steady state without conduction (y = x^4).
"""
import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spla

SIGMA = 5.670374419e-8


def csr_from_records(ra, rs, rt, q, Q, eps_area, nb):
    """Compressed rows (rowptr, col, FP32 value) of T = diag(sigma eps A) - acc Q
    with acc[a, s] = sum of q[rt] over the records (ra, rs): the entries some
    photon connected plus every diagonal, columns ascending within a row --
    the library's stored operator, bit for bit.  Also returns the exact
    column sums sigma eps A - Q sum_a acc[a, s] (FP64)."""
    nb = int(nb)
    key = ra.astype(np.uint64) * np.uint64(nb) + rs.astype(np.uint64)
    key = np.concatenate([key, np.arange(nb, dtype=np.uint64) * np.uint64(nb + 1)])
    qv = np.concatenate([q[rt].astype(np.uint64), np.zeros(nb, np.uint64)])
    order = np.argsort(key, kind="stable")
    ks = key[order]
    start = np.flatnonzero(np.concatenate([[True], ks[1:] != ks[:-1]]))
    acc = np.add.reduceat(qv[order], start)
    uk = ks[start]
    rows = (uk // np.uint64(nb)).astype(np.int64)
    cols = (uk % np.uint64(nb)).astype(np.int64)
    diag = SIGMA * np.asarray(eps_area, np.float64)
    vals = np.where(rows == cols, diag[rows], 0.0) - acc.astype(np.float64) * Q
    rp = np.zeros(nb + 1, np.uint32)
    np.cumsum(np.bincount(rows, minlength=nb), out=rp[1:])
    col_acc = np.zeros(nb, np.uint64)
    np.add.at(col_acc, cols, acc)
    colsum = diag - col_acc.astype(np.float64) * Q
    return rp, cols.astype(np.uint32), vals.astype(np.float32), colsum


def to_scipy(rowptr, col, val, n):
    return sps.csr_matrix((np.asarray(val, np.float32).astype(np.float64),
                           np.asarray(col, np.int64), np.asarray(rowptr, np.int64)),
                          shape=(n, n))


def _gmres(A, b, rtol=1e-14, restart=60, maxiter=400):
    d = A.diagonal()
    dinv = np.where(d != 0.0, 1.0 / d, 1.0)
    M = spla.LinearOperator(A.shape, matvec=lambda v: dinv * v, dtype=np.float64)
    x, info = spla.gmres(A, b, rtol=rtol, atol=0.0, restart=restart, maxiter=maxiter, M=M)
    res = np.linalg.norm(A @ x - b) / max(np.linalg.norm(b), 1e-300)
    return x, res


def rhs(source, colsum, ambient_kelvin):
    b = np.asarray(source, np.float64).copy()
    if ambient_kelvin is not None and ambient_kelvin >= 0:
        b += ambient_kelvin ** 4 * np.asarray(colsum, np.float64)
    return b


def steady_linear(T, b, fixed=None, kfix=None):
    """Steady state without conduction, T y = b linear in y = x^4 (free rows:
    T_ff y_f = b_f - T_fx y_x).  Returns (x, relative residual of the y solve)."""
    n = T.shape[0]
    fixed = np.zeros(n, bool) if fixed is None else np.asarray(fixed, bool)
    kfix = np.zeros(n) if kfix is None else np.asarray(kfix, np.float64)
    f = ~fixed
    y = np.where(fixed, kfix ** 4, 0.0)
    Tc = T.tocsc()
    Tff = Tc[f][:, f] if fixed.any() else T
    r = np.asarray(b, np.float64)[f] - (Tc[f][:, fixed] @ y[fixed] if fixed.any() else 0.0)
    yf, res = _gmres(sps.csr_matrix(Tff), r)
    y[f] = yf
    x = np.where(fixed, kfix, np.sign(y) * np.abs(y) ** 0.25)
    return x, res


def implicit_euler_step(T, b, mass, x_prev, dt, fixed=None, tol=1e-11, max_iter=30):
    """One implicit-Euler step: M (x - x_prev) + dt (T x^4 - b) = 0 on free
    bins (fixed bins keep x_prev), Newton with exact Jacobian
    M + dt T diag(4 x^3) and GMRES steps.  Returns (x, Newton iterations)."""
    n = T.shape[0]
    fixed = np.zeros(n, bool) if fixed is None else np.asarray(fixed, bool)
    f = ~fixed
    m = np.asarray(mass, np.float64)
    x = np.asarray(x_prev, np.float64).copy()
    for it in range(max_iter):
        r = m * (x - x_prev) + dt * (T @ (x ** 4) - b)
        r[fixed] = 0.0
        J = sps.diags(m) + dt * (T @ sps.diags(4.0 * x ** 3))
        J = sps.csr_matrix(J)[f][:, f]
        dx, _ = _gmres(J, -r[f])
        x[f] += dx
        if np.max(np.abs(dx)) < tol:
            return x, it + 1
    return x, max_iter
