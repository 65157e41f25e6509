"""Adaptive cross approximation restated in numpy FP64 (test infrastructure only).

Follows ``src/implementations/thermal_renderer/thermal_transport.cpp``:
``simpleAdaptiveCrossApproximation`` (full pivoting, :360-386),
``completeAdaptiveCrossApproximation`` (partial pivoting from row 0, the
reference's default, :388-461), ``adaptiveCrossApproximationPlus`` (:522-597),
``findArgmaxNotInList`` (:463-474), the residual row/column (:476-500),
``resetReferenceRow/Col`` (:502-520) and ``collumnSumCorrection`` (:309-358).

Every residual sum runs over l = 0 .. k-1 in order with one rounding per
product and per addition, the order the reference writes, so a device
implementation that keeps that order agrees bit for bit.

Pinned by (tests/test_oracle_aca.py): exact recovery of rank-r matrices in r
steps by all three variants, the cross-interpolation property (the
approximation reproduces T on every pivot row and column), the error never
below the SVD truncation bound, the argmax tie and exclusion rule, and the
column-sum correction against its definition on a hand-sized case.
"""
import numpy as np


def argmax_not_in(v, excluded):
    """Largest |v_i| with i not excluded; ties to the smallest i; only strictly
    positive magnitudes qualify and the default is 0
    (thermal_transport.cpp:463-474)."""
    a = np.abs(np.asarray(v, np.float64)).copy()
    if excluded:
        a[np.fromiter(excluded, dtype=np.int64)] = 0.0
    a[~(a > 0)] = 0.0  # NaN never wins
    i = int(np.argmax(a))
    return i if a[i] > 0 else 0


def _res_row(T, Ut, V, i, k):
    acc = np.zeros(T.shape[1])
    for l in range(k):
        acc = acc + Ut[l, i] * V[l]
    return T[i] - acc


def _res_col(T, Ut, V, j, k):
    acc = np.zeros(T.shape[0])
    for l in range(k):
        acc = acc + V[l, j] * Ut[l]
    return T[:, j] - acc


def _div(x, p):
    return x / p if p != 0 else x.copy()


def aca_partial(T, k):
    """Partial-pivot ACA (thermal_transport.cpp:388-461).  Returns Ut (k x n,
    the columns of U), V (k x n), the pivot rows (k + 1: the last is the row
    the next step would take) and the pivot columns (k)."""
    T = np.asarray(T, np.float64)
    n = T.shape[0]
    Ut = np.zeros((k, n))
    V = np.zeros((k, n))
    rows, cols = {0}, set()
    prow, pcol = [0], []
    i = 0
    for kk in range(k):
        v = _res_row(T, Ut, V, i, kk)
        j = argmax_not_in(v, cols)
        cols.add(j)
        pcol.append(j)
        u = _res_col(T, Ut, V, j, kk)
        V[kk] = _div(v, v[j])
        Ut[kk] = u
        i = argmax_not_in(u, rows)
        if kk > 0:  # the row chosen after step 0 is not entered (:430-431)
            rows.add(i)
        prow.append(i)
    return Ut, V, prow, pcol


def aca_full(T, k):
    """Full-pivot ACA (thermal_transport.cpp:360-386): R = T - sum u v^T, pivot
    at the first maximum of |R| in row-major order."""
    T = np.asarray(T, np.float64)
    n = T.shape[0]
    Ut = np.zeros((k, n))
    V = np.zeros((k, n))
    A = np.zeros_like(T)
    prow, pcol = [], []
    for kk in range(k):
        if kk:
            A = A + np.outer(Ut[kk - 1], V[kk - 1])
        R = T - A
        flat = int(np.argmax(np.abs(R)))
        pr, pc = divmod(flat, n)
        Ut[kk] = _div(R[:, pc], R[pr, pc])
        V[kk] = R[pr]
        prow.append(pr)
        pcol.append(pc)
    return Ut, V, prow, pcol


class CrtRand:
    """The MSVC C runtime's rand(): x = 214013 x + 2531011, output bits 16..30
    (RAND_MAX 32767); the reference's ACA+ draws its reference rows and
    columns with it (thermal_transport.cpp:502-520)."""

    def __init__(self, seed=1):
        self.x = seed & 0xFFFFFFFF

    def next(self):
        self.x = (self.x * 214013 + 2531011) & 0xFFFFFFFF
        return (self.x >> 16) & 0x7FFF

    def index(self, n):
        return int(self.next() / 32767.0 * (n - 1))


def _reset_reference(rng, n, used):
    r = 0
    for _ in range(1, n):
        r = rng.index(n)
        if r not in used:
            break
    return r


def aca_plus(T, k, seed=1):
    """ACA+ (thermal_transport.cpp:522-597).  U columns are normalised by the
    pivot, V rows are not."""
    T = np.asarray(T, np.float64)
    n = T.shape[0]
    Ut = np.zeros((k, n))
    V = np.zeros((k, n))
    rng = CrtRand(seed)
    rows, cols = set(), set()
    r = _reset_reference(rng, n, rows)
    c = _reset_reference(rng, n, rows)  # the reference passes the row set here
    prow, pcol = [], []
    for kk in range(k):
        v = _res_row(T, Ut, V, r, kk)
        u = _res_col(T, Ut, V, c, kk)
        i = argmax_not_in(u, rows)
        j = argmax_not_in(v, cols)
        if abs(u[i]) > abs(v[j]):
            v = _res_row(T, Ut, V, i, kk)
            j = argmax_not_in(v, cols)
            u = _res_col(T, Ut, V, j, kk)
        else:
            u = _res_col(T, Ut, V, j, kk)
            i = argmax_not_in(u, rows)
            v = _res_row(T, Ut, V, i, kk)
        rows.add(i)
        cols.add(j)
        if i == r:
            r = _reset_reference(rng, n, rows)
        elif j == c:
            c = _reset_reference(rng, n, cols)
        Ut[kk] = _div(u, u[i])
        V[kk] = v
        prow.append(i)
        pcol.append(j)
    return Ut, V, prow, pcol


def lowrank(Ut, V):
    """U V accumulated over l in order."""
    A = np.zeros((Ut.shape[1], V.shape[1]))
    for l in range(Ut.shape[0]):
        A = A + np.outer(Ut[l], V[l])
    return A


def column_sum_correction(Ut, V, C, max_uv=None):
    """thermal_transport.cpp:309-358: where a column of U V + C sums below half
    the largest |U V| entry, raise its diagonal entry to that value.  Column
    sums of U run over the rows in order.  Returns (new C, corrected mask)."""
    m = np.abs(lowrank(Ut, V)).max() if max_uv is None else max_uv
    us = np.cumsum(Ut, axis=1)[:, -1]
    cs = np.zeros(V.shape[1])
    for l in range(Ut.shape[0]):
        cs = cs + us[l] * V[l]
    cs = cs + C
    mask = cs < 0.5 * m
    Cn = C.copy()
    Cn[mask] = C[mask] + (0.5 * m - cs[mask])
    return Cn, mask


def offdiag_operator(acc, Q):
    """The matrix the reference compresses: the photon part of T without the
    emission diagonal, FP32(0 - acc Q) (thermal_transport.cpp:634-660)."""
    return (0.0 - acc.astype(np.float64) * Q).astype(np.float32).astype(np.float64)
