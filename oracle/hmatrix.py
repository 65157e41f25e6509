"""Hierarchical (H-matrix) block compression of the transport operator,
restated in numpy FP64 -- TEST INFRASTRUCTURE ONLY.

The reference compresses T with one global rank-k cross approximation and
no tolerance stop (``src/implementations/thermal_renderer/thermal_transport.cpp:
281-597``); its partial-pivot loop (``:388-461``) starts at row 0, takes
``findArgmaxNotInList`` pivots (``:463-474``: the largest magnitude among
the indices not yet used, ties to the smallest index) and leaves the
early-termination test commented out (``:439-451``).  The facade
configuration asks instead for hierarchical blocks compressed to a
relative tolerance with a rank cap.  This module states that algorithm, as
the device code (``csrc/hmatrix.hip``) runs it:

1. Bin order: 30-bit Morton codes of the bin positions (each axis scaled to
   its extent and quantised to 10 bits in FP64), ties by bin index.
2. Cluster tree: index ranges halved at the middle down to ``leaf_size``;
   each cluster's box is the FP64 min / max of its bins' positions.
3. Block tree from (root, root): a pair is far (admissible) when its boxes
   are apart and min(diam) <= eta * dist; else a pair with a leaf is near;
   else its four child pairs follow, in order.
4. Each far block with entries runs partial-pivot ACA in FP64: the
   reference's residual row / column (``:476-500``: sums over l = 0 .. k-1
   in order, one rounding per product and per sum) and pivot rule, a
   first row holding the block's largest entry, rows whose residual vanishes
   marked used without a term (the next row is then the next unused row
   with stored entries, in cyclic order), and the stop
   |u_k|^2 |v_k|^2 <= tol^2 |S_k|_F^2 with the Frobenius norm of the
   approximation updated per step, taken at a quarter of the tolerance
   (ACA's estimate follows the last term, and the exact error runs up to
   about three times that on smooth kernels).  The rank is capped at
   ``max_rank``.
5. A block that stopped is kept in low rank when the exact error
   |A - U V|_F (from the Gram matrices of U and V and the block's entries)
   is at most tol |A|_F and, when ``keep_smaller`` is set, its stored
   factors take fewer bytes than its entries (8 B each); every other
   entry stays in the sparse near field.
6. Stored factors: U (m x k) and V^T (n x k), row-major, each row scaled
   by max|row| * 2^-14 (FP32) and rounded to FP16 (through FP32), or FP32.

Every reduction a stopping decision reads (the squared norms and the dot
products of the norm update) is taken as the device takes it: 256 strided
partial sums, then a halving tree, so the steps, pivots and factors agree
bit for bit.

Pinned by (``tests/test_oracle_hmatrix.py``): the blocks partition the matrix
exactly once and every far block meets the admissibility condition; a rank-r
block is recovered exactly in r steps; on a smooth kernel (1 / (1 + r^2)
between two point clouds) every far block converges, its exact error is at
most tol, and its rank is at least the SVD rank for that tolerance and no
more than a few above it; the ACA factors cross-interpolate (the
approximation reproduces the block on every pivot row and column); the
exact error formula equals the error of the formed product; products with
the compressed operator equal the dense product of its reconstruction.
"""
import numpy as np

NT = 256  # threads of the device's block reductions
ACA_MARGIN = 0.25  # ACA stops at ACA_MARGIN * tol; blocks are accepted at tol


def tree_sum(x):
    """Sum as the device takes it: 256 strided sequential partial sums, then
    a halving tree (partial[t] += partial[t + off], off = 128 .. 1)."""
    x = np.asarray(x, np.float64)
    m = -(-len(x) // NT) * NT
    p = np.zeros(NT)
    if len(x):
        xp = np.zeros(m)
        xp[:len(x)] = x
        for row in xp.reshape(-1, NT):
            p = p + row
    off = NT // 2
    while off:
        p[:off] = p[:off] + p[off:2 * off]
        off //= 2
    return p[0]


def morton_order(pos):
    """Bin order by 30-bit Morton code of the positions (FP64), ties by index."""
    pos = np.asarray(pos, np.float64)
    lo = pos.min(axis=0)
    hi = pos.max(axis=0)
    ext = hi - lo
    q = np.zeros(pos.shape, np.uint64)
    for a in range(3):
        if ext[a] > 0:
            s = np.floor((pos[:, a] - lo[a]) / ext[a] * 1024.0)
            q[:, a] = np.clip(s, 0, 1023).astype(np.uint64)
    code = np.zeros(len(pos), np.uint64)
    for b in range(10):
        for a in range(3):
            code |= ((q[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + (2 - a))
    return np.lexsort((np.arange(len(pos)), code)).astype(np.int64)


class Cluster:
    __slots__ = ("b", "e", "lo", "hi", "kids")

    def __init__(self, b, e, lo, hi):
        self.b, self.e, self.lo, self.hi, self.kids = b, e, lo, hi, []


def cluster_tree(pos_perm, leaf_size):
    """Clusters of the permuted positions: ranges halved at the middle until
    at most leaf_size bins remain."""
    def make(b, e):
        c = Cluster(b, e, pos_perm[b:e].min(axis=0), pos_perm[b:e].max(axis=0))
        if e - b > leaf_size:
            mid = b + (e - b) // 2
            c.kids = [make(b, mid), make(mid, e)]
        return c
    return make(0, len(pos_perm))


def _diam(c):
    d = c.hi - c.lo
    return np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])


def _dist(t, s):
    g = [max(0.0, s.lo[a] - t.hi[a], t.lo[a] - s.hi[a]) for a in range(3)]
    return np.sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2])


def admissible(t, s, eta):
    d = _dist(t, s)
    return d > 0.0 and min(_diam(t), _diam(s)) <= eta * d


def block_tree(root, eta):
    """(far, near): lists of (r0, r1, c0, c1) in recursion order."""
    far, near = [], []

    def rec(t, s):
        if admissible(t, s, eta):
            far.append((t.b, t.e, s.b, s.e))
        elif not t.kids or not s.kids:
            near.append((t.b, t.e, s.b, s.e))
        else:
            for tc in t.kids:
                for sc in s.kids:
                    rec(tc, sc)
    rec(root, root)
    return far, near


def aca_block(A, tol, max_rank, has_entries=None):
    """Partial-pivot ACA of a dense FP64 block A (m x n) as stated in the
    module docstring.  Returns (U: k x m, V: k x n, pivot rows, pivot
    columns, converged)."""
    A = np.asarray(A, np.float64)
    m, n = A.shape
    if has_entries is None:
        has_entries = (A != 0).any(axis=1)
    U = np.zeros((max_rank, m))
    V = np.zeros((max_rank, n))
    used = np.zeros(m, bool)
    pcols, prows = [], []
    absA = np.abs(A)
    i = int(np.argmax(absA.max(axis=1))) if m and n else 0
    k = 0
    s2 = 0.0
    conv = False

    def next_row(i0):
        for d in range(1, m + 1):
            r = (i0 + d) % m
            if not used[r] and has_entries[r]:
                return r
        return -1

    while k < max_rank:
        r = A[i].copy()
        for l in range(k):
            r = r - U[l, i] * V[l]
        used[i] = True
        a = np.abs(r)
        if pcols:
            a[pcols] = 0.0
        a[~(a > 0)] = 0.0
        j = int(np.argmax(a)) if n else 0
        if not (n and a[j] > 0):
            i = next_row(i)
            if i < 0:
                conv = True
                break
            continue
        p = r[j]
        V[k] = r / p
        pcols.append(j)
        prows.append(i)
        c = A[:, j].copy()
        for l in range(k):
            c = c - V[l, j] * U[l]
        U[k] = c
        uu = tree_sum(U[k] * U[k])
        vv = tree_sum(V[k] * V[k])
        nu2 = uu * vv
        cross = 0.0
        for l in range(k):
            cross = cross + tree_sum(U[k] * U[l]) * tree_sum(V[k] * V[l])
        s2 = s2 + nu2 + 2.0 * cross
        k += 1
        if nu2 <= tol * tol * s2:
            conv = True
            break
        a = np.abs(U[k - 1])
        a[used] = 0.0
        a[~(a > 0)] = 0.0
        i2 = int(np.argmax(a)) if m else 0
        if m and a[i2] > 0:
            i = i2
        else:
            i = next_row(i)
            if i < 0:
                conv = True
                break
    return U[:k], V[:k], prows, pcols, conv


def exact_error2(A, U, V):
    """|A - U^T V|_F^2 from the Gram matrices and the block's entries."""
    A = np.asarray(A, np.float64)
    gu = U @ U.T
    gv = V @ V.T
    nz = np.nonzero(A)
    cross = np.sum(A[nz] * np.einsum("li,lj->ij", U, V)[nz]) if len(nz[0]) else 0.0
    return max(0.0, float(np.sum(gu * gv) - 2.0 * cross + np.sum(A[nz] ** 2)))


def quantise_rows(F, precision=1):
    """Rows of F (FP64) -> stored values as FP64: FP16 with per-row scale
    max|row| * 2^-14 (FP32), rounded through FP32; or FP32."""
    F = np.asarray(F, np.float64)
    if precision == 0:
        return F.astype(np.float32).astype(np.float64), None
    mx = np.abs(F).max(axis=1) if F.shape[1] else np.zeros(F.shape[0])
    s = (mx.astype(np.float32) * np.float32(2.0 ** -14)).astype(np.float32)
    q = np.zeros_like(F)
    ok = s > 0
    q[ok] = (F[ok] / s[ok, None].astype(np.float64)).astype(np.float32).astype(np.float16) \
        .astype(np.float64) * s[ok, None].astype(np.float64)
    return q, s


def compress(T, pos, tol=1e-4, max_rank=32, leaf_size=64, eta=1.0, keep_smaller=True,
             precision=1):
    """H-matrix of the dense FP64 matrix T (n x n, original bin order).
    Returns a dict: perm, far and near block lists, per-far-block status
    ('zero', 'noconv', 'reject', 'lowrank'), rank, exact relative error,
    the FP64 factors, and the stored (quantised) reconstruction."""
    T = np.asarray(T, np.float64)
    perm = morton_order(pos)
    Tp = T[np.ix_(perm, perm)]
    root = cluster_tree(np.asarray(pos, np.float64)[perm], leaf_size)
    far, near = block_tree(root, eta)
    out = {"perm": perm, "far": far, "near": near, "status": [], "rank": [], "err_rel": [],
           "U": [], "V": [], "prows": [], "pcols": []}
    H = np.zeros_like(Tp)
    covered = np.zeros(Tp.shape, bool)
    for (r0, r1, c0, c1) in far:
        A = Tp[r0:r1, c0:c1]
        nnz = int(np.count_nonzero(A))
        if nnz == 0:
            out["status"].append("zero")
            out["rank"].append(0)
            out["err_rel"].append(0.0)
            for key in ("U", "V", "prows", "pcols"):
                out[key].append(None)
            covered[r0:r1, c0:c1] = True
            continue
        U, V, pr, pc, conv = aca_block(A, ACA_MARGIN * tol, max_rank)
        k = U.shape[0]
        fro2 = float(np.sum(A[A != 0] ** 2))
        err = np.sqrt(exact_error2(A, U, V) / fro2)
        lr_bytes = (A.shape[0] + A.shape[1]) * k * (2 if precision else 4) + \
            ((A.shape[0] + A.shape[1]) * 4 if precision else 0)
        if not conv:
            st = "noconv"
        elif err > tol or (keep_smaller and lr_bytes >= 8 * nnz):
            st = "reject"
        else:
            st = "lowrank"
        out["status"].append(st)
        out["rank"].append(k)
        out["err_rel"].append(err)
        out["U"].append(U)
        out["V"].append(V)
        out["prows"].append(pr)
        out["pcols"].append(pc)
        if st == "lowrank":
            Uq, _ = quantise_rows(U.T, precision)
            Vq, _ = quantise_rows(V.T, precision)
            H[r0:r1, c0:c1] = Uq @ Vq.T
            covered[r0:r1, c0:c1] = True
    H[~covered] = Tp[~covered]
    Hd = np.zeros_like(T)
    Hd[np.ix_(perm, perm)] = H
    out["H"] = Hd
    return out
