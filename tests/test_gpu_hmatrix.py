"""Hierarchical block form of the sparse-layout operator (irtt_compress_hierarchical)
against the oracle restatement (oracle/hmatrix.py, pinned in
tests/test_oracle_hmatrix.py) on the same stored entries:

* bin order, far-block list and every far block's ACA (rank, pivot rows and
  columns, FP64 factors) equal the oracle's bit for bit;
* per-block exact errors equal the oracle's to 1e-9 relative plus the 5e-8
  resolution of the Gram-matrix form, and so do the keep / reject decisions
  away from the threshold;
* products with the stored form (FP32 and FP16 factors) equal the dense
  product of the oracle's reconstruction to 1e-12 of |H||x|;
* steady and adjoint solves with the H-matrix Krylov products and the stored
  entries' residuals agree with FP64 solves of the stored operator.
"""
import numpy as np
import pytest

from oracle import hmatrix as ohm
from oracle import physics, solver

pytestmark = pytest.mark.gpu

TA = 293.0


def _city(irtt, seed=6, target=2500, threads=128):
    from irtt import scenes
    sd = scenes.city_block(seed=seed, nx=2, ny=2, target_tris=target)
    s = irtt.Scene.from_data(sd, irtt.BIN_VERTEX)
    s.build_transport(threads_per_tri=threads, rays_per_thread=4, batch_seed=2,
                      layout=irtt.LAYOUT_SPARSE)
    return sd, s


def _dense(s):
    rp, col, val = s.transport_sparse()
    n = rp.size - 1
    T = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
    T[rows, col] = val.astype(np.float64)
    return T


# loose tolerance and low rank cap so that many far blocks of a Monte-Carlo
# operator stop in low rank (at 1e-4 nearly none does), exercising every path
@pytest.mark.parametrize("tol,max_rank,precision,allow_larger", [
    (0.5, 8, 0, True),
    (0.3, 16, 1, True),
    (1e-4, 32, 1, False),
])
def test_hmatrix_matches_oracle(irtt, tol, max_rank, precision, allow_larger):
    sd, s = _city(irtt)
    print(f"[hmatrix] city 2x2 n_bins={s.n_bins} tol={tol} rank<={max_rank} "
          f"precision={precision}", flush=True)
    with s:
        T = _dense(s)
        st = s.compress_hierarchical(tol=tol, max_rank=max_rank, leaf_size=32,
                                     precision=precision, allow_larger=allow_larger,
                                     keep_factors=True)
        perm, blocks, err, piv = s.hmatrix_blocks()
        facs = {b: s.hmatrix_factors(b, blocks) for b in range(len(blocks))
                if blocks[b, 5] != irtt.HBLOCK_ZERO}
        rng = np.random.default_rng(1)
        X = rng.standard_normal((3, s.n_bins))
        Y = s.apply_operator(X)
        Yt = s.apply_operator(X, transposed=True)
    print(f"[hmatrix] {st}", flush=True)
    pos = sd.xyz.astype(np.float64).reshape(-1, 3)
    assert np.array_equal(perm, ohm.morton_order(pos))
    root = ohm.cluster_tree(pos[perm], 32)
    far, near = ohm.block_tree(root, 1.0)
    assert [tuple(b) for b in blocks[:, :4].tolist()] == [tuple(f) for f in far]
    assert st["n_far"] == len(far) and st["n_near"] == len(near)
    Tp = T[np.ix_(perm, perm)]
    n_lr = 0
    for b, (r0, r1, c0, c1) in enumerate(far):
        A = Tp[r0:r1, c0:c1]
        if not A.any():
            assert blocks[b, 5] == irtt.HBLOCK_ZERO and blocks[b, 4] == 0
            continue
        U, V, pr, pc, conv = ohm.aca_block(A, ohm.ACA_MARGIN * tol, max_rank)
        k = U.shape[0]
        assert blocks[b, 4] == k, (b, blocks[b], k)
        assert list(piv[b, :k]) == pr and list(piv[b, 32:32 + k]) == pc
        Ug, Vg = facs[b]
        assert np.array_equal(Ug, U) and np.array_equal(Vg, V), b
        fro2 = np.sum(A[A != 0] ** 2)
        e = np.sqrt(ohm.exact_error2(A, U, V) / fro2)
        # the Gram form of the error cancels to ~1e-16 |A|_F^2: about 1e-8 relative
        assert abs(err[b] - e) <= 1e-9 * e + 5e-8, (b, err[b], e)
        if not conv:
            assert blocks[b, 5] == irtt.HBLOCK_NOCONV
        elif abs(e - tol) > 1e-6 * tol:
            lrb = (A.shape[0] + A.shape[1]) * k * (2 if precision else 4) + \
                ((A.shape[0] + A.shape[1]) * 4 if precision else 0)
            small = lrb < 8 * np.count_nonzero(A)
            want = irtt.HBLOCK_LOWRANK if e <= tol and (allow_larger or small) \
                else irtt.HBLOCK_REJECTED
            assert blocks[b, 5] == want, (b, blocks[b], e)
        n_lr += blocks[b, 5] == irtt.HBLOCK_LOWRANK
    assert n_lr == st["n_lowrank"]
    if tol > 0.1:
        assert n_lr > 10
    # products with the stored form vs the oracle's reconstruction of it
    # (the oracle's keep decisions are the library's, checked above)
    H = np.zeros_like(Tp)
    covered = np.zeros(Tp.shape, bool)
    for b, (r0, r1, c0, c1) in enumerate(far):
        if blocks[b, 5] == irtt.HBLOCK_LOWRANK:
            U, V = facs[b]
            Uq, _ = ohm.quantise_rows(U.T, precision)
            Vq, _ = ohm.quantise_rows(V.T, precision)
            H[r0:r1, c0:c1] = Uq @ Vq.T
            covered[r0:r1, c0:c1] = True
    H[~covered] = Tp[~covered]
    Hd = np.zeros_like(T)
    Hd[np.ix_(perm, perm)] = H
    for Yg, M in ((Y, Hd), (Yt, Hd.T)):
        ref = X @ M.T
        scale = np.abs(X) @ np.abs(M).T
        assert np.all(np.abs(Yg - ref) <= 1e-12 * scale + 1e-300)
    # stated errors: |T - H|_F / |T|_F of the stored form
    e_st = np.linalg.norm(Hd - T) / np.linalg.norm(T)
    assert abs(st["err_frob_rel_stored"] - e_st) <= 1e-6 * max(e_st, 1e-12) + 1e-13


def test_hmatrix_solves_vs_fp64(irtt):
    """H-matrix Krylov products, stored-entry residuals and refinement: the
    steady state and the adjoint state of the stored operator (FP64 dense
    solves), to 1e-7 K and 1e-8 relative, with low-rank blocks in play."""
    sd, s = _city(irtt)
    src = physics.bins(sd, 0)[3] * 1.3
    with s:
        T = _dense(s)
        cs = s.column_sums()
        st = s.compress_hierarchical(tol=0.3, max_rank=16, leaf_size=32, allow_larger=True)
        assert st["n_lowrank"] > 10
        s.set_source(src)
        p = irtt.solver_params(ambient_kelvin=TA, x_tol=1e-9, f_tol=1e-18, lin_rtol=1e-12)
        s.set_temperatures(np.zeros(T.shape[0]))
        sol = s.solve(params=p)
        x = s.temperatures()
        rhs = np.random.default_rng(3).normal(size=T.shape[0])
        lam = s.adjoint_solve(rhs, x=x, params=p)[0]
    n = T.shape[0]
    x64 = solver.steady_linear(T, solver.rhs(src, cs, TA), np.zeros(n, bool), np.zeros(n))
    J = T * (4.0 * x ** 3)[None, :]
    l64 = np.linalg.solve(J.T, rhs)
    dT = np.abs(x - x64).max()
    rel = np.abs(lam - l64).max() / np.abs(l64).max()
    print(f"[hmatrix] steady |dT| vs FP64 {dT:.2e} K, adjoint relative {rel:.2e}; {sol}")
    assert sol["converged"]
    assert dT < 1e-7 and rel < 1e-8


def test_hmatrix_batched_work_space_is_the_same(irtt):
    """The cross approximations taken in batches of blocks (a 1 MiB work
    space: many batches, kept blocks rebuilt for packing) give the same block
    outcomes and bit-identical products as all blocks at once."""
    sd, s = _city(irtt)
    rng = np.random.default_rng(7)
    X = rng.standard_normal((2, s.n_bins))
    out = []
    with s:
        for kw in ({"keep_factors": True}, {"workspace_mib": 1}):
            st = s.compress_hierarchical(tol=0.3, max_rank=16, leaf_size=32, allow_larger=True,
                                         **kw)
            _, blocks, err, _ = s.hmatrix_blocks()
            out.append((st, blocks, err, s.apply_operator(X), s.apply_operator(X, True)))
    (s1, b1, e1, y1, t1), (s2, b2, e2, y2, t2) = out
    print(f"[hmatrix] one batch {s1['build_bytes']} B, batched {s2['build_bytes']} B", flush=True)
    assert s2["build_bytes"] < s1["build_bytes"] and s2["build_bytes"] <= (1 << 20) + (64 << 10)
    assert np.array_equal(b1, b2) and np.array_equal(e1, e2)
    assert np.array_equal(y1, y2) and np.array_equal(t1, t2)
    assert s1["n_lowrank"] > 10 and s1["err_frob_rel_stored"] == s2["err_frob_rel_stored"]


def test_hmatrix_needs_sparse_layout(irtt):
    from irtt import scenes
    sd = scenes.city_block(seed=6, nx=2, ny=2, target_tris=1500)
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=16, rays_per_thread=4, batch_seed=2)
        with pytest.raises(irtt.IrttError):
            s.compress_hierarchical()
        with pytest.raises(irtt.IrttError):
            s.select_operator(irtt.OP_HMATRIX)
