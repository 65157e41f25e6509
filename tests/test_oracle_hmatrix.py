"""Pins the H-matrix restatement (oracle/hmatrix.py) by properties that do not
repeat its formulas: the block partition, admissibility, exact recovery of
low-rank blocks, convergence on a smooth kernel against the SVD rank, cross
interpolation, the exact-error identity and products of the stored form."""
import numpy as np
import pytest

from oracle import hmatrix as hm


def _clouds(n, seed=0):
    rng = np.random.default_rng(seed)
    # two surfaces 4 m apart, like facing facades
    a = np.c_[rng.uniform(0, 4, n // 2), rng.uniform(0, 4, n // 2), np.zeros(n // 2)]
    b = np.c_[rng.uniform(0, 4, n - n // 2), rng.uniform(0, 4, n - n // 2),
              np.full(n - n // 2, 4.0)]
    return np.vstack([a, b])


def _kernel(pos):
    d2 = ((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1)
    return 1.0 / (1.0 + d2)


def test_tree_sum_is_a_sum():
    rng = np.random.default_rng(1)
    for n in (0, 1, 255, 256, 257, 5000):
        x = rng.standard_normal(n)
        assert abs(hm.tree_sum(x) - x.sum()) <= 1e-12 * max(1.0, np.abs(x).sum())


def test_morton_order_is_a_permutation_and_groups_space():
    pos = _clouds(3000)
    p = hm.morton_order(pos)
    assert sorted(p.tolist()) == list(range(3000))
    # consecutive bins in the order are close: mean step far below a random order's
    step = np.linalg.norm(np.diff(pos[p], axis=0), axis=1).mean()
    rnd = np.linalg.norm(np.diff(pos, axis=0), axis=1).mean()
    assert step < 0.25 * rnd


@pytest.mark.parametrize("eta", [0.5, 1.0, 2.0])
def test_blocks_partition_and_admissibility(eta):
    pos = _clouds(1500, seed=2)
    p = hm.morton_order(pos)
    root = hm.cluster_tree(pos[p], 48)
    far, near = hm.block_tree(root, eta)
    cover = np.zeros((1500, 1500), np.int32)
    for r0, r1, c0, c1 in far + near:
        cover[r0:r1, c0:c1] += 1
    assert (cover == 1).all()
    pp = pos[p]
    for r0, r1, c0, c1 in far:
        t, s = pp[r0:r1], pp[c0:c1]
        # diameters and distance from the points themselves (boxes bound them)
        dist = np.sqrt(((t[:, None] - s[None]) ** 2).sum(-1)).min()
        dt = np.linalg.norm(t.max(0) - t.min(0))
        ds = np.linalg.norm(s.max(0) - s.min(0))
        assert min(dt, ds) <= eta * dist + 1e-12
    for r0, r1, c0, c1 in near:
        assert r1 - r0 <= 48 or c1 - c0 <= 48
    assert len(far) > 0


@pytest.mark.parametrize("r", [1, 3, 7])
def test_rank_r_block_is_recovered_in_r_steps(r):
    rng = np.random.default_rng(r)
    A = rng.standard_normal((90, r)) @ rng.standard_normal((r, 70))
    U, V, pr, pc, conv = hm.aca_block(A, 1e-10, 32)
    assert conv
    assert U.shape[0] <= r + 1
    assert np.abs(U.T @ V - A).max() <= 1e-10 * np.abs(A).max()


def test_cross_interpolation_on_pivots():
    rng = np.random.default_rng(5)
    A = _kernel(_clouds(200, seed=5))[:100, 100:]
    U, V, pr, pc, conv = hm.aca_block(A, 1e-14, 12)
    R = U.T @ V
    for i in pr:
        assert np.allclose(R[i], A[i], rtol=0, atol=1e-12)
    for j in pc:
        assert np.allclose(R[:, j], A[:, j], rtol=0, atol=1e-12)


def test_zero_rows_are_skipped_and_sparse_blocks_are_exact():
    A = np.zeros((40, 30))
    A[3, 7] = 2.0
    A[17, 2] = -1.0
    A[17, 9] = 0.5
    A[30, 29] = 4.0
    U, V, pr, pc, conv = hm.aca_block(A, 1e-4, 32)
    assert conv
    assert np.abs(U.T @ V - A).max() <= 1e-15
    assert U.shape[0] == 3  # rank of the pattern above (rows 3, 17, 30)


def test_smooth_kernel_converges_within_tolerance_and_near_svd_rank():
    pos = _clouds(1200, seed=3)
    T = _kernel(pos)
    tol = 1e-4
    out = hm.compress(T, pos, tol=tol, max_rank=32, leaf_size=64, eta=1.0, keep_smaller=False,
                      precision=0)
    p = out["perm"]
    Tp = T[np.ix_(p, p)]
    n_lr = 0
    for (r0, r1, c0, c1), st, k, err in zip(out["far"], out["status"], out["rank"],
                                            out["err_rel"]):
        A = Tp[r0:r1, c0:c1]
        assert st == "lowrank", (r0, r1, c0, c1, st, k, err)
        n_lr += 1
        assert err <= tol
        s = np.linalg.svd(A, compute_uv=False)
        # smallest rank whose truncation error meets the tolerance
        tail = np.sqrt(np.cumsum((s ** 2)[::-1])[::-1])
        svd_rank = int(np.argmax(np.r_[tail, 0.0] <= tol * np.linalg.norm(s)))
        assert svd_rank <= k <= svd_rank + 6
    assert n_lr > 10


def test_exact_error_identity():
    rng = np.random.default_rng(9)
    A = rng.standard_normal((50, 40)) * (rng.random((50, 40)) < 0.3)
    U, V, *_ = hm.aca_block(A, 1e-3, 6)
    e2 = hm.exact_error2(A, U, V)
    assert abs(e2 - np.sum((A - U.T @ V) ** 2)) <= 1e-10 * np.sum(A ** 2)


def test_stored_products_and_fp16_rounding():
    pos = _clouds(800, seed=4)
    T = _kernel(pos)
    o32 = hm.compress(T, pos, tol=1e-4, precision=0, keep_smaller=False)
    o16 = hm.compress(T, pos, tol=1e-4, precision=1, keep_smaller=False)
    x = np.random.default_rng(4).standard_normal(800)
    # FP32 factors: the tolerance bounds the product's error
    e32 = np.linalg.norm(o32["H"] @ x - T @ x) / np.linalg.norm(T @ x)
    assert e32 < 1e-4
    # FP16 factors: rows rounded to 2^-12 relative of their largest entry
    q, s = hm.quantise_rows(np.array([[1.0, 1e-3, -0.3337]]))
    assert s[0] == np.float32(2.0 ** -14)
    assert abs(q[0, 2] + 0.3337) <= 0.3337 * 2.0 ** -11
    e16 = np.linalg.norm(o16["H"] - T) / np.linalg.norm(T)
    assert 1e-6 < e16 < 2e-3
