"""Edge cases of the hot path against the oracle: a tree of one leaf, zero-area
emitters next to non-traceable triangles, rays whose direction has exact zero
components (the slab test's 1/0 -> 1e30 path), and an empty scene."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STAT_KEYS = ("photons", "rays", "absorbed", "miss", "too_close", "backface", "degenerate",
             "forced_absorb", "node_visits", "tri_tests")


def _transport_vs_oracle(irtt, oracle, sd, bin_mode, threads, rays, seed):
    with irtt.Scene.from_data(sd, bin_mode) as s:
        st = s.build_transport(threads_per_tri=threads, rays_per_thread=rays, batch_seed=seed)
        acc, Q = s.transport_fixed()
    sc = oracle.OracleScene.from_irtt(sd)
    p = oracle.make_params(threads, rays, batch_seed=seed, bin_mode=bin_mode)
    oacc, _, ost = oracle.transport(sc, p)
    _, _, oQ = oracle.photon_weights(sc, p)
    assert Q == oQ
    for k in STAT_KEYS:
        assert st[k] == ost[k], (k, st[k], ost[k])
    assert np.array_equal(acc, oacc)
    return st, acc


def test_one_leaf_tree(irtt, oracle):
    """Only the ground is traceable (a tree of one leaf: queries are answered
    without traversal); a small non-traceable triangle above it faces down and
    emits onto it, the ground's photons escape."""
    from irtt import scenes
    xyz = np.array([[-5, -5, 0], [5, -5, 0], [0, 5, 0],          # ground, normal +z
                    [0, 0, 1], [0.2, 0.3, 1], [0.3, 0, 1]], np.float32)  # normal -z
    idx = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    sd = scenes._uniform_scene("one_leaf", xyz, idx, rho_d=0.2)
    sd.flags[1] = scenes.MAT_DIFFUSE_EMISSION  # not traceable
    print("[edge] one-leaf tree", flush=True)
    for bin_mode in (0, 1):
        st, acc = _transport_vs_oracle(irtt, oracle, sd, bin_mode, 256, 8, 5)
        print(bin_mode, st)
        assert st["node_visits"] == 0 and st["tri_tests"] == st["rays"]
        assert st["absorbed"] > 0 and st["miss"] > 0


def test_degenerate_and_non_traceable_triangles(irtt, oracle):
    """City block with zero-area triangles (a repeated vertex) among the
    emitters and some walls not traceable: accumulators and every counter
    bit-exact, degenerate emitters counted and dropped."""
    from irtt import scenes
    sd = scenes.city_block(seed=4, nx=2, ny=2, target_tris=2000)
    rng = np.random.default_rng(7)
    deg = rng.choice(sd.nt, 40, replace=False)
    sd.idx[deg, 2] = sd.idx[deg, 0]
    rest = np.setdiff1d(np.arange(sd.nt), deg)
    nontr = rng.choice(rest, 60, replace=False)
    sd.flags[nontr] &= ~np.uint8(scenes.MAT_TRACEABLE)
    print("[edge] degenerate + non-traceable", flush=True)
    st, _ = _transport_vs_oracle(irtt, oracle, sd, 0, 16, 4, 11)
    print(st)
    assert st["degenerate"] == 40 * 16 * 4


def test_axis_parallel_rays(irtt, oracle):
    """Rays along +-x, +-y, +-z (two direction components exactly zero), from
    grid vertices, edge midpoints and random points of a closed box and of two
    facing plates: hits bit-exact vs the oracle's tree walk and equal to brute
    force (exact vertex / edge hits included)."""
    from irtt import scenes
    rng = np.random.default_rng(2)
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]],
                    np.float32)
    for sd in (scenes.closed_box(4), scenes.parallel_plates(n=8)):
        sc = oracle.OracleScene.from_irtt(sd)
        grid = np.round(rng.uniform(0, 1, size=(600, 3)) * 8) / 8   # on the 1/8 grid
        mid = 0.5 * (sd.xyz[sd.idx[:, 0]] + sd.xyz[sd.idx[:, 1]])
        pts = np.concatenate([sd.xyz, mid, grid, rng.uniform(-0.5, 1.5, size=(600, 3))])
        o = np.repeat(pts, len(axes), axis=0)
        d = np.tile(axes, (len(pts), 1))
        rays = np.concatenate([o, d], 1).astype(np.float32)
        print(f"[edge] axis-parallel rays {sd.name}: {len(rays)}", flush=True)
        with irtt.Scene(sd.xyz, sd.idx, sd.materials()) as s:
            gp, gt, g1, g2 = s.trace_rays(rays)
        ob = oracle.OracleBVH(sc)
        op, ot, o1, o2, _ = ob.trace(rays)
        bp, bt, _, _, ties = oracle.trace_brute(sc, rays)
        assert np.array_equal(gp, op) and np.array_equal(gt, ot)
        assert np.array_equal(g1, o1) and np.array_equal(g2, o2)
        hit = gp >= 0
        print(f"hits {hit.sum()} of {len(rays)}, brute-force ties {ties}")
        assert np.array_equal(gp, bp)
        assert np.array_equal(gt[hit], bt[hit])


def test_empty_scene_is_rejected(irtt):
    """No triangles: scene creation fails with an error, nothing is traced."""
    with pytest.raises(Exception):
        irtt.Scene(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32),
                   np.zeros(0, irtt.MATERIAL_DTYPE))
