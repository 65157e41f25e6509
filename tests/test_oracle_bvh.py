"""Pins the oracle LBVH and traversal: structural invariants of the Karras
hierarchy and closest-hit equality with brute force (including ties)."""
import numpy as np
import pytest

from conftest import random_rays


def _scene(oracle, sd):
    return oracle.OracleScene.from_irtt(sd)


def _check_structure(bvh, sc):
    n = bvh.n_leaves
    nodes = bvh.nodes.view(np.uint32)
    boxes = bvh.nodes[:, :12].copy().view(np.float32).reshape(-1, 2, 6)
    leaf_seen = np.zeros(n, int)
    node_seen = np.zeros(max(n - 1, 0), int)
    tri_box = np.zeros((n, 6), np.float32)
    tris = bvh.tris.copy()
    pv = tris[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]].view(np.float32).reshape(n, 3, 3)
    tri_box[:, :3] = pv.min(1)
    tri_box[:, 3:] = pv.max(1)
    # leaf records carry the triangle's own vertices
    for j in range(n):
        t = bvh.tri_of_leaf[j]
        assert tris[j, 3] == t
        assert np.array_equal(pv[j], sc.xyz[sc.idx[t]])
    assert np.all(np.diff(bvh.morton.astype(np.int64)) >= 0)
    stack = [(0, None)]
    while stack:
        i, box = stack.pop()
        node_seen[i] += 1
        for c in range(2):
            ch = int(nodes[i, 12 + c])
            cb = boxes[i, c]
            if box is not None:
                assert np.all(cb[:3] >= box[:3]) and np.all(cb[3:] <= box[3:])
            if ch & 0x80000000:
                leaf = ch & 0x7FFFFFFF
                leaf_seen[leaf] += 1
                assert np.array_equal(cb, tri_box[leaf])
            else:
                stack.append((ch, cb))
    assert np.all(leaf_seen == 1) and np.all(node_seen == 1)


@pytest.mark.parametrize("which", ["plates", "box", "city_small"])
def test_lbvh_invariants(oracle, which):
    from irtt import scenes
    sd = {"plates": lambda: scenes.parallel_plates(),
          "box": lambda: scenes.closed_box(3),
          "city_small": lambda: scenes.city_block(seed=2, nx=2, ny=2, target_tris=3000)}[which]()
    sc = _scene(oracle, sd)
    bvh = oracle.OracleBVH(sc)
    assert bvh.n_leaves == sd.nt
    _check_structure(bvh, sc)


def test_bvh_traversal_equals_brute_force(oracle):
    from irtt import scenes
    sd = scenes.city_block(seed=3, nx=2, ny=2, target_tris=2500)
    sc = _scene(oracle, sd)
    bvh = oracle.OracleBVH(sc)
    rng = np.random.default_rng(0)
    lo = sd.xyz.min(0) - 5
    hi = sd.xyz.max(0) + 5
    rays = random_rays(rng, 4000, lo, hi)
    # grazing / edge rays: aim exactly at vertices and edge midpoints
    tgt = np.concatenate([sd.xyz[rng.integers(0, sd.nv, 1000)],
                          0.5 * (sd.xyz[sd.idx[:1000, 0]] + sd.xyz[sd.idx[:1000, 1]])])
    o = tgt + rng.normal(size=tgt.shape) * 10
    d = tgt - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.concatenate([rays, np.concatenate([o, d], 1).astype(np.float32)])
    p1, t1, u1, v1, _ = bvh.trace(rays)
    p2, t2, u2, v2, ties = oracle.trace_brute(sc, rays)
    assert np.array_equal(p1, p2)
    assert np.array_equal(t1, t2) and np.array_equal(u1, u2) and np.array_equal(v1, v2)
    hit = p1 >= 0
    edge = hit & ((u1 == 0) | (v1 == 0) | (u1 + v1 == 1))
    print(f"rays {len(rays)} hits {hit.sum()} ties {ties} exact-edge hits {edge.sum()}")
    assert hit.sum() > 1000
