"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bar: bit-exact for the LBVH (Morton codes, topology, boxes), for closest-hit
triangle ids / distances / barycentrics, and for the fixed-point transport
accumulators (integer photon-weight sums); view factors within 4 Monte-Carlo
standard errors of the closed form."""
import numpy as np
import pytest

from conftest import random_rays

pytestmark = pytest.mark.gpu


def _scenes():
    from irtt import scenes
    return {
        "plates": lambda: scenes.parallel_plates(),
        "box": lambda: scenes.closed_box(3, emissivity=0.8),
        "city_small": lambda: scenes.city_block(seed=2, nx=2, ny=2, target_tris=3000),
        "city_full": lambda: scenes.city_block(seed=0),
        "city_deep": lambda: scenes.city_block(seed=2, nx=4, ny=4, target_tris=8000),
    }


@pytest.mark.parametrize("which", ["plates", "box", "city_small", "city_full"])
def test_lbvh_bitexact(irtt, oracle, which):
    sd = _scenes()[which]()
    print(f"[lbvh] {which} nt={sd.nt}", flush=True)
    with irtt.Scene(sd.xyz, sd.idx, sd.materials()) as s:
        nodes, tris, morton, depth = s.get_bvh()
    ob = oracle.OracleBVH(oracle.OracleScene.from_irtt(sd))
    assert np.array_equal(morton, ob.morton)
    assert np.array_equal(tris, ob.tris)
    assert np.array_equal(nodes, ob.nodes)
    assert depth == ob.max_depth


def test_trace_rays_bitexact(irtt, oracle):
    from irtt import scenes
    sd = scenes.city_block(seed=3, nx=2, ny=2, target_tris=2500)
    sc = oracle.OracleScene.from_irtt(sd)
    rng = np.random.default_rng(0)
    rays = random_rays(rng, 20000, sd.xyz.min(0) - 5, sd.xyz.max(0) + 5)
    tgt = np.concatenate([sd.xyz[rng.integers(0, sd.nv, 2000)],
                          0.5 * (sd.xyz[sd.idx[:2000, 0]] + sd.xyz[sd.idx[:2000, 1]])])
    o = tgt + rng.normal(size=tgt.shape) * 10
    d = tgt - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.concatenate([rays, np.concatenate([o, d], 1).astype(np.float32)])
    print(f"[trace] rays={len(rays)}", flush=True)
    with irtt.Scene(sd.xyz, sd.idx, sd.materials()) as s:
        gp, gt, g1, g2 = s.trace_rays(rays)
    ob = oracle.OracleBVH(sc)
    op, ot, o1, o2, _ = ob.trace(rays)
    bp, bt, b1, b2, ties = oracle.trace_brute(sc, rays)
    assert np.array_equal(gp, op) and np.array_equal(gt, ot)
    assert np.array_equal(g1, o1) and np.array_equal(g2, o2)
    hit = gp >= 0
    edge = hit & ((g1 == 0) | (g2 == 0) | (g1 + g2 == 1))
    print(f"hits {hit.sum()} brute-force ties {ties} exact edge/vertex hits {edge.sum()} "
          f"mismatch vs brute force {(gp != bp).sum()}")
    assert np.array_equal(gp, bp)


CASES = [
    # name, bin mode, threads_per_tri, rays_per_thread
    ("plates", 0, 64, 8),
    ("plates", 1, 64, 8),
    ("box", 1, 128, 4),
    ("city_small", 0, 16, 4),
    ("city_deep", 0, 8, 4),   # 18-level tree, deeper than the LDS part of the photon stack
]


@pytest.mark.parametrize("which,bin_mode,threads,rays", CASES)
def test_transport_fixed_point_bitexact(irtt, oracle, which, bin_mode, threads, rays):
    sd = _scenes()[which]()
    print(f"[transport] {which} bin={bin_mode} photons/tri={threads * rays}", flush=True)
    with irtt.Scene.from_data(sd, bin_mode) as s:
        st = s.build_transport(threads_per_tri=threads, rays_per_thread=rays, batch_seed=1234)
        acc, Q = s.transport_fixed()
        T = s.transport()
        colsum = s.column_sums()
        area, eps_area, _ = s.bin_data()
    sc = oracle.OracleScene.from_irtt(sd)
    p = oracle.make_params(threads, rays, batch_seed=1234, bin_mode=bin_mode)
    oacc, _, ost = oracle.transport(sc, p)
    _, _, oQ = oracle.photon_weights(sc, p)
    assert Q == oQ
    for k in ("photons", "rays", "absorbed", "miss", "too_close", "backface", "degenerate",
              "forced_absorb", "node_visits", "tri_tests"):
        assert st[k] == ost[k], (k, st[k], ost[k])
    assert np.array_equal(acc, oacc)
    # FP32 operator = diag(sigma eps A) - acc * Q, and exact column sums
    diag = 5.670374419e-8 * eps_area
    want = (np.diag(diag) - oacc.astype(np.float64) * Q).astype(np.float32)
    assert np.array_equal(T, want)
    assert np.allclose(colsum, diag - oacc.sum(0).astype(np.float64) * Q, rtol=0, atol=1e-12)


def test_view_factor_plates_full_size(irtt):
    """Config 0 geometry at its full photon count (2^20 per triangle, 2^30
    total), black plates in triangle bins: absorbed fraction on B = F_AB."""
    from irtt import scenes
    from oracle.closed_forms import parallel_rect_view_factor
    sd = scenes.parallel_plates(emissivity=1.0)
    na = sd.meta["plate_a_tris"]
    print("[viewfactor] 2^30 photons", flush=True)
    with irtt.Scene.from_data(sd, irtt.BIN_TRIANGLE) as s:
        st = s.build_transport(threads_per_tri=131072, rays_per_thread=8, batch_seed=0)
        acc, Q = s.transport_fixed()
    per = 131072 * 8
    # power absorbed on B over power emitted by A (sigma * eps * area, eps = 1)
    frac = float(acc[na:, :na].sum()) * Q / (5.670374419e-8 * 1.0)
    F = parallel_rect_view_factor(1, 1, 1)
    sig = np.sqrt(F * (1 - F) / (na * per))
    print(f"F_AB mc={frac:.6f} closed={F:.6f} sigma={sig:.2e} rays={st['rays']} "
          f"trace_ms={st['trace_ms']:.1f} Mrays/s={st['rays'] / st['trace_ms'] / 1e3:.1f}")
    assert abs(frac - F) < 4 * sig
    # reciprocity A_A F_AB = A_B F_BA (equal plates: F_AB = F_BA), the two
    # estimates independent; per triangle pair, A_j F_ij - A_i F_ji against
    # its Monte-Carlo noise (counts n_ij, n_ji: sigma ~ sqrt(n_ij + n_ji))
    frac_ba = float(acc[:na, na:].sum()) * Q / 5.670374419e-8
    rec = abs(frac - frac_ba)
    print(f"reciprocity: F_AB={frac:.6f} F_BA={frac_ba:.6f} |diff|={rec:.2e} "
          f"(sigma of the difference {np.sqrt(2) * sig:.2e})")
    assert rec < 4 * np.sqrt(2) * sig
    blk = acc[na:, :na]
    # every triangle has the same area, hence the same fixed-point photon
    # weight q: the gcd of the entries (counts x q)
    q = float(np.gcd.reduce(blk[blk > 0].ravel()))
    n_ij = blk.astype(np.float64) / q
    n_ji = acc[:na, na:].T.astype(np.float64) / q
    z = (n_ij - n_ji) / np.sqrt(np.maximum(n_ij + n_ji, 1.0))
    print(f"pairwise reciprocity z-scores: rms {np.sqrt((z ** 2).mean()):.3f}, "
          f"max {np.abs(z).max():.2f} over {z.size} pairs")
    assert 0.8 < np.sqrt((z ** 2).mean()) < 1.2
