"""IRTT_TRANSPORT_KEEP_PATHS: the transport build's photon kernel writes the
path records the adjoint trace's replay would write, and the trace uses them
instead of replaying the photons.  Checked against a build that keeps only
the absorbers (the trace then replays every photon): the gradients are
bit-identical (exact integer sums of the same per-record terms), every
counter is equal, the operator itself is unchanged -- for every term set, both
layouts, and when the kept records do not all fit (the trace replays from the
first photon whose record was dropped)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

AREA, CONT, EDGE, DROPPED, GAIN, HIT, PROJ, COLL = 1, 2, 4, 8, 16, 32, 64, 128
TERM_SETS = {
    "reference": AREA | CONT | EDGE,
    "extensions": AREA | CONT | EDGE | GAIN | HIT | PROJ | COLL,
    "reference_dropped": AREA | CONT | EDGE | DROPPED,
    "edges_only": EDGE | GAIN,
    "no_edges": AREA | CONT,
}
COUNTERS = ("photons", "absorbed", "segments", "continuous", "edge_hits", "occluded", "rays")


def _rows(n, seed):
    rng = np.random.default_rng(seed)
    lam = rng.normal(size=n)
    x4 = (290.0 + 10.0 * rng.random(n)) ** 4
    return np.stack([lam, -0.5 * lam]), np.stack([x4, 0.7 * x4])


def _operator(s, layout):
    if layout == "dense":
        return (s.transport(),)
    return s.transport_sparse()


def _check_same(irtt, a, b, layout):
    (st_a, op_a, tr_a), (st_b, op_b, tr_b) = a, b
    for k in ("rays", "absorbed", "miss", "too_close", "backface", "node_visits", "tri_tests"):
        assert st_a[k] == st_b[k], (k, st_a[k], st_b[k])
    for x, y in zip(op_a, op_b):
        assert np.array_equal(x, y)
    for name in TERM_SETS:
        (g_a, s_a), (g_b, s_b) = tr_a[name], tr_b[name]
        assert s_a["kept_paths"] == 0, (name, s_a)
        for k in COUNTERS:
            assert s_a[k] == s_b[k], (name, k, s_a[k], s_b[k])
        assert np.array_equal(g_a, g_b), (name, np.abs(g_a - g_b).max())
        print(f"{layout} {name}: bit-identical, kept_paths={s_b['kept_paths']}, "
              f"segments {s_b['segments']}, edge hits {s_b['edge_hits']}")


@pytest.mark.parametrize("layout", ["dense", "sparse"])
def test_kept_paths_equal_replay(irtt, layout):
    from irtt import scenes
    sd = scenes.city_block(seed=5, nx=3, ny=3, target_tris=4000, sun_model="emitter")
    lay = irtt.LAYOUT_DENSE if layout == "dense" else irtt.LAYOUT_SPARSE
    lam, x4 = _rows(sd.nv, 11)
    runs = []
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        # the second kept build sizes its records from the first one's count
        for keep in (irtt.TRANSPORT_KEEP_ABSORBERS, irtt.TRANSPORT_KEEP_PATHS,
                     irtt.TRANSPORT_KEEP_PATHS):
            st = s.build_transport(threads_per_tri=64, rays_per_thread=4, batch_seed=7,
                                   layout=lay, flags=keep)
            tr = {n: s.adjoint_trace(lam, x4, flags=f) for n, f in TERM_SETS.items()}
            runs.append((st, _operator(s, layout), tr))
        # another edge ratio than the kept records': replayed
        _, other = s.adjoint_trace(lam, x4, edge_ratio=0.08)
        assert other["kept_paths"] == 0
    for run in runs[1:]:
        _check_same(irtt, runs[0], run, layout)
    assert all(tr[1]["kept_paths"] == 1 for tr in runs[2][2].values())


def test_kept_paths_overflow_replays_the_rest(irtt):
    """A closed box of 90 % reflectors: about ten accepted segments per photon,
    so the first kept build's records (sized at one per photon) overflow and
    the trace replays from the first photon with a dropped record; the next
    build sizes its records from the count and keeps them all."""
    from irtt import scenes
    sd = scenes.closed_box(n=4, emissivity=0.1)
    lam, x4 = _rows(sd.nv, 12)
    runs = []
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        for keep in (irtt.TRANSPORT_KEEP_ABSORBERS, irtt.TRANSPORT_KEEP_PATHS,
                     irtt.TRANSPORT_KEEP_PATHS):
            st = s.build_transport(threads_per_tri=256, rays_per_thread=8, batch_seed=3,
                                   flags=keep)
            tr = {n: s.adjoint_trace(lam, x4, flags=f) for n, f in TERM_SETS.items()}
            runs.append((st, _operator(s, "dense"), tr))
    for run in runs[1:]:
        _check_same(irtt, runs[0], run, "dense")
    assert all(tr[1]["kept_paths"] == 2 for tr in runs[1][2].values())
    assert all(tr[1]["kept_paths"] == 1 for tr in runs[2][2].values())
