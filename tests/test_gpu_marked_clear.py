"""Repeated dense builds of one scene: the photon kernel marks each 16-entry
row chunk of the fixed-point accumulators it adds to; the next build zeroes
only the marked chunks, and the finalize pass reads only those.

A chunk that was added to and not marked, or marked and not cleared, would
leave the previous build's photon counts in the next operator.  The builds
therefore alternate batch seeds (different photons, different occupied
chunks) and switch layouts, and every fixed-point block, FP32 operator and
column sum must equal, bit for bit, what a fresh scene gives for the same
seed (a fresh scene clears its accumulators whole).  Bit equality is the bar:
the accumulators are integers."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bin_mode", ["vertex", "triangle"])
def test_repeated_dense_builds_bit_identical(bin_mode):
    import irtt
    from irtt import scenes
    mode = irtt.BIN_VERTEX if bin_mode == "vertex" else irtt.BIN_TRIANGLE
    sd = scenes.city_block(seed=3, nx=2, ny=2, target_tris=3000)
    # (seed, rays per thread, layout): the sparse build in the middle
    # releases the dense blocks, the dense build after it starts afresh
    plan = [(11, 8, "dense"), (12, 4, "dense"), (11, 8, "dense"), (13, 8, "sparse"),
            (12, 4, "dense"), (14, 2, "dense"), (11, 8, "dense")]
    ref = {}
    for seed, rpt, _ in plan:
        if (seed, rpt) not in ref:
            with irtt.Scene.from_data(sd, mode) as s:
                st = s.build_transport(irtt.transport_params(64, rpt, batch_seed=seed))
                acc, q = s.transport_fixed()
                ref[(seed, rpt)] = (acc.copy(), q, s.transport().copy(), s.column_sums().copy(), st)
    with irtt.Scene.from_data(sd, mode) as s:
        for k, (seed, rpt, layout) in enumerate(plan):
            lay = irtt.LAYOUT_SPARSE if layout == "sparse" else irtt.LAYOUT_DENSE
            st = s.build_transport(irtt.transport_params(64, rpt, batch_seed=seed, layout=lay))
            r_acc, r_q, r_T, r_col, r_st = ref[(seed, rpt)]
            assert st["rays"] == r_st["rays"] and st["absorbed"] == r_st["absorbed"]
            assert np.array_equal(s.column_sums(), r_col), f"build {k}: column sums differ"
            if layout == "sparse":
                continue
            acc, q = s.transport_fixed()
            assert q == r_q
            assert np.array_equal(acc, r_acc), f"build {k} (seed {seed}): accumulators differ"
            assert np.array_equal(s.transport(), r_T), f"build {k}: operator differs"
