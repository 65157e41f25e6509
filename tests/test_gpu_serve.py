"""The request loop on the device: each design's temperatures, objective and
dL/dp are bit-identical whether it runs alone, in a mix of concurrent
requests (up to a 52k-triangle city block beside four small ones), or
through the scene API directly (every kernel on the path is deterministic,
the adjoint trace's sums exact); admission respects the HBM budget; a
request's adjoint terms follow its flags; requests that share an operator
(one group) give the results of the same requests run one by one, to the
solver tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(2, 2, 1500, 0), (2, 3, 3000, 1), (3, 3, 5000, 2), (2, 2, 2000, 3), (3, 2, 4000, 4)]


def _designs():
    from irtt import scenes
    out = []
    for nx, ny, tris, seed in SIZES:
        sd = scenes.city_block(seed=seed, nx=nx, ny=ny, target_tris=tris)
        out.append((sd, scenes.height_parametrization(sd)))
    return out


def _params(irtt):
    tp = irtt.transport_params(threads_per_tri=64, rays_per_thread=4, batch_seed=5)
    sp = irtt.solver_params(ambient_kelvin=293.0, x_tol=1e-6, f_tol=1e-12)
    return tp, sp


def _same_gradient(a, b):
    """dL/dp: the adjoint trace sums in exact fixed point, so bit-identical."""
    np.testing.assert_array_equal(a, b)


def _submit(sv, irtt, sd, par, tp, sp):
    return sv.submit(sd, transport=tp, solver=sp, kelvin0=sd.kelvin, fixed=sd.fixed,
                     target=300.0, objective_weight=1e-3, params=par)


def test_results_same_alone_in_the_mix_and_direct(irtt):
    from irtt import scenes
    tp, sp = _params(irtt)
    designs = _designs()
    big = scenes.city_block(seed=9, target_tris=50000)
    designs.insert(2, (big, scenes.height_parametrization(big, 64)))
    assert big.nt > 50000
    alone = []
    with irtt.Server(slots=1) as sv:
        for sd, par in designs:
            alone.append(sv.wait(_submit(sv, irtt, sd, par, tp, sp)))
    with irtt.Server(slots=4) as sv:
        tickets = [_submit(sv, irtt, sd, par, tp, sp) for sd, par in designs]
        mix = [sv.wait(t) for t in tickets]
    for (sa, a), (sm, m), (sd, _) in zip(alone, mix, designs):
        print(f"nt={sd.nt} bins={sa['n_bins']}: alone {sa['run_ms']:.1f} ms, in the mix "
              f"{sm['run_ms']:.1f} ms run / {sm['latency_ms']:.1f} ms latency; "
              f"objective {a['objective']:.6e}, |dp|max {np.abs(a['dp']).max():.3e}")
        np.testing.assert_array_equal(a["kelvin"], m["kelvin"])
        assert a["objective"] == m["objective"]
        _same_gradient(a["dp"], m["dp"])
    # the same design through the scene API directly
    sd, par = designs[0]
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.set_vertex_state(sd.kelvin, sd.fixed)
        s.build_transport(tp)
        s.solve(sp)
        x = s.temperatures()
    np.testing.assert_array_equal(x, alone[0][1]["kelvin"])
    total_run = sum(st["run_ms"] for st, _ in mix)
    wall = max(st["latency_ms"] for st, _ in mix)
    print(f"mix: sum of run times {total_run:.1f} ms over {wall:.1f} ms wall")
    assert total_run > wall  # requests overlapped on the device


def test_admission_by_memory_budget(irtt):
    """A budget that holds the largest design once (plus a little) admits
    the others as memory frees; all finish and the results do not change."""
    tp, sp = _params(irtt)
    designs = _designs()
    with irtt.Server(slots=1) as sv:
        ref = [sv.wait(_submit(sv, irtt, sd, par, tp, sp)) for sd, par in designs]
    biggest = max(st["footprint_bytes"] for st, _ in ref)
    with irtt.Server(slots=4, memory_budget=int(biggest * 1.2)) as sv:
        tickets = [_submit(sv, irtt, sd, par, tp, sp) for sd, par in designs]
        got = [sv.wait(t) for t in tickets]
    waited = [st["queued_ms"] for st, _ in got]
    print(f"budget {biggest * 1.2 / 2**20:.0f} MiB, queue waits (ms): "
          f"{[round(w, 1) for w in waited]}")
    for (_, a), (_, b) in zip(ref, got):
        np.testing.assert_array_equal(a["kelvin"], b["kelvin"])
        _same_gradient(a["dp"], b["dp"])
    assert max(waited) > 0.0


def test_request_adjoint_flags(irtt):
    """The request's adjoint flags reach the trace: dL/dp with the hit-area
    and projected-edge terms equals the scene API's design gradient with the
    same flags, and differs from the default terms."""
    from irtt import scenes
    tp, sp = _params(irtt)
    sd, par = _designs()[1]
    flags = irtt.ADJ_AREA | irtt.ADJ_CONTINUOUS | irtt.ADJ_EDGE | irtt.ADJ_HIT_AREA | \
        irtt.ADJ_EDGE_PROJECTED
    with irtt.Server(slots=2) as sv:
        _, a = sv.wait(sv.submit(sd, transport=tp, solver=sp, kelvin0=sd.kelvin, fixed=sd.fixed,
                                 target=300.0, objective_weight=1e-3, params=par,
                                 adjoint_flags=flags))
        _, b = sv.wait(_submit(sv, irtt, sd, par, tp, sp))
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.set_vertex_state(sd.kelvin, sd.fixed)
        q = irtt.TransportParams()
        import ctypes
        ctypes.pointer(q)[0] = tp
        q.flags |= irtt.TRANSPORT_KEEP_ABSORBERS
        s.build_transport(q)
        s.solve(sp)
        x = s.temperatures()
        area = s.bin_data()[0]
        dp, _, _, _ = s.design_gradient(1e-3 * area * (x - 300.0), params=par, solver=sp,
                                        flags=flags)
    np.testing.assert_array_equal(a["dp"], dp)
    assert not np.array_equal(a["dp"], b["dp"])


def test_group_shares_the_operator(irtt):
    """Six designs on one geometry that differ in source (the materials' sun
    flux scaled, plus internal gains) run as one group: one scene and build,
    forward and adjoint solves in lockstep; each result equals the same
    request run alone to the solver tolerance (temperatures 1e-6 K, dL/dp
    1e-6 of its largest component)."""
    from irtt import scenes
    tp, sp0 = _params(irtt)
    sp = irtt.solver_params(ambient_kelvin=293.0, x_tol=1e-9, f_tol=1e-20, lin_rtol=1e-12)
    sd = scenes.city_block(seed=3, nx=2, ny=3, target_tris=3000)
    par = scenes.height_parametrization(sd)
    base = scenes.bin_sources(sd, sd.heat_flux)
    rng = np.random.default_rng(2)
    area = sd.triangle_areas()
    sources = [base * f + rng.uniform(0, 10, sd.nv) for f in (1.0, 0.5, 0.8, 1.2, 0.3, 0.9)]

    def submit(sv, src, group):
        return sv.submit(sd, transport=tp, solver=sp, kelvin0=sd.kelvin, fixed=sd.fixed,
                         target=300.0, objective_weight=1e-3, params=par, source=src,
                         group=group)
    with irtt.Server(slots=1) as sv:
        alone = [sv.wait(submit(sv, src, 0)) for src in sources]
    with irtt.Server(slots=1) as sv:
        # a blocker first so the group queues up behind it
        blk = submit(sv, sources[0], 0)
        tickets = [submit(sv, src, 77) for src in sources]
        sv.wait(blk)
        grouped = [sv.wait(t) for t in tickets]
    runs = {round(st["run_ms"], 6) for st, _ in grouped}
    print(f"group of {len(sources)}: run {sorted(runs)} ms; alone "
          f"{[round(st['run_ms'], 1) for st, _ in alone]} ms")
    # at least one request rode on another's scene (its memory counted there)
    assert sum(st["footprint_bytes"] == 0 for st, _ in grouped) >= 1
    for (_, a), (_, g) in zip(alone, grouped):
        assert np.abs(a["kelvin"] - g["kelvin"]).max() < 1e-6
        assert abs(a["objective"] - g["objective"]) <= 1e-6 * abs(a["objective"])
        assert np.abs(a["dp"] - g["dp"]).max() <= 1e-6 * np.abs(a["dp"]).max()
    assert area.size == sd.nt
