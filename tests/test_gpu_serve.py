"""The request loop on the device: each design's temperatures, objective and
dL/dp are bit-identical whether it runs alone, in a mix of concurrent
requests, or through the scene API directly (every kernel on the path is
deterministic); admission respects the HBM budget."""
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
    """dL/dp sums the adjoint trace's FP64 atomic additions, whose order is
    not fixed: equal to 1e-12 of the largest component (temperatures and
    objectives are bit-identical)."""
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()


def _submit(sv, irtt, sd, par, tp, sp):
    return sv.submit(sd, transport=tp, solver=sp, kelvin0=sd.kelvin, fixed=sd.fixed,
                     target=300.0, objective_weight=1e-3, params=par)


def test_results_same_alone_in_the_mix_and_direct(irtt):
    tp, sp = _params(irtt)
    designs = _designs()
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
