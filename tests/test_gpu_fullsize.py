"""The headline workload at full size against the oracle and against FP64.

City block (8 x 8 buildings, seed 0, 51 786 scene triangles, 36 475 vertex
bins) with the collimated sun quad (45 deg elevation, 800 W/m^2; 12 946
repeated triangles, 4 fixed bins), 1040 photons per triangle (67.3 M ~
2^26), the bench's configuration:
* the stored operator, dense and sparse layouts, bit for bit against the
  oracle tracer's per-photon records aggregated in exact integer sums (every
  entry, every column sum, every counter);
* the steady solves the bench times (FP32 form, FP16 form, 32 designs in
  lockstep with FP16 MFMA products -- sun intensity and internal gains per
  design, the sun entering as the source its operator column produces) against
  an FP64 solve of the same stored operator on the CPU (oracle/sparse_fp64.py)
  over the free bins, bar 0.05 K;
* the transient form (FP16 operator, FP64-accumulated FP32-operator
  residuals) over six implicit-Euler steps with the sun's intensity rising
  (its fixed temperature changed between steps), each step against an FP64
  Newton step of the same operator, bar 0.05 K."""
import os

import numpy as np
import pytest

from oracle import physics, sparse_fp64

pytestmark = pytest.mark.gpu

TA = 293.0
TPT, RPT = 130, 8
BAR_K = 0.05


def _say(msg):
    print(f"[fullsize] {msg}", flush=True)


@pytest.fixture(scope="module")
def headline(irtt, oracle):
    from irtt import scenes
    sd = scenes.city_block(seed=0, sun_model="emitter")
    _say(f"city block {sd.nt} triangles: GPU dense build")
    s = irtt.Scene.from_data(sd, irtt.BIN_VERTEX)
    st = s.build_transport(threads_per_tri=TPT, rays_per_thread=RPT, batch_seed=1234)
    rp, col, val = s.transport_sparse()
    colsum = s.column_sums()
    _, ea, mass = s.bin_data()
    _say(f"{st['photons']} photons, {col.size} entries; oracle tracer")
    sc = oracle.OracleScene.from_irtt(sd)
    p = oracle.make_params(TPT, RPT, batch_seed=1234, bin_mode=0)
    threads = min(16, os.cpu_count() or 1)
    _, recs, ost = oracle.transport(sc, p, dense=False, records=st["photons"],
                                    n_threads=threads, with_tri=True)
    _, q, Q = oracle.photon_weights(sc, p)
    _say("aggregating the oracle's records")
    orp, ocol, oval, ocolsum = sparse_fp64.csr_from_records(*recs, q, Q, ea, s.n_bins)
    del recs
    yield dict(sd=sd, s=s, st=st, ost=ost, gpu=(rp, col, val, colsum),
               oracle=(orp, ocol, oval, ocolsum), ea=ea, mass=mass)
    s.close()


def test_headline_operator_bit_exact(headline):
    st, ost = headline["st"], headline["ost"]
    for k in ("photons", "rays", "absorbed", "miss", "too_close", "backface", "degenerate",
              "forced_absorb", "node_visits", "tri_tests"):
        assert st[k] == ost[k], (k, st[k], ost[k])
    rp, col, val, colsum = headline["gpu"]
    orp, ocol, oval, ocolsum = headline["oracle"]
    _say(f"dense layout: {col.size} entries vs oracle {ocol.size}")
    assert np.array_equal(rp, orp)
    assert np.array_equal(col, ocol)
    assert np.array_equal(val, oval)
    assert np.array_equal(colsum, ocolsum)


def test_headline_sparse_layout_bit_exact(irtt, headline):
    sd = headline["sd"]
    _say("sparse layout build")
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s2:
        st = s2.build_transport(threads_per_tri=TPT, rays_per_thread=RPT, batch_seed=1234,
                                layout=irtt.LAYOUT_SPARSE)
        rp, col, val = s2.transport_sparse()
        colsum = s2.column_sums()
    orp, ocol, oval, ocolsum = headline["oracle"]
    assert st["nnz"] == ocol.size
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.array_equal(val, oval) and np.array_equal(colsum, ocolsum)


def test_headline_solves_vs_fp64(irtt, headline):
    sd, s = headline["sd"], headline["s"]
    orp, ocol, oval, ocolsum = headline["oracle"]
    n = s.n_bins
    fixed = sd.fixed.astype(bool)
    free = ~fixed
    kfix = sd.kelvin.astype(np.float64)
    T = sparse_fp64.to_scipy(orp, ocol, oval, n)
    _, _, _, src = physics.bins(sd, 0)
    b = sparse_fp64.rhs(src, ocolsum, TA)
    _say("FP64 steady solve on the CPU")
    x64, res = sparse_fp64.steady_linear(T, b, fixed, kfix)
    assert res < 1e-13
    sp = irtt.solver_params(steady=True, ambient_kelvin=TA, x_tol=1e-7, f_tol=1e-14,
                            lin_rtol=1e-10)
    state = np.where(fixed, sd.kelvin, 0.0).astype(np.float32)
    errs = {}
    for name, op in (("fp32", irtt.OP_DENSE), ("fp16", irtt.OP_DENSE_F16)):
        _say(f"GPU steady solve, {name} form")
        s.select_operator(op)
        s.set_vertex_state(state, sd.fixed)
        stt = s.solve(sp)
        assert stt["converged"]
        errs[name] = float(np.abs(s.temperatures() - x64)[free].max())
    # 32 designs in lockstep: sun intensity x U[0.5, 1.25] and internal gains
    # U[0, 20] W/m^2; the sun enters as its column's source with the sun
    # bins at 0 K (GPU: one operator product; FP64: the oracle's operator)
    R = 32
    rng = np.random.default_rng(11)
    sv = np.concatenate([q["vertices"] for q in sd.meta["suns"]])
    y = np.zeros(n)
    y[sv] = kfix[sv] ** 4
    s.select_operator(irtt.OP_DENSE)
    col = -s.apply_operator(y)[0]
    col[sv] = 0.0
    col64 = -(T @ y)
    col64[sv] = 0.0
    area = s.bin_data()[0]
    scales = np.concatenate([[1.0], rng.uniform(0.5, 1.25, R - 1)])
    internal = rng.uniform(0.0, 20.0, (R, n)) * area[None, :]
    internal[0] = 0.0
    internal[:, sv] = 0.0
    srcs = scales[:, None] * col[None, :] + internal
    k0 = kfix.copy()
    k0[sv] = 0.0
    _say("GPU batched steady solves (FP16 MFMA products)")
    s.select_operator(irtt.OP_DENSE_F16)
    s.set_vertex_state(np.where(fixed, k0, 0.0).astype(np.float32), sd.fixed)
    X, _ = s.solve_batch(srcs, params=sp)
    s.select_operator(irtt.OP_DENSE)
    worst = 0.0
    for k in (0, 1, R // 2, R - 1):
        bk = sparse_fp64.rhs(scales[k] * col64 + internal[k], ocolsum, TA)
        xk, _ = sparse_fp64.steady_linear(T, bk, fixed, k0)
        worst = max(worst, float(np.abs(X[k] - xk)[free].max()))
    errs["fp16_batch"] = worst
    _say(f"max |dT| vs FP64: {errs} (T in [{x64[free].min():.2f}, {x64[free].max():.2f}] K)")
    for k, e in errs.items():
        assert e < BAR_K, (k, e)


def test_transient_city_vs_fp64(irtt, headline):
    """Six implicit-Euler steps of 5 min from 293 K in the FP16 form, the
    sun's flux rising 200 -> 800 W/m^2 over them (fixed temperature of the sun
    bins set between steps)."""
    from irtt import sun
    sd, s = headline["sd"], headline["s"]
    orp, ocol, oval, ocolsum = headline["oracle"]
    T = sparse_fp64.to_scipy(orp, ocol, oval, s.n_bins)
    area, ea, mass, src = physics.bins(sd, 0)
    fixed = sd.fixed.astype(bool)
    dt, steps = 300.0, 6
    p = irtt.solver_params(steady=False, dt=dt, ambient_kelvin=TA, x_tol=1e-6, f_tol=1e-14,
                           lin_rtol=1e-8)
    s.select_operator(irtt.OP_DENSE_F16)
    x0 = np.where(fixed, sd.kelvin, TA).astype(np.float32)
    s.set_vertex_state(x0, sd.fixed)
    x64 = x0.astype(np.float64)
    worst = []
    for k in range(steps):
        v, kel = sun.sun_vertex_kelvin(sd, [200.0 + 120.0 * k])
        x = s.temperatures()
        x[v] = kel
        s.set_vertex_state(x.astype(np.float32), sd.fixed)
        s.set_temperatures(x)
        st = s.solve(p)
        assert st["converged"]
        x64[v] = kel
        x64, it = sparse_fp64.implicit_euler_step(T, sparse_fp64.rhs(src, ocolsum, TA), mass,
                                                  x64, dt, fixed)
        worst.append(float(np.abs(s.temperatures() - x64)[~fixed].max()))
        _say(f"step {k + 1}: max |dT| vs FP64 {worst[-1]:.2e} K ({it} FP64 Newton steps)")
    s.select_operator(irtt.OP_DENSE)
    assert max(worst) < BAR_K
