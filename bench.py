"""Benchmark: photon-traced transport build on the synthetic city block.

Workload (one step): the full forward photon transport of the 8 x 8 city block
(~51.8k triangles, vertex bins, 1296 photons per triangle = 67.1M photons
~ 2^26): fixed-point accumulator clear, photon tracing through the on-device
LBVH, FP32 dense operator + diagonal + exact column sums.  Geometry and the
BVH are resident in HBM before the timed region.

Prints one JSON line: value = traced ray segments per second over the whole
step (Mrays/s); ``roofline`` for the photon kernel (algorithmic bytes per ray
segment from the kernel's own node/triangle visit counts, HIP-event kernel
time); ``cpu_baseline`` = the CPU oracle (C + OpenMP restatement of the
reference tracer) on a bounded sample of the same scene, rank 0 only.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "inverse-radiative-thermal-transport_amd"))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s transport build; solves/s at ≤0.05K vs FP64; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
NODE_BYTES = 64      # two child boxes + two child refs
TRI_BYTES = 40       # 3 x float3 vertices + id
RAY_BYTES = 48       # 32 B ray in + 16 B hit out
BIN_BYTES = 8        # one 64-bit accumulator add per absorbed photon
# Per k_photons launch on the default workload (city block seed 0 with the
# traced sun quad, 130 x 8 photons per triangle, dense layout), from
# rocprofv3 counters; commands and raw values in
# profiles/k_photons_city_block_2p26.md.
# DRAM bytes: FETCH_SIZE (x 2.0) + WRITE_SIZE (x 1.0), the factors from a
# 1 GiB device copy (bench.py --calibrate-copy 1024).
# (The chunk marks' byte stores account for about 3 GB of it: each is a
# partial write of an 83 MB array that L2 reads and writes back whole.)
TRAFFIC_BYTES_CITY = (3667984 * 2.0 + 1924083 * 1.0) * 1024
# VALU wave-instructions issued (SQ_INSTS_VALU) and lane utilisation
# (SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)); the work is the same
# every launch (fixed seeds), so the count is a property of the workload.
VALU_INSTS_CITY = 5563105516
VALU_LANE_UTIL_CITY = 146286162270 / (64.0 * 5625671467)
# VALU issue peak: one wave64 instruction per cycle per CU (four SIMD16 each
# taking four cycles per instruction), 256 CUs at 2.4 GHz
VALU_PEAK_G = 256 * 2.4


def _log(msg):
    """Progress on stderr (the result line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _host_cpu():
    """(nproc, cores this process may use, CPU model string)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return os.cpu_count() or 1, usable, model


def cpu_baseline(sd, threads_per_tri, rays_per_thread, check=None, target_s=12.0):
    """Times the oracle tracer (all photons of a strided subset of source
    triangles) on the host cores, and, with `check` (the stored operator and
    the GPU solutions), an FP64 CPU solve of the same operator: the
    reference's CPU solve, and the independent FP64 answer each GPU form's
    max |dT| is measured against.  Returns a dict for the JSON line."""
    import numpy as np
    import oracle
    nproc, usable, model = _host_cpu()
    # the GPU box allots 16 cores to a one-GPU job, whatever nproc shows
    cores = max(1, min(16, usable))
    sc = oracle.OracleScene.from_irtt(sd)
    p = oracle.make_params(threads_per_tri, rays_per_thread, batch_seed=1234, bin_mode=0)
    stride = 512
    # all photons of every stride-th source triangle; shrink the stride until
    # the sample takes about target_s
    while True:
        t0 = time.perf_counter()
        _, _, st = oracle.transport(sc, p, 0, sd.nt, dense=False, records=1, n_threads=cores,
                                    tri_stride=stride)
        dt = time.perf_counter() - t0
        if dt >= 0.3 * target_s or stride == 1:
            break
        stride = max(1, int(stride * dt / target_s))
    rays, photons = st["rays"], st["photons"]
    out = {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
           "nproc": nproc, "usable_cores": usable, "cpu_model": model,
           "sample": f"{photons} photons ({rays} ray segments): all "
                     f"{threads_per_tri * rays_per_thread} photons of every {stride}-th source "
                     f"triangle of the same city block, CPU oracle (C + OpenMP, {cores} "
                     f"threads, LBVH build included), {dt:.1f} s"}
    if check is not None:
        out["fp64"] = fp64_check(sd, check)
    return out


def fp64_check(sd, c):
    """max |dT| of every GPU solve form against an FP64 solve of the same
    stored operator on the CPU (scipy, Jacobi-preconditioned GMRES to a
    relative residual of 1e-14; oracle/sparse_fp64.py), over the scene's
    free bins (the sun quads are fixed)."""
    import numpy as np
    from oracle import physics, sparse_fp64
    rp, col, val, colsum = c["operator"]
    T = sparse_fp64.to_scipy(rp, col, val, len(colsum))
    fixed = sd.fixed.astype(bool)
    free = ~fixed
    kfix = sd.kelvin.astype(np.float64)
    src = physics.bins(sd, 0)[3]
    out = {}
    t0 = time.perf_counter()
    x64, res = sparse_fp64.steady_linear(T, sparse_fp64.rhs(src, colsum, 293.0), fixed, kfix)
    out["steady_solve_s"] = round(time.perf_counter() - t0, 2)
    out["steady_residual"] = float(res)
    for k in ("fp32", "fp16", "sparse"):
        if k in c:
            out[f"max_dT_{k}_K"] = float(np.abs(c[k] - x64)[free].max())
    if "batch" in c:
        scales, internal, X = c["batch"]
        sv = np.concatenate([q["vertices"] for q in sd.meta["suns"]])
        y = np.zeros(len(colsum))
        y[sv] = kfix[sv] ** 4
        col64 = -(T @ y)
        col64[sv] = 0.0
        k0 = kfix.copy()
        k0[sv] = 0.0
        worst = 0.0
        for k in sorted({0, 1, len(scales) // 2, len(scales) - 1}):
            b = sparse_fp64.rhs(scales[k] * col64 + internal[k], colsum, 293.0)
            xk, _ = sparse_fp64.steady_linear(T, b, fixed, k0)
            worst = max(worst, float(np.abs(X[k] - xk)[free].max()))
        out["max_dT_fp16_batch_K"] = worst
    if "transient" in c:
        (trp, tcol, tval, tcs), sdt, kel, sv, H, dt = c["transient"]
        Tt = sparse_fp64.to_scipy(trp, tcol, tval, len(tcs))
        _, _, mass, srct = physics.bins(sdt, 0)
        tfixed = sdt.fixed.astype(bool)
        x = np.where(tfixed, sdt.kelvin.astype(np.float64), 293.0)
        d = []
        t0 = time.perf_counter()
        for k in range(H.shape[0]):
            x[sv] = kel[k]
            x, _ = sparse_fp64.implicit_euler_step(Tt, sparse_fp64.rhs(srct, tcs, 293.0), mass,
                                                   x, dt, tfixed)
            d.append(float(np.abs(H[k] - x)[~tfixed].max()))
        out["transient_steps"] = H.shape[0]
        out["transient_fp64_s_per_step"] = round((time.perf_counter() - t0) / H.shape[0], 2)
        out["max_dT_transient_fp16_per_step_K"] = [float(f"{v:.3e}") for v in d]
    out["note"] = ("FP64 CPU solve of the GPU's stored operator (every product and sum FP64); "
                   "bar 0.05 K")
    return out


def _state0(sd, suns=True):
    """Initial temperatures: fixed bins (the sun quads) at their own values
    (or 0 K with suns=False, when the sun enters as a per-design source),
    every other bin 0 K."""
    import numpy as np
    k = np.where(sd.fixed.astype(bool), sd.kelvin, 0.0).astype(np.float32)
    if not suns:
        for q in sd.meta.get("suns", []):
            k[q["vertices"]] = 0.0
    return k


def design_sources(irtt, s, sd, R, seed=11):
    """R steady-state designs sharing one operator: design r has the sun at
    scale_r x its flux (scale_0 = 1, the rest U[0.5, 1.25]) and internal heat
    gains of U[0, 20] W/m^2 per bin area (none for design 0).  The sun enters
    as the source -T[:, sun] y_sun it produces (one operator product), with
    the sun bins held at 0 K.  Returns (sources R x n, scales, internal R x n)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    area, _, _ = s.bin_data()
    sv = np.concatenate([q["vertices"] for q in sd.meta["suns"]])
    y = np.zeros(s.n_bins)
    y[sv] = sd.kelvin[sv].astype(np.float64) ** 4
    col = -s.apply_operator(y)[0]
    col[sv] = 0.0
    scales = np.concatenate([[1.0], rng.uniform(0.5, 1.25, R - 1)])
    internal = rng.uniform(0.0, 20.0, (R, s.n_bins)) * area[None, :]
    internal[0] = 0.0
    internal[:, sv] = 0.0
    return scales[:, None] * col[None, :] + internal, scales, internal


def solve_rates(irtt, s, sd, reps=3, batch=32):
    """Forward steady solve + adjoint solve on the built operator, FP32 form
    and FP16 form (FP16 Krylov products, FP32-operator Newton residuals).
    Returns solves/s of each and the largest temperature difference; the
    FP32-form solve converges the FP64-accumulated residual of the FP32
    operator, i.e. it is the FP64 solution of that operator to the Newton
    tolerance (|r|^2 < f_tol)."""
    import numpy as np
    sp = irtt.solver_params(steady=True, ambient_kelvin=293.0, x_tol=1e-7, f_tol=1e-14,
                            lin_rtol=1e-10)
    _, ea, _ = s.bin_data()
    free = ~sd.fixed.astype(bool)
    out = {}
    xs = {}
    for name, op in (("fp32", irtt.OP_DENSE), ("fp16", irtt.OP_DENSE_F16)):
        s.select_operator(op)
        times = []
        for _ in range(reps):
            s.set_vertex_state(_state0(sd), sd.fixed)
            t0 = time.perf_counter()
            fst = s.solve(sp)
            x = s.temperatures()
            # objective gradient of sum_i A_i (x_i - mean x)^2 over the scene's
            # bins (thermal_objective.cpp:47-57)
            rhs = np.where(free, ea * (x - x[free].mean()), 0.0)
            lam, ast = s.adjoint_solve(rhs, params=sp)
            times.append(time.perf_counter() - t0)
        xs[name] = x
        t = min(times)
        out[name] = {"solves_per_s": round(1.0 / t, 2), "ms": round(t * 1e3, 2),
                     "forward_ms": round(fst["solve_ms"], 2),
                     "adjoint_ms": round(ast["solve_ms"], 2),
                     "newton_iters": fst["newton_iters"],
                     "linear_iters": fst["linear_iters"] + ast["linear_iters"],
                     "residual_sq": fst["residual_sq"]}
    # served form: R designs sharing the operator (sun intensity and internal
    # gains) solved in lockstep, forward and adjoint, FP16 MFMA products with
    # FP32-operator refinement
    R = batch
    s.select_operator(irtt.OP_DENSE)
    srcs, scales, internal = design_sources(irtt, s, sd, R)
    s.select_operator(irtt.OP_DENSE_F16)
    s.set_vertex_state(_state0(sd, suns=False), sd.fixed)
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        X, fst = s.solve_batch(srcs, params=sp)
        rhs = np.where(free[None, :], ea[None, :] * (X - X[:, free].mean(axis=1, keepdims=True)),
                       0.0)
        lam, ast = s.adjoint_solve_batch(X, rhs, params=sp)
        times.append(time.perf_counter() - t0)
    t = min(times)
    s.select_operator(irtt.OP_DENSE)
    out["fp16_batch"] = {"solves_per_s": round(R / t, 2), "ms": round(t * 1e3, 2), "designs": R,
                         "forward_ms": round(fst["solve_ms"], 2),
                         "adjoint_ms": round(ast["solve_ms"], 2),
                         "newton_iters": fst["newton_iters"],
                         "linear_iters": fst["linear_iters"] + ast["linear_iters"],
                         "residual_sq_max": fst["residual_sq"]}
    out["x_fp32"] = xs["fp32"]
    out["x_fp16"] = xs["fp16"]
    out["batch_solutions"] = (scales, internal, X)
    out["max_dT_fp16_vs_fp32_K"] = float(np.abs(xs["fp16"] - xs["fp32"])[free].max())
    out["max_dT_fp16_batch_vs_fp32_K"] = float(np.abs(X[0] - xs["fp32"])[free].max())
    out["speedup_fp16"] = round(out["fp16"]["solves_per_s"] / out["fp32"]["solves_per_s"], 3)
    out["speedup_fp16_batch"] = round(out["fp16_batch"]["solves_per_s"] /
                                      out["fp32"]["solves_per_s"], 3)
    out["unit"] = "forward steady + adjoint solves per second (city block operator resident)"
    return out


def hmatrix_form(irtt, s, sd, sp, tol=1e-4, max_rank=32, reps=3):
    """Hierarchical block form of a scene's sparse-layout operator at `tol`
    and rank <= `max_rank` (FP16 factors with FP32 row scales): block
    outcomes, stored bytes against the sparse layout, |T - H|_F / |T|_F,
    compression time, and steady forward + adjoint solves with the H-matrix
    Krylov products (stored-entry residuals) against the same solves on the
    compressed rows, max |dT| between the two."""
    import numpy as np
    _, ea, _ = s.bin_data()
    free = ~np.asarray(sd.fixed, bool)

    def solves():
        best = None
        for _ in range(reps):
            s.set_vertex_state(_state0(sd), sd.fixed)
            t0 = time.perf_counter()
            fst = s.solve(sp)
            x = s.temperatures()
            lam, ast = s.adjoint_solve(np.where(free, ea * (x - x[free].mean()), 0.0), params=sp)
            t = time.perf_counter() - t0
            if best is None or t < best[0]:
                best = (t, x, fst["linear_iters"] + ast["linear_iters"])
        return best

    s.select_operator(irtt.OP_SPARSE)
    t_sp, x_sp, it_sp = solves()
    st = s.compress_hierarchical(tol=tol, max_rank=max_rank, precision=irtt.STORE_F16)
    t_h, x_h, it_h = solves()
    s.select_operator(irtt.OP_SPARSE)
    keys = ("n_clusters", "n_leaves", "n_far", "n_near", "n_lowrank", "n_zero", "n_noconv",
            "n_reject", "max_rank_used")
    out = {k: int(st[k]) for k in keys}
    out.update({
        "tol": tol, "max_rank": max_rank, "mean_rank": round(st["mean_rank"], 2),
        "nnz": int(st["nnz"]), "nnz_near": int(st["nnz_near"]),
        "nnz_in_lowrank": int(st["nnz_lowrank"]),
        "sparse_MB": round(st["bytes_sparse"] / 1e6, 1),
        "hmatrix_MB": round(st["bytes_hmatrix"] / 1e6, 1),
        "lowrank_MB": round(st["bytes_lowrank"] / 1e6, 2),
        "far_field_norm_share": round(st["fro_far_rel"], 4),
        "err_frob_rel_fp64_factors": float(f"{st['err_frob_rel']:.3e}"),
        "err_frob_rel_stored": float(f"{st['err_frob_rel_stored']:.3e}"),
        "compress_ms": round(st["build_ms"], 1), "aca_work_MB": round(st["build_bytes"] / 1e6, 1),
        "solves": {"sparse_ms": round(t_sp * 1e3, 2), "sparse_linear_iters": int(it_sp),
                   "hmatrix_ms": round(t_h * 1e3, 2), "hmatrix_linear_iters": int(it_h),
                   "max_dT_hmatrix_vs_sparse_K": float(np.abs(x_h - x_sp)[free].max())}})
    return out


def sparse_layout(irtt, sd, p, x_ref, reps=3, batch=32):
    """The same photon transport stored in the sparse layout (device hash
    table -> compressed rows and columns, entries bit-identical to the dense
    ones): build time, stored entries, and forward steady + adjoint solves/s
    on it, single and R designs in lockstep, with max |dT| against the
    dense FP32-form solution x_ref."""
    import ctypes
    import numpy as np
    q = irtt.TransportParams()
    ctypes.pointer(q)[0] = p
    q.layout = irtt.LAYOUT_SPARSE
    sp = irtt.solver_params(steady=True, ambient_kelvin=293.0, x_tol=1e-7, f_tol=1e-14,
                            lin_rtol=1e-10)
    out = {}
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_bvh()
        s.build_transport(q)
        walls, stats = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            stats.append(s.build_transport(q))
            walls.append(time.perf_counter() - t0)
        st = stats[-1]
        n = s.n_bins
        out["build"] = {"ms": round(min(walls) * 1e3, 2),
                        "photon_kernel_ms": round(min(x["trace_ms"] for x in stats), 3),
                        "rows_columns_ms": round(min(x["finalize_ms"] for x in stats), 3),
                        "mrays_per_s": round(st["rays"] / (min(walls) * 1e6), 1),
                        "nnz": int(st["nnz"]), "fill": round(st["nnz"] / float(n) ** 2, 5),
                        "operator_MB": round(st["nnz"] * 16 / 1e6, 1),
                        "dense_operator_MB": round(n * ((n + 63) // 64 * 64) * 4 / 1e6, 1)}
        _, ea, _ = s.bin_data()
        free = ~sd.fixed.astype(bool)
        times = []
        for _ in range(reps):
            s.set_vertex_state(_state0(sd), sd.fixed)
            t0 = time.perf_counter()
            fst = s.solve(sp)
            x = s.temperatures()
            lam, ast = s.adjoint_solve(np.where(free, ea * (x - x[free].mean()), 0.0), params=sp)
            times.append(time.perf_counter() - t0)
        t = min(times)
        out["x"] = x
        out["single"] = {"solves_per_s": round(1.0 / t, 2), "ms": round(t * 1e3, 2),
                         "linear_iters": fst["linear_iters"] + ast["linear_iters"],
                         "max_dT_vs_dense_fp32_K": float(np.abs(x - x_ref)[free].max())}
        srcs, _, _ = design_sources(irtt, s, sd, batch)
        s.set_vertex_state(_state0(sd, suns=False), sd.fixed)
        times = []
        for _ in range(2):
            t0 = time.perf_counter()
            X, fst = s.solve_batch(srcs, params=sp)
            rhs = np.where(free[None, :],
                           ea[None, :] * (X - X[:, free].mean(axis=1, keepdims=True)), 0.0)
            lam, ast = s.adjoint_solve_batch(X, rhs, params=sp)
            times.append(time.perf_counter() - t0)
        t = min(times)
        out["batch"] = {"solves_per_s": round(batch / t, 2), "ms": round(t * 1e3, 2),
                        "designs": batch,
                        "linear_iters": fst["linear_iters"] + ast["linear_iters"],
                        "max_dT_vs_dense_fp32_K": float(np.abs(X[0] - x_ref)[free].max())}
        # adjoint trace of the same photons (dL/dvertex with the rows
        # (lambda, x^4) and (-Ta^4, lambda)), absorbers kept by the build
        q.flags = irtt.TRANSPORT_KEEP_ABSORBERS
        s.build_transport(q)
        s.set_vertex_state(_state0(sd), sd.fixed)
        s.solve(sp)
        x = s.temperatures()
        lam, _ = s.adjoint_solve(np.where(free, ea * (x - x[free].mean()), 0.0), params=sp)
        rows_l = np.stack([lam, np.full(len(x), -293.0 ** 4)])
        rows_x = np.stack([x ** 4, lam])
        tr = []
        for _ in range(reps):
            g_replay, ast = s.adjoint_trace(rows_l, rows_x)
            tr.append(ast)
        ms = min(a["trace_ms"] for a in tr)
        out["adjoint_trace"] = {"ms": round(ms, 2), "photons": tr[-1]["photons"],
                                "segments": tr[-1]["segments"],
                                "edge_hits": tr[-1]["edge_hits"],
                                "occluded_edge_hits": tr[-1]["occluded"],
                                "segments_per_s_M": round(tr[-1]["segments"] / ms / 1e3, 1),
                                "forward_photon_kernel_ms": round(st["trace_ms"], 2)}
        # the same with the paths kept by the build (no replay in the trace):
        # builds (the first sizes the records at one per photon, the next
        # ones from its count), then traces of the last build's records
        q.flags = irtt.TRANSPORT_KEEP_PATHS
        kb, kwall = [], []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            kb.append(s.build_transport(q))
            kwall.append(time.perf_counter() - t0)
        ktr = []
        for _ in range(reps):
            g_kept, kst = s.adjoint_trace(rows_l, rows_x)
            ktr.append(kst)
        kms = min(a["trace_ms"] for a in ktr)
        out["adjoint_trace"]["kept_paths"] = {
            "trace_ms": round(kms, 2), "kept": ktr[-1]["kept_paths"],
            "build_ms": round(min(kwall[1:]) * 1e3, 2),
            "photon_kernel_ms": round(min(x["trace_ms"] for x in kb[1:]), 2),
            "build_ms_without": round(min(walls) * 1e3, 2),
            "gradient_bit_identical": bool(np.array_equal(g_kept, g_replay)),
            "build_plus_trace_ms": round(min(kwall[1:]) * 1e3 + kms, 2),
            "build_plus_trace_ms_replay": round(min(walls) * 1e3 + ms, 2)}
        _log("sparse layout: hierarchical blocks")
        out["hmatrix"] = hmatrix_form(irtt, s, sd, sp)
    out["unit"] = ("sparse layout of the city block operator: build, and forward steady + "
                   "adjoint solves per second (FP32 entries, FP64 sums)")
    return out


def facade(irtt, h=1e-3, h_wide=3e-2, gd_steps=50, photons=1 << 30, rays_per_thread=8,
           step_size=0.05, fd_params=(6, 93, 162, 167, 0, 109, 240, 255), fd_seeds=4):
    """Procedural facade (config 3): 20 buildings, 401k triangles plus the
    collimated sun quad (45 deg elevation, 135 deg azimuth, 800 W/m^2) so
    that balconies and fins shade walls and ground; 256 balcony / fin
    extrusion depths; about 2^30 photons into the sparse layout; steady
    solve; design gradient of the mean ground temperature (adjoint solve,
    adjoint trace, source term, projection) with the reference's terms and
    with the extensions (hit-area Jacobian, projected edge factor, beam
    transfer); central differences with common random numbers at h (the
    configured step) and h_wide on `fd_params` depths (balconies of sunlit
    +y / -x facades, of shaded -y facades, fins); then `gd_steps` Adam steps
    on the depths (projected onto [0, 1.5] m).  The h_wide differences and
    both gradients are repeated on `fd_seeds` independent photon sets and
    compared as means with standard errors."""
    import numpy as np
    from irtt import params as P, scenes
    from irtt.design import FacadeDesign
    t0 = time.perf_counter()
    _log("facade: generating the scene")
    sd = scenes.facade(sun_model="emitter")
    gen_s = time.perf_counter() - t0
    tpt = max(1, int(round(photons / (sd.nt * rays_per_thread))))
    tp = irtt.transport_params(tpt, rays_per_thread, batch_seed=1234,
                               layout=irtt.LAYOUT_SPARSE, flags=irtt.TRANSPORT_KEEP_ABSORBERS)
    sp = irtt.solver_params(steady=True, ambient_kelvin=293.0, x_tol=1e-6, f_tol=1e-12,
                            lin_rtol=1e-10)
    ref_flags = irtt.ADJ_AREA | irtt.ADJ_CONTINUOUS | irtt.ADJ_EDGE
    ext_flags = ref_flags | irtt.ADJ_HIT_AREA | irtt.ADJ_EDGE_PROJECTED | irtt.ADJ_COLLIMATED
    out = {"triangles": sd.nt, "scene_triangles": sd.meta["scene_triangles"], "bins": sd.nv,
           "params": int(sd.meta["n_params"]), "photons": tpt * rays_per_thread * sd.nt,
           "generate_s": round(gen_s, 2)}
    with FacadeDesign(irtt, sd, tp, sp, adjoint_flags=ref_flags) as fz:
        d0 = fz.depths.copy()
        fz.forward(d0)
        L0 = fz.forward(d0)
        fwd = dict(fz.timing)
        dp_ext = fz.gradient(flags=ext_flags)
        dp_ref = fz.gradient(flags=ref_flags)
        out.update({"objective_K": round(L0, 6), "forward": {k: (round(v, 2) if isinstance(
            v, float) else v) for k, v in fwd.items()},
            "gradient_ms": round(fz.timing["gradient_ms"], 1)})
        _log("facade: hierarchical blocks")
        out["hmatrix"] = hmatrix_form(irtt, fz.s, sd, sp, reps=1)
        ks = [int(k) for k in fd_params]
        _log(f"facade: objective {L0:.4f} K, gradient {fz.timing['gradient_ms']:.0f} ms; "
             f"FD on {len(ks)} depths")
        fd = fz.fd_gradient(d0, h, dims=ks)
        fdw = fz.fd_gradient(d0, h_wide, dims=ks)
        kind = {k: ("fin" if k >= 240 else
                    ("sunlit balcony" if (k // 3) % 4 in (2, 3) else "shaded balcony"))
                for k in ks}
        rows = [{"param": k, "kind": kind[k], "fd": round(float(fd[k]), 5),
                 "fd_wide": round(float(fdw[k]), 5),
                 "adjoint_reference_terms": round(float(dp_ref[k]), 5),
                 "adjoint_extended": round(float(dp_ext[k]), 5)} for k in ks]

        def agree(g, f):
            return int(np.sum(np.sign(g[ks]) == np.sign(f[ks])))
        # statistics over independent photon sets: at each seed the central
        # difference (common random numbers on both sides) and both adjoint
        # gradients; means with standard errors and the z-score of each
        # adjoint form against the FD mean
        base_seed = fz.tp.batch_seed
        per = {"fd": [], "ref": [], "ext": []}
        for i in range(fd_seeds):
            fz.tp.batch_seed = base_seed + 7919 * i
            if i == 0:
                per["fd"].append(fdw[ks])
                per["ref"].append(dp_ref[ks])
                per["ext"].append(dp_ext[ks])
                continue
            _log(f"facade: FD statistics, photon set {i + 1} of {fd_seeds}")
            fz.forward(d0)
            per["ext"].append(fz.gradient(flags=ext_flags)[ks])
            per["ref"].append(fz.gradient(flags=ref_flags)[ks])
            per["fd"].append(fz.fd_gradient(d0, h_wide, dims=ks)[ks])
        fz.tp.batch_seed = base_seed
        stat = {}
        for key, v in per.items():
            v = np.array(v)
            stat[key] = (v.mean(0), v.std(0, ddof=1) / np.sqrt(len(v)) if len(v) > 1
                         else np.zeros(len(ks)))

        def z(key):
            # paired over the photon sets: adjoint and FD share each set's photons
            d = np.array(per[key]) - np.array(per["fd"])
            if len(d) < 2:
                return np.full(len(ks), np.inf)
            se = d.std(0, ddof=1) / np.sqrt(len(d))
            return np.where(se > 0, d.mean(0) / np.where(se > 0, se, 1.0), np.inf)
        zr, ze = z("ref"), z("ext")
        seed_rows = [{"param": k, "fd_mean": round(float(stat["fd"][0][i]), 5),
                      "fd_se": round(float(stat["fd"][1][i]), 5),
                      "ref_mean": round(float(stat["ref"][0][i]), 5),
                      "ref_se": round(float(stat["ref"][1][i]), 5),
                      "ext_mean": round(float(stat["ext"][0][i]), 5),
                      "ext_se": round(float(stat["ext"][1][i]), 5),
                      "z_ref": round(float(zr[i]), 2), "z_ext": round(float(ze[i]), 2)}
                     for i, k in enumerate(ks)]
        out["fd_statistics"] = {
            "h_m": h_wide, "photon_sets": fd_seeds, "rows": seed_rows,
            "within_3se_reference_terms": int(np.sum(np.abs(zr) <= 3.0)),
            "within_3se_extended": int(np.sum(np.abs(ze) <= 3.0)),
            "sign_agreement_of_means": {"reference_terms": int(np.sum(
                np.sign(stat["ref"][0]) == np.sign(stat["fd"][0]))), "extended": int(np.sum(
                    np.sign(stat["ext"][0]) == np.sign(stat["fd"][0])))},
            "note": "per photon set (seed): central difference with common random numbers "
                    "and both adjoint gradients at the same photons; mean and standard "
                    "error over the sets; z = mean of (adjoint - FD) over the sets / its "
                    "standard error (paired)"}
        # the reference's FD mode 2 (transport only, lambda and x held at d0)
        # beside the adjoint trace of the same term, term by term
        _log("facade: transport-only FD and the trace's terms")
        tfd, _ = fz.transport_fd(d0, h_wide, dims=ks)
        lam_t, x4_t = fz.lam.copy(), fz.x.astype(np.float64) ** 4
        trip = fz.par.triplets(d0)
        terms = {"area": irtt.ADJ_AREA, "continuous": irtt.ADJ_CONTINUOUS,
                 "edge": irtt.ADJ_EDGE, "edge_projected": irtt.ADJ_EDGE | irtt.ADJ_EDGE_PROJECTED,
                 "hit_area": irtt.ADJ_HIT_AREA, "reference_terms": ref_flags,
                 "extended": ext_flags}
        tterm = {}
        for name, fl in terms.items():
            g, _ = fz.s.adjoint_trace(lam_t, x4_t, flags=fl)
            tterm[name] = [round(float(v), 5) for v in irtt.project_gradient(g, *trip)[ks]]
        out["transport_fd_terms"] = {
            "h_m": h_wide, "params": ks, "fd": [round(float(v), 5) for v in tfd[ks]],
            "trace": tterm,
            "note": "-lambda^T dT/dp x^4 with lambda and x held at the design (FD mode 2, "
                    "common random numbers) and the projected adjoint trace of the same "
                    "quantity per term set: the occluder's gain (hit_area) and the shadow's "
                    "loss (edge_projected) are each several times the net derivative"}
        out["fd_check"] = {
            "h_m": h, "h_wide_m": h_wide, "rows": rows,
            "sign_agreement_reference_terms": {"fd": agree(dp_ref, fd),
                                               "fd_wide": agree(dp_ref, fdw)},
            "sign_agreement_extended": {"fd": agree(dp_ext, fd), "fd_wide": agree(dp_ext, fdw)},
            "note": "extended = reference terms + hit-triangle area Jacobian + projected edge "
                    "factor + beam transfer of the sun's first segments (IRTT_ADJ_HIT_AREA, "
                    "IRTT_ADJ_EDGE_PROJECTED, IRTT_ADJ_COLLIMATED); FD with common random "
                    "numbers (same seed and photon counts)"}
        # the net derivative is a small difference (a deeper sunlit balcony
        # gains what the wall below loses); the descent uses the reference's
        # terms, whose omissions both point the same way (no occluder gain,
        # too small grazing edge factor)
        t1 = time.perf_counter()
        opt = P.Adam(step_size=step_size)
        ms = {"transport_ms": [], "solve_ms": [], "gradient_ms": []}

        def log(k, L, timing):
            for key in ms:
                ms[key].append(timing[key])
            if k % 10 == 0:
                _log(f"facade descent step {k}: objective {L:.4f} K")
        hist = fz.run(d0, opt, gd_steps, log=log, lo=0.0, hi=1.5)
        dt = time.perf_counter() - t1
        L = hist["objective"]
        out["descent"] = {"optimizer": "Adam", "step_size_m": step_size, "steps": gd_steps,
                          "gradient": "reference terms",
                          "objective_K_every_5": [round(float(v), 5) for v in L[::5]],
                          "start_K": round(float(L[0]), 5), "last_K": round(float(L[-1]), 5),
                          "best_K": round(float(opt.best_value), 5),
                          "s_per_step": round(dt / max(gd_steps, 1), 2),
                          "mean_ms": {k: round(float(np.mean(v)), 1) for k, v in ms.items()},
                          "depth_bounds_m": [0.0, 1.5]}
    return out


def transient(irtt, seed, p, steps=288, dt=300.0, check_steps=6, site=(48.21, 16.37, 2),
              date=(2024, 6, 21)):
    """The city block over 24 h: `steps` implicit-Euler steps of dt, per-patch
    areal heat capacity U[1e4, 1e5] J/(m^2 K), ambient 293 K.  The sun moves
    over the day: one collimated sun quad per daylight hour at the solar
    position of `site` (latitude, east longitude, UTC offset) on `date`
    (irtt.sun.solar_position), all in one operator; at time t the quads of
    the two enclosing hours emit the clear-sky direct irradiance DNI(t)
    split linearly between them (their fixed temperatures change between
    steps, the operator does not).  FP16 form (FP16 Krylov products,
    FP32-operator residuals with FP64 accumulation) against the FP32 form,
    max |dT| between them per step."""
    import numpy as np
    from irtt import scenes, sun
    lat, lon, utc = site
    base = scenes.city_block(seed=seed, sun_model=None)
    hours, suns = [], []
    for h in range(24):
        el, az, zen = sun.solar_position(lat, lon, utc, *date, h)
        if el > 1.0:
            hours.append(h)
            suns.append((sun.sun_direction(el, az, compass=True), sun.clear_sky_dni(zen)))
    per_sun = 2 * int(round(base.nt / (4 * len(hours))))
    sd = sun.add_suns(base, suns, triangles_per_sun=per_sun)
    hidx = {h: i for i, h in enumerate(hours)}
    fluxes = np.zeros((steps, len(hours)))
    for k in range(steps):
        t = (k + 1) * dt / 3600.0
        h0 = int(np.floor(t))
        el, _, zen = sun.solar_position(lat, lon, utc, *date, h0 % 24, (t - h0) * 60.0)
        if el <= 0.0:
            continue
        dni, w = sun.clear_sky_dni(zen), t - h0
        for h, wt in ((h0, 1.0 - w), (h0 + 1, w)):
            if h in hidx:
                fluxes[k, hidx[h]] += wt * dni
    sv = np.concatenate([q["vertices"] for q in sd.meta["suns"]])
    kel = np.stack([sun.sun_vertex_kelvin(sd, f)[1] for f in fluxes])
    sp = irtt.solver_params(steady=False, dt=dt, ambient_kelvin=293.0, x_tol=1e-6, f_tol=1e-14,
                            lin_rtol=1e-8)
    out, hist = {}, {}
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        t0 = time.perf_counter()
        st = s.build_transport(p)
        out["build_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        operator = s.transport_sparse() + (s.column_sums(),)
        for name, op in (("fp32", irtt.OP_DENSE), ("fp16", irtt.OP_DENSE_F16)):
            s.select_operator(op)
            s.set_vertex_state(np.where(sd.fixed.astype(bool), sd.kelvin, 293.0)
                               .astype(np.float32), sd.fixed)
            H = np.empty((steps, s.n_bins))
            t0 = time.perf_counter()
            iters = 0
            for k in range(steps):
                # this step's sun temperatures (fixed bins), the rest of the
                # state carried over exactly
                x = s.temperatures()
                x[sv] = kel[k]
                s.set_vertex_state(x.astype(np.float32), sd.fixed)
                s.set_temperatures(x)
                r = s.solve(sp)
                iters += r["linear_iters"]
                H[k] = s.temperatures()
            t = time.perf_counter() - t0
            hist[name] = H
            out[name] = {"ms_per_step": round(t * 1e3 / steps, 2), "linear_iters": iters}
    free = ~sd.fixed.astype(bool)
    out["first_steps"] = (operator, sd, kel[:check_steps], sv, hist["fp16"][:check_steps], dt)
    d = np.abs(hist["fp16"] - hist["fp32"])[:, free].max(axis=1)
    out.update({"steps": steps, "dt_s": dt, "site": {"lat": lat, "lon": lon, "utc": utc,
                                                      "date": "%04d-%02d-%02d" % date},
                "sun_hours": hours, "triangles": sd.nt, "sun_triangles": per_sun * len(hours),
                "photons": st["photons"],
                "max_dT_fp16_vs_fp32_K": float(d.max()),
                "max_dT_per_step_K": [float(f"{v:.3e}") for v in d[::24]],
                "T_range_K": [round(float(hist["fp32"][:, free].min()), 2),
                              round(float(hist["fp32"][:, free].max()), 2)],
                "unit": "implicit-Euler steps of the city block with a moving sun, FP32-form "
                        "solution = FP64-accumulated residual of the FP32 operator "
                        "(max_dT_per_step every 2 h)"})
    return out


def _city_designs(n, tmin, tmax, seed):
    """n city-block design variants with the traced sun, log-uniform
    tmin..tmax scene triangles, random heights and seeds, each with its 64
    building-height parameters."""
    import numpy as np
    from irtt import scenes
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        t = float(np.exp(rng.uniform(np.log(tmin), np.log(tmax))))
        g = int(np.clip(round(np.sqrt(t / 800.0)), 2, 8))
        sd = scenes.city_block(seed=1000 + k, nx=g, ny=g, target_tris=int(t), sun_model="emitter")
        out.append((sd, scenes.height_parametrization(sd, 64)))
    return out


def _stream(irtt, designs, arrivals, tp, sp, slots):
    """Submits design k at arrivals[k] seconds (Poisson, or all at once) to
    a server and waits for all; returns (statuses, wall seconds)."""
    with irtt.Server(slots=slots) as sv:
        sd, par = designs[0]   # one warm request: code loading is not a request's latency
        sv.wait(sv.submit(sd, transport=tp, solver=sp, kelvin0=sd.kelvin, fixed=sd.fixed,
                          target=300.0, objective_weight=1e-3, params=par))
        tickets = []
        t0 = time.perf_counter()
        for (sd, par), at in zip(designs, arrivals):
            dt = at - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets.append(sv.submit(sd, transport=tp, solver=sp, kelvin0=sd.kelvin,
                                     fixed=sd.fixed, target=300.0, objective_weight=1e-3,
                                     params=par))
        stats = [sv.wait(tk)[0] for tk in tickets]
        return stats, time.perf_counter() - t0


def serving(irtt, n_req=200, rate=8.0, slots=4, tmin=2000, tmax=200000, seed=7, tpt=64, rpt=8):
    """Poisson stream of design variants through irtt_serve_*: city blocks
    with the traced sun, log-uniform tmin..tmax scene triangles, random
    heights and seeds, each a forward steady solve plus adjoint over 64
    building-height parameters, operators in the sparse layout.  The same
    n_req designs at `rate` requests/s and at saturation (all submitted at
    once).  Returns latency percentiles and throughput of each."""
    import numpy as np
    t0 = time.perf_counter()
    designs = _city_designs(n_req, tmin, tmax, seed)
    gen_s = time.perf_counter() - t0
    tp = irtt.transport_params(threads_per_tri=tpt, rays_per_thread=rpt, batch_seed=1234,
                               layout=irtt.LAYOUT_SPARSE)
    sp = irtt.solver_params(ambient_kelvin=293.0, x_tol=1e-4, f_tol=1e-8)
    rng = np.random.default_rng(seed + 1)
    out = {"requests": n_req, "slots": slots, "generate_s": round(gen_s, 1),
           "triangles": {"min": int(min(d[0].nt for d in designs)),
                         "median": int(np.median([d[0].nt for d in designs])),
                         "max": int(max(d[0].nt for d in designs))},
           "photons_per_triangle": tpt * rpt}
    for name, arrivals in (("at_rate", np.cumsum(rng.exponential(1.0 / rate, n_req))),
                           ("saturation", np.zeros(n_req))):
        _log(f"serving: {n_req} requests, {name}")
        stats, wall = _stream(irtt, designs, arrivals, tp, sp, slots)
        lat = np.array([st["latency_ms"] for st in stats])
        run = np.array([st["run_ms"] for st in stats])
        out[name] = {"p50_ms": round(float(np.percentile(lat, 50)), 1),
                     "p99_ms": round(float(np.percentile(lat, 99)), 1),
                     "mean_ms": round(float(lat.mean()), 1),
                     "mean_run_ms": round(float(run.mean()), 1),
                     "mean_queue_ms": round(float(np.mean([st["queued_ms"] for st in stats])), 1),
                     "designs_per_s": round(n_req / wall, 2), "wall_s": round(wall, 2)}
        if name == "at_rate":
            out[name]["arrival_rate_per_s"] = rate
    out["work"] = ("per request: LBVH, %d photons per triangle into the sparse layout, steady "
                   "solve, adjoint solve, adjoint trace, dL/dp over 64 building heights"
                   % (tpt * rpt))
    return out


def served_vs_plain(irtt, n=12, tpt=130, rpt=8, slots=4):
    """n geometry variants of the 52k-triangle city block (random heights
    and seeds, traced sun), each a transport build, forward steady solve and
    design gradient (adjoint solve, trace, projection onto 64 heights).
    Plain: the FP32 GPU path one design after another (dense FP32 operator,
    scene API).  Served: the same requests through irtt_serve_* (sparse
    layout, FP32 entries with FP64 sums, `slots` designs on the device at
    once).  Returns designs/s of each, their ratio and the largest
    temperature difference between the two answers."""
    import numpy as np
    from irtt import scenes
    designs = []
    for k in range(n):
        sd = scenes.city_block(seed=2000 + k, sun_model="emitter")
        designs.append((sd, scenes.height_parametrization(sd, 64)))
    sp = irtt.solver_params(ambient_kelvin=293.0, x_tol=1e-7, f_tol=1e-14, lin_rtol=1e-10)
    tpd = irtt.transport_params(tpt, rpt, batch_seed=1234, flags=irtt.TRANSPORT_KEEP_ABSORBERS)
    _log(f"served vs plain: {n} city-block variants, plain FP32 path")
    xs = []
    t0 = time.perf_counter()
    for sd, par in designs:
        with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
            s.set_vertex_state(sd.kelvin, sd.fixed)
            s.build_transport(tpd)
            s.solve(sp)
            x = s.temperatures()
            area = s.bin_data()[0]
            s.design_gradient(1e-3 * area * (x - 300.0), params=par, solver=sp)
            xs.append(x)
    plain = n / (time.perf_counter() - t0)
    _log("served vs plain: served path")
    tps = irtt.transport_params(tpt, rpt, batch_seed=1234, layout=irtt.LAYOUT_SPARSE)
    stats, wall = _stream(irtt, designs, np.zeros(n), tps, sp, slots)
    served = n / wall
    dmax = 0.0
    with irtt.Server(slots=slots) as sv:
        for (sd, par), x in zip(designs[:2], xs[:2]):
            _, o = sv.wait(sv.submit(sd, transport=tps, solver=sp, kelvin0=sd.kelvin,
                                     fixed=sd.fixed, target=300.0, objective_weight=1e-3,
                                     params=par))
            dmax = max(dmax, float(np.abs(o["kelvin"] - x).max()))
    return {"designs": n, "triangles": designs[0][0].nt, "photons_per_triangle": tpt * rpt,
            "plain_fp32_designs_per_s": round(plain, 2), "served_designs_per_s": round(served, 2),
            "ratio": round(served / plain, 2), "served_p99_ms": round(float(np.percentile(
                [st["latency_ms"] for st in stats], 99)), 1),
            "max_dT_served_vs_plain_K": dmax,
            "work": "per design: transport build, forward steady solve, adjoint solve, adjoint "
                    "trace, dL/dp over 64 heights; plain = dense FP32 operator one design at a "
                    "time, served = sparse layout, %d concurrent" % slots}


def calibrate_copy(mib, reps=3):
    """Copies mib MiB device to device `reps` times (hipMemcpy's copy kernel):
    rocprofv3's FETCH_SIZE / WRITE_SIZE per dispatch of this copy, divided by
    mib MiB, give the factors applied to the kernels' counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    for _ in range(reps):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) != 0:  # device to device
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"copied {mib} MiB x {reps}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--threads-per-tri", type=int, default=130)
    ap.add_argument("--rays-per-thread", type=int, default=8)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-transient", action="store_true")
    ap.add_argument("--no-sparse", action="store_true")
    ap.add_argument("--no-facade", action="store_true",
                    help="skip the 400k-triangle facade design section (about two minutes)")
    ap.add_argument("--facade-gd", type=int, default=50)
    ap.add_argument("--facade-only", action="store_true",
                    help="run only the facade section and print it")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=8.0)
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the serving sections and print them")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device three times (a plain copy kernel "
                         "of known size, for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by source triangles across GPUs by design only)",
              file=sys.stderr)
        sys.exit(2)

    import irtt
    from irtt import scenes

    if args.facade_only:
        print(json.dumps({"facade": facade(irtt, gd_steps=args.facade_gd)}), flush=True)
        return
    if args.serving_only:
        print(json.dumps({"serving": serving(irtt, n_req=args.serve_requests,
                                             rate=args.serve_rate),
                          "served_vs_plain": served_vs_plain(irtt)}), flush=True)
        return
    _log("city block: scene and warm-up builds")
    sd = scenes.city_block(seed=args.seed, sun_model="emitter")
    s = irtt.Scene.from_data(sd, irtt.BIN_VERTEX)
    s.build_bvh()
    p = irtt.transport_params(args.threads_per_tri, args.rays_per_thread, batch_seed=1234)
    for _ in range(args.warmup):
        s.build_transport(p)
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(s.build_transport(p))
    t1 = time.perf_counter()
    # a cold build for comparison: a fresh scene clears all of its accumulators
    # and writes all of T (repeated builds of one scene reset only the chunks
    # the previous build touched)
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s_cold:
        s_cold.build_bvh()
        tc = time.perf_counter()
        s_cold.build_transport(p)
        cold_ms = (time.perf_counter() - tc) * 1e3
    # the stored operator, for the FP64 check of the solves below
    check = {"operator": s.transport_sparse() + (s.column_sums(),)}
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    st = stats[-1]
    rays = st["rays"]
    value = rays / (ms_per_step * 1e3)  # Mrays/s
    trace_ms = sum(x["trace_ms"] for x in stats) / len(stats)
    n_node = st["node_visits"] / rays
    n_tri = st["tri_tests"] / rays
    b_launch = rays * (RAY_BYTES + n_node * NODE_BYTES + n_tri * TRI_BYTES) + \
        (st["absorbed"] + st["forced_absorb"]) * BIN_BYTES
    achieved = b_launch / (trace_ms * 1e-3) / 1e9
    default = args.seed == 0 and args.threads_per_tri == 130 and args.rays_per_thread == 8
    # the photon kernel is bound by VALU issue (lanes of a wave diverge
    # between node visits, leaf tests and shading): achieved = its VALU
    # wave-instructions per second against the issue peak; the HBM figures
    # beside it show that DRAM is not the bound (the 5.8 MB tree is served
    # from L2 / MALL)
    valu = VALU_INSTS_CITY / (trace_ms * 1e-3) / 1e9 if default else None
    traffic = TRAFFIC_BYTES_CITY if default else None
    roofline = {
        "bound": "valu", "achieved": round(valu, 1) if valu else None, "peak": VALU_PEAK_G,
        "unit": "G VALU wave-instructions/s", "frac": round(valu / VALU_PEAK_G, 4) if valu else None,
        "traffic": round(traffic / 1e9, 3) if traffic else None,
        "traffic_unit": "GB DRAM per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
        "kernel": "k_photons", "kernel_ms": round(trace_ms, 3),
        "valu_lane_utilisation": round(VALU_LANE_UTIL_CITY, 4) if default else None,
        "valu_insts_per_launch": VALU_INSTS_CITY if default else None,
        "hbm": {"dram_GBs": round(traffic / (trace_ms * 1e-3) / 1e9, 1) if traffic else None,
                "dram_frac": round(traffic / (trace_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                if traffic else None,
                "algorithmic_GBs": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "algorithmic_bytes_per_ray": round(b_launch / rays, 1),
                "node_visits_per_ray": round(n_node, 2), "tri_tests_per_ray": round(n_tri, 2),
                "note": "algorithmic bytes (48 + 64 n_node + 40 n_tri per segment) exceed what "
                        "DRAM carries: node fetches are L2 / MALL hits"}}
    # steady solve on the built operator (reported beside the metric)
    sp = irtt.solver_params(steady=True, ambient_kelvin=293.0, x_tol=1e-7, f_tol=1e-14)
    s.set_vertex_state(sd.kelvin, sd.fixed)
    sol = s.solve(sp)
    _log(f"transport {ms_per_step:.2f} ms per build; solves")
    solves = solve_rates(irtt, s, sd)
    x_fp32 = solves.pop("x_fp32")
    check["fp32"] = x_fp32
    check["fp16"] = solves.pop("x_fp16")
    check["batch"] = solves.pop("batch_solutions")
    _log("transient")
    trans = None if args.no_transient else transient(irtt, args.seed, p)
    if trans is not None:
        check["transient"] = trans.pop("first_steps")
    sparse = None
    if not args.no_sparse:
        s.close()  # the dense operator (16 GB with its accumulators) is no longer needed
        _log("sparse layout")
        sparse = sparse_layout(irtt, sd, p, x_fp32)
        check["sparse"] = sparse.pop("x")
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "city block 8x8 buildings, seed %d, %d triangles (%d scene + %d "
                               "of the collimated sun quad at 45 deg elevation, 800 W/m^2), "
                               "%d vertex bins, %d photons (%d per triangle), forward transport "
                               "build" % (args.seed, sd.nt, sd.meta["scene_triangles"],
                                          sd.nt - sd.meta["scene_triangles"], s.n_bins,
                                          st["photons"],
                                          args.threads_per_tri * args.rays_per_thread),
                   "bins": s.n_bins, "triangles": sd.nt, "photons": st["photons"],
                   "ray_segments": rays},
        "roofline": roofline,
        "breakdown_ms": {"photon_kernel": round(trace_ms, 3),
                         "finalize": round(sum(x["finalize_ms"] for x in stats) / len(stats), 3),
                         "bvh_build": round(st["bvh_ms"], 3),
                         "cold_build": round(cold_ms, 3)},
        "steady_solve": {"ms": round(sol["solve_ms"], 1), "newton_iters": sol["newton_iters"],
                         "linear_iters": sol["linear_iters"], "converged": sol["converged"],
                         "power_in_W": round(sol["power_in"], 3),
                         "power_out_W": round(sol["power_out"], 3),
                         "balance_rel": float(f"{sol['balance_rel']:.3e}")},
        "energy": {"emitted_W_per_K4": st["emitted_w"], "absorbed_W_per_K4": st["absorbed_w"],
                   "escaped_W_per_K4": st["escaped_w"], "dropped_W_per_K4": st["dropped_w"],
                   "rel_err": float(f"{st['energy_rel']:.3e}"),
                   "integer_identity": bool(st["energy_exact"]),
                   "note": "FP64 on the device: photons x weights = absorbed + escaped + dropped "
                           "(exact in fixed point); rel_err is the FP32 rounding of the photon "
                           "weights"},
        "solves": solves,
    }
    if trans is not None:
        line["transient"] = trans
    if sparse is not None:
        line["sparse"] = sparse
    if not args.no_serving:
        # measured before the facade and serving sections, whose tens of GB
        # of differently sized buffers would otherwise sit in the device
        # memory cache that both the plain and the served path draw from
        s.close()
        line["served_vs_plain"] = served_vs_plain(irtt)
    if not args.no_facade:
        s.close()
        line["facade"] = facade(irtt, gd_steps=args.facade_gd)
    if not args.no_serving:
        _log("serving")
        line["serving"] = serving(irtt, n_req=args.serve_requests, rate=args.serve_rate)
    if not args.no_cpu_baseline:
        _log("CPU baseline and FP64 checks")
        line["cpu_baseline"] = cpu_baseline(sd, args.threads_per_tri, args.rays_per_thread,
                                            check=check)
        fp = line["cpu_baseline"]["fp64"]
        line["solves"]["max_dT_vs_fp64_K"] = {k[7:-2]: v for k, v in fp.items()
                                             if k.startswith("max_dT_") and k.endswith("_K")}
    s.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
