"""Sparse transport layout (IRTT_LAYOUT_SPARSE): compressed rows and columns
built from a device hash table of photon hits.

Pinned against the oracle's fixed-point accumulators: the stored entries are
exactly the (absorber, emitter) pairs some photon connected plus every
diagonal, each FP32(delta_ij sigma eps_i A_i - acc_ij Q), rows in order and
columns ascending — the numbers the dense layout stores, bit for bit.
Products, solves and adjoint solves on the sparse operator are checked
against FP64 numpy on the same FP32 entries (tolerances in each test)."""
import numpy as np
import pytest

from oracle import physics, solver

pytestmark = pytest.mark.gpu

SIGMA = 5.670374419e-8
TA = 293.0


def _scenes():
    from irtt import scenes
    return {
        "plates": lambda: scenes.parallel_plates(),
        "box": lambda: scenes.closed_box(3, emissivity=0.8),
        "city_small": lambda: scenes.city_block(seed=2, nx=2, ny=2, target_tris=3000),
        "city_deep": lambda: scenes.city_block(seed=2, nx=4, ny=4, target_tris=8000),
    }


def _csr_from_acc(acc, Q, eps_area):
    """Compressed rows of FP32(diag - acc Q): nonzero accumulators plus the diagonal."""
    n = acc.shape[0]
    mask = acc != 0
    np.fill_diagonal(mask, True)
    rows, cols = np.nonzero(mask)
    diag = SIGMA * eps_area
    vals = (np.where(rows == cols, diag[rows], 0.0) - acc[rows, cols].astype(np.float64) * Q)
    rp = np.zeros(n + 1, np.uint32)
    np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
    return rp, cols.astype(np.uint32), vals.astype(np.float32)


@pytest.mark.parametrize("which,bin_mode,threads,rays", [
    ("plates", 1, 64, 4),
    ("box", 0, 128, 4),
    ("city_small", 0, 16, 4),
    ("city_deep", 1, 8, 4),
])
def test_sparse_layout_matches_oracle(irtt, oracle, which, bin_mode, threads, rays):
    sd = _scenes()[which]()
    print(f"[sparse] {which} bin={bin_mode} photons/tri={threads * rays}", flush=True)
    with irtt.Scene.from_data(sd, bin_mode) as s:
        st = s.build_transport(threads_per_tri=threads, rays_per_thread=rays, batch_seed=1234,
                               layout=irtt.LAYOUT_SPARSE)
        rp, col, val = s.transport_sparse()
        T = s.transport()
        colsum = s.column_sums()
        _, eps_area, _ = s.bin_data()
        with pytest.raises(irtt.IrttError):
            s.transport_fixed()
    sc = oracle.OracleScene.from_irtt(sd)
    p = oracle.make_params(threads, rays, batch_seed=1234, bin_mode=bin_mode)
    oacc, _, ost = oracle.transport(sc, p)
    _, _, Q = oracle.photon_weights(sc, p)
    for k in ("photons", "rays", "absorbed", "miss", "too_close", "backface", "degenerate",
              "forced_absorb", "node_visits", "tri_tests"):
        assert st[k] == ost[k], (k, st[k], ost[k])
    wrp, wcol, wval = _csr_from_acc(oacc, Q, eps_area)
    fill = wcol.size / float(oacc.shape[0]) ** 2
    print(f"  nnz {wcol.size} ({100 * fill:.2f} % of n^2), trace {st['trace_ms']:.2f} ms, "
          f"rows+columns {st['finalize_ms']:.2f} ms")
    assert st["layout"] == irtt.LAYOUT_SPARSE and st["nnz"] == wcol.size
    assert np.array_equal(rp, wrp)
    assert np.array_equal(col, wcol)
    assert np.array_equal(val, wval)
    # scattered back it is the dense layout's matrix
    want = (np.diag(SIGMA * eps_area) - oacc.astype(np.float64) * Q).astype(np.float32)
    assert np.array_equal(T, want)
    assert np.allclose(colsum, SIGMA * eps_area - oacc.sum(0).astype(np.float64) * Q, rtol=0,
                       atol=1e-12)


def test_dense_layout_compressed_equals_sparse_layout(irtt):
    """irtt_get_transport_sparse of the dense layout (compressed on the host)
    and of the sparse layout (built on the device) agree entry for entry."""
    sd = _scenes()["city_small"]()
    out = []
    for layout in (irtt.LAYOUT_DENSE, irtt.LAYOUT_SPARSE):
        with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
            s.build_transport(threads_per_tri=32, rays_per_thread=4, batch_seed=7, layout=layout)
            out.append(s.transport_sparse())
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def _city(irtt, layout):
    from irtt import scenes
    sd = scenes.city_block(seed=6, nx=2, ny=2, target_tris=2500)
    s = irtt.Scene.from_data(sd, irtt.BIN_VERTEX)
    s.build_transport(threads_per_tri=128, rays_per_thread=4, batch_seed=2, layout=layout)
    return sd, s


@pytest.mark.parametrize("R", [1, 5, 37, 64, 100])
def test_sparse_products_are_fp64_products(irtt, R):
    """Y = T X and T^T X on compressed rows / columns: FP32 entries widened
    exactly, FP64 products and sums, so |Y - Y64| <= 1e-13 (|T||X|)."""
    sd, s = _city(irtt, irtt.LAYOUT_SPARSE)
    with s:
        T = s.transport().astype(np.float64)
        rng = np.random.default_rng(R)
        X = rng.normal(size=(R, T.shape[0]))
        Y = s.apply_operator(X)
        Yt = s.apply_operator(X, transposed=True)
    A = np.abs(T)
    for got, want, bound in ((Y, X @ T.T, np.abs(X) @ A.T), (Yt, X @ T, np.abs(X) @ A)):
        ratio = (np.abs(got - want) / bound).max()
        print(f"R={R}: max |err| / (|T||x|) = {ratio:.2e}")
        assert ratio < 1e-13


def test_sparse_steady_and_adjoint_solves_vs_fp64(irtt):
    """Steady state without conduction is linear in y = x^4, so y = T^-1 b in
    FP64 is exact; the adjoint of the exact Jacobian T diag(4x^3) is solved
    densely in FP64.  Tolerances: 1e-7 K and 1e-8 relative."""
    sd, s = _city(irtt, irtt.LAYOUT_SPARSE)
    src = physics.bins(sd, 0)[3] * 1.3
    with s:
        T = s.transport().astype(np.float64)
        cs = s.column_sums()
        s.set_source(src)
        p = irtt.solver_params(ambient_kelvin=TA, x_tol=1e-9, f_tol=1e-18, lin_rtol=1e-12)
        s.set_temperatures(np.zeros(T.shape[0]))
        st = s.solve(params=p)
        x = s.temperatures()
        rng = np.random.default_rng(3)
        rhs = rng.normal(size=T.shape[0])
        lam = s.adjoint_solve(rhs, x=x, params=p)[0]
    n = T.shape[0]
    x64 = solver.steady_linear(T, solver.rhs(src, cs, TA), np.zeros(n, bool), np.zeros(n))
    J = T * (4.0 * x ** 3)[None, :]
    l64 = np.linalg.solve(J.T, rhs)
    dT = np.abs(x - x64).max()
    rel = np.abs(lam - l64).max() / np.abs(l64).max()
    print(f"steady |dT| vs FP64 {dT:.2e} K, adjoint relative {rel:.2e}; {st}")
    assert dT < 1e-7 and rel < 1e-8


def test_sparse_batch_solves_vs_fp64(irtt):
    sd, s = _city(irtt, irtt.LAYOUT_SPARSE)
    R = 33
    src0 = physics.bins(sd, 0)[3]
    rng = np.random.default_rng(5)
    src = src0[None, :] * (0.2 + 1.6 * rng.random((R, 1))) + 20.0 * rng.random((R, src0.size))
    with s:
        T = s.transport().astype(np.float64)
        cs = s.column_sums()
        p = irtt.solver_params(ambient_kelvin=TA, x_tol=1e-9, f_tol=1e-18, lin_rtol=1e-12)
        X, st = s.solve_batch(src, params=p)
        x = 270.0 + 60.0 * rng.random((R, T.shape[0]))
        rhs = rng.normal(size=(R, T.shape[0]))
        lam, _ = s.adjoint_solve_batch(x, rhs, lin_rtol=1e-12)
    n = T.shape[0]
    err = max(np.abs(X[k] - solver.steady_linear(T, solver.rhs(src[k], cs, TA),
                                                 np.zeros(n, bool), np.zeros(n))).max()
              for k in range(R))
    rel = 0.0
    for k in range(R):
        l64 = np.linalg.solve((T * (4.0 * x[k] ** 3)[None, :]).T, rhs[k])
        rel = max(rel, np.abs(lam[k] - l64).max() / np.abs(l64).max())
    print(f"batch steady max |dT| {err:.2e} K, adjoint relative {rel:.2e}; {st}")
    assert st["converged"] and err < 1e-6 and rel < 1e-8


def test_sparse_layout_offers_only_its_own_form(irtt):
    sd, s = _city(irtt, irtt.LAYOUT_SPARSE)
    with s:
        for op in (irtt.OP_DENSE, irtt.OP_DENSE_F16, irtt.OP_ACA):
            with pytest.raises(irtt.IrttError):
                s.select_operator(op)
        with pytest.raises(irtt.IrttError):
            s.compress()
        s.select_operator(irtt.OP_SPARSE)
    sd, s = _city(irtt, irtt.LAYOUT_DENSE)
    with s:
        with pytest.raises(irtt.IrttError):
            s.select_operator(irtt.OP_SPARSE)


def test_sparse_build_is_reproducible_at_full_city_size(irtt):
    """The 51.8k-triangle city block at 64 photons per triangle, built twice
    (hash-table insertion order differs run to run): identical compressed
    rows; column sums of the stored FP32 entries agree with the exact
    integer column sums to FP32 rounding."""
    from irtt import scenes
    sd = scenes.city_block(seed=0)
    got = []
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        for _ in range(2):
            st = s.build_transport(threads_per_tri=16, rays_per_thread=4, batch_seed=1234,
                                   layout=irtt.LAYOUT_SPARSE)
            got.append(s.transport_sparse())
        cs = s.column_sums()
        n = s.n_bins
    rp, col, val = got[0]
    for a, b in zip(got[0], got[1]):
        assert np.array_equal(a, b)
    assert np.all(np.diff(rp.astype(np.int64)) >= 1)
    rows = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
    key = rows.astype(np.int64) * n + col
    assert np.all(np.diff(key) > 0)
    sums = np.bincount(col, weights=val.astype(np.float64), minlength=n)
    scale = np.bincount(col, weights=np.abs(val.astype(np.float64)), minlength=n)
    print(f"city block: {n} bins, nnz {col.size} ({100 * col.size / n / n:.2f} %), "
          f"trace {st['trace_ms']:.1f} ms, rows+columns {st['finalize_ms']:.1f} ms")
    assert np.all(np.abs(sums - cs) <= 1e-6 * scale + 1e-300)
