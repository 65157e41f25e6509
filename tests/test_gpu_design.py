"""Design gradients on the procedural facade (irtt_source_gradient,
irtt_design_gradient, irtt/design.py).

* The source term lambda^T db/dv (b = flux x area per triangle, a third to
  each corner bin) against central differences of lambda^T b(x) computed in
  FP64 numpy: the area derivative is a closed form, so the match is to
  rounding (1e-6 relative).
* The adjoint chain through the source: with the operator held fixed and the
  per-bin source b(depths) recomputed for each depth (no Monte Carlo change
  between evaluations), dL/d(depth) by central differences of full Newton
  solves equals the projected source part of irtt_design_gradient
  (1e-4 relative: solve tolerance and FD truncation).
* irtt_design_gradient = projection of (adjoint trace with the rows
  (lambda, x^4), (-Ta^4, lambda)) + source term, each computed separately.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TA = 293.0


def _facade():
    from irtt import scenes
    return scenes.facade(nx=2, ny=1, hmin=10.0, hmax=12.0, target_tris=20000,
                         n_fin_buildings=2, seed=3)


def _lam_b(sd, X, lam):
    I = sd.idx.astype(np.int64)
    a = 0.5 * np.linalg.norm(np.cross(X[I[:, 1]] - X[I[:, 0]], X[I[:, 2]] - X[I[:, 0]]), axis=1)
    w = sd.heat_flux.astype(np.float64) * a / 3.0
    return w @ (lam[I[:, 0]] + lam[I[:, 1]] + lam[I[:, 2]])


def test_source_gradient_vs_fd(irtt):
    sd = _facade()
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        rng = np.random.default_rng(0)
        lam = rng.normal(size=s.n_bins)
        g = s.source_gradient(lam)
        s.set_source(np.zeros(s.n_bins))
        assert not s.source_gradient(lam).any()  # an explicit per-bin power is fixed
    X = sd.xyz.astype(np.float64)
    h = 1e-4
    worst = 0.0
    for v in rng.choice(sd.nv, 40, replace=False):
        for k in range(3):
            Xp, Xm = X.copy(), X.copy()
            Xp[v, k] += h
            Xm[v, k] -= h
            fd = (_lam_b(sd, Xp, lam) - _lam_b(sd, Xm, lam)) / (2 * h)
            worst = max(worst, abs(fd - g[v, k]) / max(abs(fd), 1e-9 * np.abs(g).max()))
    print(f"source gradient: max relative error vs FD {worst:.2e}")
    assert worst < 1e-6


def test_design_gradient_source_chain_vs_fd(irtt):
    """Operator fixed, source b(depths): dL/dp from central differences of
    converged solves vs the source part of the design gradient."""
    from irtt import scenes
    from irtt.design import ground_weights
    sd = _facade()
    par = scenes.facade_parametrization(sd)
    d0 = np.asarray(sd.meta["depths"], np.float64)
    sp = irtt.solver_params(ambient_kelvin=TA, x_tol=1e-10, f_tol=1e-20, lin_rtol=1e-13)
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=32, rays_per_thread=8, batch_seed=5,
                          layout=irtt.LAYOUT_SPARSE)
        w = ground_weights(s, sd)
        s.set_vertex_state(sd.kelvin, sd.fixed)
        s.solve(sp)
        dp, gx, lam, st = s.design_gradient(w, params=par, solver=sp)
        src_part = irtt.project_gradient(s.source_gradient(lam), *par)

        m = sd.meta
        I = sd.idx.astype(np.int64)
        flux = sd.heat_flux.astype(np.float64)

        def L(d):
            # per-bin source of the geometry at depths d (FP64 positions),
            # operator unchanged
            X = m["base_xyz"].copy()
            X[m["param_vertex"]] += m["param_dir"].astype(np.float64) * d[m["param_index"]][:, None]
            a = 0.5 * np.linalg.norm(np.cross(X[I[:, 1]] - X[I[:, 0]], X[I[:, 2]] - X[I[:, 0]]),
                                     axis=1)
            b = np.zeros(sd.nv)
            for c in range(3):
                np.add.at(b, I[:, c], flux * a / 3.0)
            s.set_source(b)
            s.set_vertex_state(sd.kelvin, sd.fixed)
            s.solve(sp)
            return float(w @ s.temperatures())

        rel = []
        for k in (0, 5, 13, 24, 25):
            h = 1e-3
            dk = d0.copy()
            dk[k] += h
            lp = L(dk)
            dk[k] -= 2 * h
            lm = L(dk)
            fd = (lp - lm) / (2 * h)
            rel.append(abs(fd - src_part[k]) / abs(fd))
            print(f"p{k}: source part {src_part[k]:+.6e}, FD {fd:+.6e}")
    print(f"max relative difference {max(rel):.2e}")
    assert max(rel) < 1e-4


def test_design_gradient_is_trace_plus_source(irtt):
    from irtt import scenes
    from irtt.design import ground_weights
    sd = _facade()
    par = scenes.facade_parametrization(sd)
    sp = irtt.solver_params(ambient_kelvin=TA, x_tol=1e-8, f_tol=1e-16)
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=16, rays_per_thread=8, batch_seed=5,
                          layout=irtt.LAYOUT_SPARSE)
        w = ground_weights(s, sd)
        s.set_vertex_state(sd.kelvin, sd.fixed)
        s.solve(sp)
        x = s.temperatures()
        dp, gx, lam, st = s.design_gradient(w, params=par, solver=sp)
        lam2, _ = s.adjoint_solve(w, params=sp)
        gt, _ = s.adjoint_trace(np.stack([lam2, np.full(len(x), -TA ** 4)]),
                                np.stack([x ** 4, lam2]))
        g2 = gt + s.source_gradient(lam2)
    assert np.array_equal(lam, lam2)
    # the trace is bit-reproducible; the library adds the source term in its
    # own order: compare to rounding
    scale = np.abs(g2).max()
    assert np.abs(gx - g2).max() <= 1e-9 * scale
    dp2 = irtt.project_gradient(g2, *par)
    assert np.allclose(dp, dp2, rtol=1e-8, atol=1e-12 * np.abs(dp2).max())
