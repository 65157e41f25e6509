"""Pins the oracle adjoint trace (thermal_tracing_unified.glsl:292-599):
the area and geometric-term derivatives against central finite differences of
their defining formulas, the cancellation of the area terms in a closed
enclosure, and the whole vertex gradient against a central finite difference
of the transport functional with common random numbers."""
import numpy as np
import pytest

SIG = 5.670374419e-8


def _area(a, b, c):
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a))


def test_area_gradient_fd(oracle):
    rng = np.random.default_rng(0)
    for _ in range(50):
        v = rng.normal(size=(3, 3))
        g = oracle.area_gradient(*v.astype(np.float32))
        v64 = v.astype(np.float32).astype(np.float64)
        for k in range(3):
            for c in range(3):
                e = np.zeros((3, 3))
                e[k, c] = 1e-6
                fd = (_area(*(v64 + e)) - _area(*(v64 - e))) / 2e-6
                assert abs(fd - g[k][c]) < 1e-4 * (1 + abs(fd))


def _lnG(v, bi, bj, eps):
    ai, bi_, ci, aj, bj_, cj = v
    pi = ai * (1 - bi[0] - bi[1]) + bi_ * bi[0] + ci * bi[1]
    pj = aj * (1 - bj[0] - bj[1]) + bj_ * bj[0] + cj * bj[1]
    d = pj - pi
    Ni = np.cross(bi_ - ai, ci - ai)
    Nj = np.cross(bj_ - aj, cj - aj)
    r = np.linalg.norm(d)
    G = -(Ni @ d) * (Nj @ d) / (np.linalg.norm(Ni) * np.linalg.norm(Nj) * (r + eps) ** 4)
    return np.log(abs(G)), G


def test_geom_term_gradient_fd(oracle):
    """d ln|G| / dv of the segment between two facing triangles, all six
    vertices, against central differences in FP64."""
    rng = np.random.default_rng(1)
    checked = 0
    for _ in range(60):
        tri_i = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float) + rng.normal(size=(3, 3)) * 0.2
        tri_j = np.array([[0, 0, 2], [0, 1, 2], [1, 0, 2]], float) + rng.normal(size=(3, 3)) * 0.2
        tri_j += rng.normal(size=3) * 0.5
        v = np.concatenate([tri_i, tri_j]).astype(np.float32)
        bi = rng.dirichlet([1, 1, 1])[1:]
        bj = rng.dirichlet([1, 1, 1])[1:]
        G, g = oracle.geom_term_grad(v, bi.astype(np.float32), bj.astype(np.float32), 1e-2)
        v64 = v.astype(np.float64)
        bi64 = bi.astype(np.float32).astype(np.float64)
        bj64 = bj.astype(np.float32).astype(np.float64)
        lg, G64 = _lnG(v64, bi64, bj64, 1e-2)
        assert abs(G - G64) <= 1e-4 * abs(G64)
        for k in range(6):
            for c in range(3):
                e = np.zeros((6, 3))
                e[k, c] = 1e-6
                fd = (_lnG(v64 + e, bi64, bj64, 1e-2)[0] - _lnG(v64 - e, bi64, bj64, 1e-2)[0]) / 2e-6
                assert abs(fd - g[k, c]) < 2e-3 * (1 + abs(fd)), (k, c, fd, g[k, c])
        checked += 1
    assert checked == 60


def test_area_terms_cancel_in_closed_box(oracle):
    """lambda = x^4 = 1: emission and absorption area terms cancel photon by
    photon except for dropped photons, so the gradient is tiny."""
    from irtt import scenes
    sd = scenes.closed_box(2, emissivity=0.7)
    sc = oracle.OracleScene.from_irtt(sd)
    p = oracle.make_params(64, 4, batch_seed=3, bin_mode=0)
    ones = np.ones(sd.nv)
    g, st = oracle.adjoint(sc, p, ones, ones, flags=oracle.ADJ_AREA)
    g_emit_only, _ = oracle.adjoint(sc, p, ones, ones, flags=oracle.ADJ_AREA)
    scale = SIG * (1.0 / 8) / (64 * 4)  # one photon weight
    dropped = st["photons"] - st["absorbed"]
    assert np.abs(g).max() <= (dropped + 1) * scale * 10
    assert st["absorbed"] > 0.99 * st["photons"]


def _phi(oracle, sd, params, lam, x4):
    """phi = lam^T T x4 with T = diag(sigma eps A) - acc * Q (vertex bins)."""
    from oracle import physics
    sc = oracle.OracleScene.from_irtt(sd)
    acc, _, _ = oracle.transport(sc, params)
    _, _, Q = oracle.photon_weights(sc, params)
    _, ea, _, _ = physics.bins(sd, 0)
    T = np.diag(SIG * ea) - acc.astype(np.float64) * Q
    return lam @ T @ x4


def test_vertex_gradient_matches_fd_plates(oracle):
    """Moving plate B away from plate A: d(lam^T T x4)/dz by central FD (common
    random numbers) vs the adjoint trace's vertex gradient summed over plate B
    (the trace returns -d phi / d vertex)."""
    from irtt import scenes
    sd = scenes.parallel_plates(n=4, emissivity=1.0)
    nva = sd.meta["plate_a_verts"]
    params = oracle.make_params(4096, 8, batch_seed=11, bin_mode=0)
    rng = np.random.default_rng(0)
    lam = np.zeros(sd.nv)
    lam[nva:] = 1.0 + 0.2 * rng.random(sd.nv - nva)   # reward absorption on B
    x4 = np.zeros(sd.nv)
    x4[:nva] = 1.0 + 0.2 * rng.random(nva)            # emission from A only
    h = 2e-3
    sp, sm = [scenes.parallel_plates(n=4, emissivity=1.0) for _ in range(2)]
    sp.xyz[nva:, 2] += h
    sm.xyz[nva:, 2] -= h
    fd = (_phi(oracle, sp, params, lam, x4) - _phi(oracle, sm, params, lam, x4)) / (2 * h)
    sc = oracle.OracleScene.from_irtt(sd)
    g, st = oracle.adjoint(sc, params, lam, x4)
    adj = -g[nva:, 2].sum()
    print(f"d phi/dz: FD {fd:.6e}, adjoint {adj:.6e}, ratio {adj / fd:.4f}; {st}")
    # off-diagonal T is minus the absorbed power: less reaches B as it moves away
    assert fd > 0
    assert abs(adj - fd) < 0.1 * abs(fd)
