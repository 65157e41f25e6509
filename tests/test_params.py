"""Design parametrizations, constraints, objective and optimisers
(irtt/params.py, irtt/design.py host logic; no device work).

* Every vertex map's Jacobian against central differences of its positions
  (FP64, h = 1e-6: truncation O(h^2), so 1e-7 relative).
* The rotation map against closed forms: a rotation by pi/2 about z maps
  x -> y, R is orthonormal with det 1, R(r) R(-r) = I, and the small-angle
  branch agrees with the full formula across the switch.
* Parametrization.triplets pulls a world gradient back through a model
  matrix: sum_k dxdp_k . g equals d(g . x(p))/dp by central differences.
* InPolygonXZ and CollisionXZ penalties: gradient vs central differences of
  the penalty, zero inside / apart, and the reference's polygon string.
* Adam: bias-corrected first step is -step_size sign(g) (eps aside); on a
  separable quadratic it converges; GD and Adam keep the best state.
* Objective: value and gradient of 1/2 sum w (x - t)^2, "solve" targets.
"""
import numpy as np
import pytest

from irtt import params as P
from irtt.design import LinearParametrization


def _fd_jacobian(fmap, x0, p, h=1e-6):
    J = np.zeros((x0.shape[0], fmap.n_params, 3))
    for k in range(fmap.n_params):
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        J[:, k, :] = (fmap.positions(x0, pp) - fmap.positions(x0, pm)) / (2 * h)
    return J


MAPS = {
    "translate": (P.ParametricMap(), [0.3, -0.2, 0.5]),
    "axis": (P.AxisTranslate((0.0, 0.6, 0.8)), [0.7]),
    "skew": (P.AxisSkew(), [0.4]),
    "rotate": (P.Rotate(), [0.3, -0.5, 0.9]),
    "rotate_small": (P.Rotate(), [1e-3, -2e-3, 5e-4]),
    "translate_rotate": (P.TranslateRotate(), [0.2, 0.1, -0.4, 1.0, 2.0, -3.0]),
    "translate_rotate_xz": (P.TranslateRotateXZ(0.5), [0.3, 0.8, -0.2]),
    "scale": (P.UniformScale(), [0.25]),
    "nop": (P.Nop(), [0.1]),
    "bend_inside": (P.Bend(), [0.1, 0.2, -0.1, 0.5, 0.0, 1.0, 0.2, 0.3]),
    "bend_clamped": (P.Bend(), [0.1, 0.2, -0.1, 0.5, 0.0, 1.0, 0.6, 0.3]),
    "sine_modes": (P.SineModes(), [0.2, -0.3, 0.1, 0.05]),
}


@pytest.mark.parametrize("name", sorted(MAPS))
def test_jacobian_vs_fd(name):
    fmap, p = MAPS[name]
    rng = np.random.default_rng(1)
    x0 = rng.uniform(-3, 3, (50, 3))
    p = np.asarray(p, np.float64)
    J = fmap.jacobian(x0, p)
    assert J.shape == (50, fmap.n_params, 3)
    Jfd = _fd_jacobian(fmap, x0, p)
    scale = max(np.abs(Jfd).max(), 1.0)
    assert np.abs(J - Jfd).max() < 1e-7 * scale, name


def test_rotation_closed_forms():
    R = P.rotation_matrix([0.0, 0.0, np.pi / 2])
    assert np.allclose(R @ [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], atol=1e-15)
    rng = np.random.default_rng(2)
    for _ in range(10):
        r = rng.normal(size=3)
        R = P.rotation_matrix(r)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1.0) < 1e-14
        assert np.allclose(R @ P.rotation_matrix(-r), np.eye(3), atol=1e-14)
        assert np.allclose(R @ r, r, atol=1e-14)  # the axis is fixed
    # both branches at the switch |r|^2 = 1e-12 agree to first order
    r = np.array([1e-6, -0.5e-6, 0.3e-6])
    x0 = rng.normal(size=(5, 3))
    Jsmall = P.rotation_jacobian(r * 0.999, x0)
    Jfull = P.rotation_jacobian(r * 1.001, x0)
    assert np.abs(Jsmall - Jfull).max() < 1e-5


def test_translate_rotate_xz_matches_y_rotation():
    m = P.TranslateRotateXZ(2.0)
    x0 = np.array([[1.0, 2.0, 0.0]])
    # rotation_scale * p1 = pi/2 about y maps +x to -z
    x = m.positions(x0, [0.0, np.pi / 4, 0.0])
    assert np.allclose(x, [[0.0, 2.0, -1.0]], atol=1e-15)
    assert np.allclose(m.positions(x0, [1.0, 0.0, -2.0]), [[2.0, 2.0, -2.0]])


def test_make_map_modes():
    assert P.make_map(P.NOT_MOVING) is None
    assert isinstance(P.make_map(P.TRANSLATE_ONLY), P.ParametricMap)
    assert P.make_map(P.TRANSLATE_ROTATE).n_params == 6
    assert P.make_map(P.TRANSLATE_ROTATE_XZ, 0.1).s == 0.1
    assert P.make_map(P.CUSTOM97).n_params == 8
    assert P.make_map(P.CUSTOM99).n_params == 4
    with pytest.raises(ValueError):
        P.make_map(42)


def test_parametrization_triplets_pull_back():
    rng = np.random.default_rng(3)
    rest = rng.uniform(-2, 2, (40, 3))
    M = np.eye(4)
    M[:3, :3] = P.rotation_matrix([0.2, -0.4, 0.1]) * 1.5
    M[:3, 3] = [1.0, -2.0, 0.5]
    par = P.Parametrization(rest, [(np.arange(0, 15), P.TranslateRotate(), M),
                                   (np.arange(20, 30), P.UniformScale(), None),
                                   (np.arange(30, 40), P.Bend(), None)])
    assert par.n_params == 15
    assert np.allclose(par.positions(np.zeros(15)), rest)  # p = 0 is the rest state
    p = rng.normal(scale=0.2, size=15)
    g = rng.normal(size=(40, 3))
    v, k, d, n = par.triplets(p)
    dp = np.zeros(n)
    np.add.at(dp, k.astype(np.int64), (d.astype(np.float64) * g[v.astype(np.int64)]).sum(1))
    h = 1e-6
    fd = np.array([(np.sum(g * par.positions(p + h * e)) - np.sum(g * par.positions(p - h * e)))
                   / (2 * h) for e in np.eye(15)])
    assert np.allclose(dp, fd, rtol=1e-5, atol=1e-5 * np.abs(fd).max())
    x = par.positions(p)
    assert np.array_equal(x[15:20], rest[15:20])  # unparametrised vertices stay
    with pytest.raises(ValueError):
        P.Parametrization(rest, [(np.arange(0, 10), P.ParametricMap(), None),
                                 (np.arange(5, 12), P.Nop(), None)])
    with pytest.raises(ValueError):
        par.positions(np.zeros(3))


def test_linear_parametrization():
    rest = np.zeros((6, 3))
    lp = LinearParametrization(rest, [1, 2, 2, 5], [0, 0, 1, 1],
                               [[0, 0, 1], [1, 0, 0], [0, 1, 0], [0, 0, 2]], 2)
    x = lp.positions([0.5, 2.0])
    assert np.allclose(x[1], [0, 0, 0.5]) and np.allclose(x[2], [0.5, 2.0, 0])
    assert np.allclose(x[5], [0, 0, 4.0]) and not x[[0, 3, 4]].any()


def _fd_penalty(con, xyz, h=1e-6):
    g = np.zeros_like(xyz)
    for i in range(xyz.shape[0]):
        for c in range(3):
            xp, xm = xyz.copy(), xyz.copy()
            xp[i, c] += h
            xm[i, c] -= h
            g[i, c] = (con.evaluate(xp, np.zeros_like(xyz)) -
                       con.evaluate(xm, np.zeros_like(xyz))) / (2 * h)
    return g


def test_in_polygon_xz():
    poly = "0,0; 4,0; 4,3; 2,5; 0,3; 0,0"   # closed, last = first
    c = P.InPolygonXZ(np.arange(6), 10.0, poly)
    assert c.poly.shape == (5, 2)
    xyz = np.array([[1.0, 7.0, 1.0],     # inside (y is ignored)
                    [5.0, 0.0, 1.0],     # right of the polygon: d = 1
                    [-1.0, 0.0, -1.0],   # outside near a corner
                    [3.5, 0.0, 4.5],     # outside the roof edge
                    [2.0, 0.0, 4.0],     # inside under the roof
                    [6.0, 1.0, 6.0]])
    g = np.zeros_like(xyz)
    val = c.evaluate(xyz, g)
    assert not g[[0, 4]].any()
    assert np.allclose(g[1], [10.0, 0.0, 0.0])
    assert np.allclose(g[2], [-10.0, 0.0, -10.0])
    assert not g[:, 1].any()
    fd = _fd_penalty(c, xyz)
    assert np.allclose(g, fd, atol=1e-4)
    assert val > 0.5 * 10.0 * (1.0 + 2.0)  # at least the two axis-aligned cases


def test_collision_xz():
    # two unit squares (2 triangles each) in the XZ plane, 1.2 apart: radii
    # sqrt(2)/2 each, so they overlap by 1.2 - sqrt(2) < 0
    sq = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], np.float64)
    xyz = np.concatenate([sq, sq + [1.2, 0.3, 0.0], sq + [10.0, 0.0, 10.0]])
    idx = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7], [8, 9, 10], [8, 10, 11]])
    c = P.CollisionXZ(xyz, idx, [np.arange(4), np.arange(4, 8), np.arange(8, 12)], 5.0)
    assert np.allclose([m[2] for m in c.models], np.sqrt(2) / 2)
    g = np.zeros_like(xyz)
    val = c.evaluate(xyz, g)
    o = 1.2 - np.sqrt(2)
    assert abs(val - 0.5 * 5.0 * o * o) < 1e-12
    assert not g[8:].any()                       # the far model is untouched
    assert np.allclose(g[:4].sum(0), [-5.0 * o, 0, 0])   # pushes the pair apart
    assert np.allclose(g[4:8].sum(0), [5.0 * o, 0, 0])
    assert np.allclose(g, _fd_penalty(c, xyz), atol=1e-5)
    cons = P.Constraints([c, P.InPolygonXZ(np.arange(12), 1.0, [[-1, -1], [20, -1], [20, 20]])])
    g2 = np.zeros_like(xyz)
    assert cons.evaluate(xyz, g2) >= val


def test_objective():
    w = np.array([1.0, 2.0, 0.0, 1.0])
    a = np.array([0.5, 1.0, 3.0, 2.0])
    obj = P.Objective(w, a, [300.0, np.nan, 290.0, 310.0])
    x = np.array([301.0, 305.0, 280.0, 300.0])
    assert obj.missing == 1
    assert obj.eval(x) == (0.0, pytest.approx(np.zeros(4)))
    obj.update_targets(x)
    assert obj.missing == 0
    L, g = obj.eval(x + 1.0)
    wa = w * a
    d = np.array([2.0, 1.0, 1.0, -9.0])
    assert L == pytest.approx(0.5 * np.sum(wa * d * d))
    assert np.allclose(g, wa * d)
    assert obj.eval(x, step=3)[0] == 0.0


def test_adam_and_best_state():
    opt = P.Adam(step_size=0.1)
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([4.0, -0.5, 0.0])
    q = opt.step(p, g, 10.0)
    # first step: m = (1-b1) g, v = (1-b2) g^2, corrected -> -a g / |g|
    assert np.allclose(q, p - 0.1 * np.array([1.0, -1.0, 0.0]), atol=1e-7)
    # separable quadratic 1/2 sum c (p - t)^2
    c = np.array([1.0, 10.0, 0.1])
    t = np.array([0.5, -1.0, 2.0])
    opt = P.Adam(step_size=0.05)
    p = np.zeros(3)
    for _ in range(2000):
        p = opt.step(p, c * (p - t), 0.5 * np.sum(c * (p - t) ** 2))
    assert np.abs(p - t).max() < 1e-2
    assert opt.best_value <= 0.5 * np.sum(c * (p - t) ** 2) + 1e-3
    gd = P.GradientDescent(0.5)
    p = np.array([4.0])
    vals = []
    for _ in range(5):
        v = float((p[0] - 1.0) ** 2)
        vals.append(v)
        p = gd.step(p, 2.0 * (p - 1.0) * 1.5, v)   # overshooting steps
    assert gd.best_value == min(vals)
    gd.reset(0.1)
    assert gd.step_size == 0.1


def test_parse_polygon():
    pts = P.parse_polygon("1.5,2; 3,-4 ;bad; 5,6,7")
    assert np.allclose(pts, [[1.5, 2.0], [3.0, -4.0], [5.0, 6.0]])
