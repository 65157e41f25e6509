"""Design parametrizations, penalty constraints, objectives and optimisers
around the device gradient (irtt_design_gradient / irtt_project_gradient).

Host-side logic with no device work: every function here acts on at most
V x 3 vertex arrays and P parameters.  The reference keeps the same pieces in
src/implementations/thermal_renderer/thermal_params.{hpp,cpp} (vertex maps
with Jacobians, :41-130 / :160-480; penalty constraints, :135-180 /
:480-724) and thermal_objective.{hpp,cpp} (weighted squared deviation,
thermal_objective.cpp:1-78; gradient descent and Adam with the best state
seen, thermal_objective.hpp:19-78).

Conventions.  A vertex map x = f(x0, p) acts on the rest positions x0 of one
group of vertices (model space); ``jacobian`` returns d x / d p as an array of
shape (n, n_params, 3) -- one row per parameter, as the reference's
getJacobian fills a (n_params x 3) matrix.  A group may carry a 4x4 model
matrix M (world = M [x; 1]); gradients with respect to world positions are
pulled back with the transpose of M's linear part
(thermal_params.cpp:142-143), which ``Parametrization.triplets`` folds into
the d(world x)/dp rows handed to irtt_project_gradient.
"""
import numpy as np

# optimize-mode numbers of the reference (thermal_params.hpp:17-29)
NOT_MOVING = 0
TRANSLATE_ONLY = 1
TRANSLATE_ROTATE = 2
ROTATE_ONLY = 3
TRANSLATE_ROTATE_XZ = 4
UNIFORM_SCALE = 5
TRANSLATE_AXIS = 6
SKEW_AXIS = 7
NOP_TEST1 = 96
CUSTOM97 = 97
CUSTOM99 = 99


def _cross_matrix(v):
    """[v]x for an array of vectors (..., 3) -> (..., 3, 3)."""
    v = np.asarray(v, np.float64)
    z = np.zeros(v.shape[:-1])
    return np.stack([np.stack([z, -v[..., 2], v[..., 1]], -1),
                     np.stack([v[..., 2], z, -v[..., 0]], -1),
                     np.stack([-v[..., 1], v[..., 0], z], -1)], -2)


def rotation_matrix(r):
    """exp([r]x): rotation by |r| about r / |r| (Rodrigues' formula).  Below
    |r|^2 = 1e-12 the first-order form I + [r]x, as thermal_params.cpp:160-200
    switches."""
    r = np.asarray(r, np.float64)
    t2 = r @ r
    K = _cross_matrix(r)
    if t2 <= 1e-12:
        return np.eye(3) + K
    t = np.sqrt(t2)
    return np.eye(3) + (np.sin(t) / t) * K + ((1.0 - np.cos(t)) / t2) * (K @ K)


def rotation_jacobian(r, x0):
    """d(exp([r]x) x0) / dr for rest positions x0 (n, 3) -> (n, 3, 3), row i
    the derivative with respect to r_i.  Compact form of Gallego and Yezzi
    (J. Math. Imaging Vis. 51, 2015, eq. 9):
        dR/dr_i = (r_i [r]x + [r x ((I - R) e_i)]x) R / |r|^2,
    and at |r|^2 <= 1e-12 the derivative of I + [r]x, e_i x x0."""
    r = np.asarray(r, np.float64)
    x0 = np.asarray(x0, np.float64).reshape(-1, 3)
    E = np.eye(3)
    t2 = r @ r
    if t2 <= 1e-12:
        return np.stack([np.cross(E[i], x0) for i in range(3)], 1)
    R = rotation_matrix(r)
    K = _cross_matrix(r)
    Rx = x0 @ R.T
    out = np.empty((x0.shape[0], 3, 3))
    for i in range(3):
        Mi = (r[i] * K + _cross_matrix(np.cross(r, (E - R)[:, i]))) / t2
        out[:, i, :] = Rx @ Mi.T
    return out


class ParametricMap:
    """Translation x = x0 + p[0:3] (the reference's ParametricFunction base,
    thermal_params.hpp:41-46)."""
    n_params = 3

    def positions(self, x0, p):
        return np.asarray(x0, np.float64) + np.asarray(p, np.float64)[:3]

    def jacobian(self, x0, p):
        n = np.asarray(x0).reshape(-1, 3).shape[0]
        return np.broadcast_to(np.eye(3), (n, 3, 3)).copy()


Translate = ParametricMap


class AxisTranslate(ParametricMap):
    """x = x0 + axis p[0] (thermal_params.hpp:48-57; axis (0, 0, 1) there)."""
    n_params = 1

    def __init__(self, axis=(0.0, 0.0, 1.0)):
        self.axis = np.asarray(axis, np.float64)

    def positions(self, x0, p):
        return np.asarray(x0, np.float64) + self.axis * float(p[0])

    def jacobian(self, x0, p):
        n = np.asarray(x0).reshape(-1, 3).shape[0]
        return np.broadcast_to(self.axis, (n, 1, 3)).copy()


class AxisSkew(ParametricMap):
    """x = x0 + translation (projection . x0) p[0]: a shear that moves each
    vertex along `translation` in proportion to its coordinate along
    `projection` (thermal_params.hpp:59-69; (0,1,0) and (1,0,0) there)."""
    n_params = 1

    def __init__(self, projection=(0.0, 1.0, 0.0), translation=(1.0, 0.0, 0.0)):
        self.proj = np.asarray(projection, np.float64)
        self.trans = np.asarray(translation, np.float64)

    def positions(self, x0, p):
        x0 = np.asarray(x0, np.float64)
        return x0 + np.outer(x0 @ self.proj * float(p[0]), self.trans)

    def jacobian(self, x0, p):
        x0 = np.asarray(x0, np.float64).reshape(-1, 3)
        return (np.outer(x0 @ self.proj, self.trans))[:, None, :]


class Rotate(ParametricMap):
    """x = exp([p]x) x0, rotation about the origin (thermal_params.hpp:71-75)."""
    n_params = 3

    def positions(self, x0, p):
        return np.asarray(x0, np.float64) @ rotation_matrix(p[:3]).T

    def jacobian(self, x0, p):
        return rotation_jacobian(p[:3], x0)


class TranslateRotate(ParametricMap):
    """x = exp([p0:3]x) x0 + p3:6 (thermal_params.hpp:77-82)."""
    n_params = 6

    def positions(self, x0, p):
        p = np.asarray(p, np.float64)
        return np.asarray(x0, np.float64) @ rotation_matrix(p[:3]).T + p[3:6]

    def jacobian(self, x0, p):
        x0 = np.asarray(x0, np.float64).reshape(-1, 3)
        J = np.zeros((x0.shape[0], 6, 3))
        J[:, :3] = rotation_jacobian(p[:3], x0)
        J[:, 3:] = np.eye(3)
        return J


class TranslateRotateXZ(ParametricMap):
    """Rotation about y by rotation_scale * p[1] followed by a translation by
    (p[0], 0, p[2]) (thermal_params.hpp:84-92, thermal_params.cpp:330-357)."""
    n_params = 3

    def __init__(self, rotation_scale=1.0):
        self.s = float(rotation_scale)

    def _rot(self, a):
        c, s = np.cos(a), np.sin(a)
        return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])

    def positions(self, x0, p):
        x = np.asarray(x0, np.float64) @ self._rot(self.s * float(p[1])).T
        x[..., 0] += p[0]
        x[..., 2] += p[2]
        return x

    def jacobian(self, x0, p):
        x0 = np.asarray(x0, np.float64).reshape(-1, 3)
        a = self.s * float(p[1])
        c, s = np.cos(a), np.sin(a)
        dR = self.s * np.array([[-s, 0.0, c], [0.0, 0.0, 0.0], [-c, 0.0, -s]])
        J = np.zeros((x0.shape[0], 3, 3))
        J[:, 0, 0] = 1.0
        J[:, 1, :] = x0 @ dR.T
        J[:, 2, 2] = 1.0
        return J


class UniformScale(ParametricMap):
    """x = (1 + p[0]) x0, scaling about the origin (thermal_params.hpp:94-99)."""
    n_params = 1

    def positions(self, x0, p):
        return (1.0 + float(p[0])) * np.asarray(x0, np.float64)

    def jacobian(self, x0, p):
        return np.asarray(x0, np.float64).reshape(-1, 1, 3).copy()


class Nop(ParametricMap):
    """One parameter that moves nothing: its gradient must come out zero
    (the reference's NOP test, thermal_params.hpp:102-108)."""
    n_params = 1

    def positions(self, x0, p):
        return np.asarray(x0, np.float64).copy()

    def jacobian(self, x0, p):
        return np.zeros((np.asarray(x0).reshape(-1, 3).shape[0], 1, 3))


class Bend(ParametricMap):
    """Rigid motion plus a quadratic bend along x (the reference's mode 97,
    thermal_params.cpp:359-390): x = exp([p0:3]x) x0 + p3:6, then
    x_x += a/2 (x0_x - b)^2 with a = clamp(p6, -amax, amax), b = p7.

    The Jacobian is the derivative of this map: the a-row vanishes while p6
    lies outside [-amax, amax], and the b-row uses the clamped a.  (The
    reference's rows keep the unclamped a in the b-row and zero the a-row
    only below -amax.)"""
    n_params = 8

    def __init__(self, amax=0.3):
        self.amax = float(amax)

    def positions(self, x0, p):
        p = np.asarray(p, np.float64)
        x0 = np.asarray(x0, np.float64)
        x = x0 @ rotation_matrix(p[:3]).T + p[3:6]
        a = np.clip(p[6], -self.amax, self.amax)
        x[..., 0] += 0.5 * a * (x0[..., 0] - p[7]) ** 2
        return x

    def jacobian(self, x0, p):
        p = np.asarray(p, np.float64)
        x0 = np.asarray(x0, np.float64).reshape(-1, 3)
        J = np.zeros((x0.shape[0], 8, 3))
        J[:, :3] = rotation_jacobian(p[:3], x0)
        J[:, 3:6] = np.eye(3)
        inside = -self.amax <= p[6] <= self.amax
        a = np.clip(p[6], -self.amax, self.amax)
        if inside:
            J[:, 6, 0] = 0.5 * (x0[:, 0] - p[7]) ** 2
        J[:, 7, 0] = -a * (x0[:, 0] - p[7])
        return J


class SineModes(ParametricMap):
    """x_x = x0_x + sum_k f_k tanh(s_k p_k) m_k(x0): four saturating
    deformation modes along x, each a product of sines of the rest y and z
    coordinates (the reference's mode 99, thermal_params.cpp:392-430, whose
    mode shapes, amplitudes and steepnesses are the defaults here).  Mode k:
    amplitude f_k, steepness s_k and the sine factors (wy_k, wz_k): m_k =
    sin(wy_k y) sin(wz_k z) (a zero wavenumber multiplies by y instead)."""
    n_params = 4

    def __init__(self, amplitude=(0.0, 3.0, 2.0, 0.8), steepness=(0.5, 1.0, 2.0, 8.0),
                 wy=(0.0, 5 * np.pi / 204, 5 * np.pi / 102, 25 * np.pi / 204),
                 wz=(None, 5 * np.pi / 102, 5 * np.pi / 51, 5 * np.pi / 102)):
        self.f = np.asarray(amplitude, np.float64)
        self.s = np.asarray(steepness, np.float64)
        self.wy = wy
        self.wz = wz

    def _modes(self, x0):
        y, z = x0[:, 1], x0[:, 2]
        m = np.empty((x0.shape[0], 4))
        m[:, 0] = y  # the first mode is a shear in y
        for k in range(1, 4):
            m[:, k] = np.sin(self.wy[k] * y) * np.sin(self.wz[k] * z)
        return m

    def positions(self, x0, p):
        x0 = np.asarray(x0, np.float64).reshape(-1, 3)
        p = np.asarray(p, np.float64)[:4]
        x = x0.copy()
        x[:, 0] += self._modes(x0) @ (self.f * np.tanh(self.s * p))
        return x

    def jacobian(self, x0, p):
        x0 = np.asarray(x0, np.float64).reshape(-1, 3)
        p = np.asarray(p, np.float64)[:4]
        J = np.zeros((x0.shape[0], 4, 3))
        J[:, :, 0] = self._modes(x0) * (self.f * self.s * (1.0 - np.tanh(self.s * p) ** 2))
        return J


def make_map(mode, rotation_scale=1.0):
    """ParametricMap for one of the reference's optimize-mode numbers
    (thermal_params.cpp:8-76); None for NOT_MOVING."""
    table = {TRANSLATE_ONLY: ParametricMap, ROTATE_ONLY: Rotate,
             TRANSLATE_ROTATE: TranslateRotate, UNIFORM_SCALE: UniformScale,
             NOP_TEST1: Nop, CUSTOM97: Bend, CUSTOM99: SineModes,
             TRANSLATE_AXIS: AxisTranslate, SKEW_AXIS: AxisSkew}
    if mode == NOT_MOVING:
        return None
    if mode == TRANSLATE_ROTATE_XZ:
        return TranslateRotateXZ(rotation_scale)
    if mode not in table:
        raise ValueError(f"unknown optimize mode {mode}")
    return table[mode]()


class Parametrization:
    """Scene-wide parameter vector over groups of vertices, each moved by its
    own map (thermal_params.cpp:8-156).  groups: list of (vertex indices,
    ParametricMap, model matrix 4x4 or None).  rest_xyz: world positions of
    all vertices at p = 0 (with a model matrix, the rest positions in model
    space are recovered through its inverse)."""

    def __init__(self, rest_xyz, groups):
        self.rest = np.asarray(rest_xyz, np.float64).reshape(-1, 3)
        self.groups = []
        off = 0
        for verts, fmap, M in groups:
            verts = np.asarray(verts, np.int64)
            if verts.size and (verts.min() < 0 or verts.max() >= self.rest.shape[0]):
                raise ValueError("group vertex out of range")
            if len(np.unique(verts)) != verts.size:
                raise ValueError("a group lists a vertex twice")
            M = np.eye(4) if M is None else np.asarray(M, np.float64)
            x0 = (self.rest[verts] - M[:3, 3]) @ np.linalg.inv(M[:3, :3]).T
            self.groups.append((verts, fmap, M, x0, off))
            off += fmap.n_params
        claimed = np.concatenate([g[0] for g in self.groups]) if self.groups else np.zeros(0)
        if len(np.unique(claimed)) != claimed.size:
            raise ValueError("a vertex belongs to two groups")
        self.n_params = off
        self.p = np.zeros(off)

    def positions(self, p):
        """World positions of every vertex at parameters p (others stay at
        rest).  NaN positions are rejected per vertex, as
        thermal_params.cpp:112-118 does."""
        p = self._check(p)
        x = self.rest.copy()
        for verts, fmap, M, x0, off in self.groups:
            loc = fmap.positions(x0, p[off:off + fmap.n_params])
            w = loc @ M[:3, :3].T + M[:3, 3]
            bad = ~np.isfinite(w).all(1)
            w[bad] = self.rest[verts[bad]]
            x[verts] = w
        self.p = p.copy()
        return x

    def triplets(self, p):
        """(vertex, param, dxdp, n_params) with dxdp = d(world x)/dp for
        irtt_project_gradient / irtt_design_gradient: dp[param] += dxdp . g."""
        p = self._check(p)
        vs, ps, ds = [], [], []
        for verts, fmap, M, x0, off in self.groups:
            J = fmap.jacobian(x0, p[off:off + fmap.n_params]) @ M[:3, :3].T
            k = fmap.n_params
            vs.append(np.repeat(verts, k))
            ps.append(np.tile(np.arange(off, off + k), verts.size))
            ds.append(J.reshape(-1, 3))
        if not vs:
            return (np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros((0, 3), np.float32),
                    self.n_params)
        d = np.concatenate(ds)
        d[~np.isfinite(d)] = 0.0
        return (np.concatenate(vs).astype(np.uint32), np.concatenate(ps).astype(np.uint32),
                d.astype(np.float32), self.n_params)

    def _check(self, p):
        p = np.asarray(p, np.float64).ravel()
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        return p


def parse_polygon(cdata):
    """Polygon points from the reference's constraint-data string
    "x0,z0;x1,z1;..." (thermal_params.cpp:520-540) or an (m, 2) array."""
    if isinstance(cdata, str):
        pts = []
        for tok in cdata.split(";"):
            f = [s for s in tok.split(",") if s.strip()]
            if len(f) >= 2:
                pts.append((float(f[0]), float(f[1])))
        return np.asarray(pts, np.float64).reshape(-1, 2)
    return np.asarray(cdata, np.float64).reshape(-1, 2)


class InPolygonXZ:
    """Soft constraint keeping vertices inside a polygon of the XZ plane:
    for a vertex outside (even-odd crossing rule), penalty w/2 d^2 with d the
    distance to the nearest boundary point c, gradient w (x - c) on x and z
    (thermal_params.cpp:542-580)."""

    def __init__(self, vertices, weight, polygon):
        self.verts = np.asarray(vertices, np.int64)
        self.w = float(weight)
        poly = parse_polygon(polygon)
        keep = np.ones(len(poly), bool)
        keep[1:] = (np.diff(poly, axis=0) != 0).any(1)  # drop repeated points
        poly = poly[keep]
        if len(poly) > 1 and (poly[0] == poly[-1]).all():
            poly = poly[:-1]
        if len(poly) < 3:
            raise ValueError("a polygon needs three distinct points")
        self.poly = poly

    def evaluate(self, xyz, grad):
        q = np.asarray(xyz, np.float64)[self.verts][:, [0, 2]]
        a = self.poly
        b = np.roll(a, -1, axis=0)
        e = b - a
        # nearest point on each edge
        t = np.clip(((q[:, None, :] - a[None]) * e[None]).sum(-1) / (e * e).sum(-1)[None], 0.0, 1.0)
        c = a[None] + t[..., None] * e[None]
        d2 = ((q[:, None, :] - c) ** 2).sum(-1)
        k = d2.argmin(1)
        near = c[np.arange(len(q)), k]
        # even-odd crossings of a ray towards +x
        ya, yb = a[None, :, 1], b[None, :, 1]
        straddle = (ya > q[:, None, 1]) != (yb > q[:, None, 1])
        with np.errstate(divide="ignore", invalid="ignore"):
            xc = a[None, :, 0] + (q[:, None, 1] - ya) * (b[None, :, 0] - a[None, :, 0]) / (yb - ya)
        inside = ((straddle & (q[:, None, 0] < xc)).sum(1) % 2) == 1
        out = ~inside
        g = self.w * (q - near)
        g[inside] = 0.0
        np.add.at(grad, (self.verts, 0), g[:, 0])
        np.add.at(grad, (self.verts, 2), g[:, 1])
        return float(0.5 * self.w * d2[np.arange(len(q)), k][out].sum())


def _vertex_area_weights(xyz, idx, verts):
    """Area/3 per corner of every triangle whose corners are all in `verts`,
    normalised to sum 1 (thermal_params.cpp:600-640)."""
    mask = np.zeros(xyz.shape[0], bool)
    mask[verts] = True
    tri = idx[mask[idx].all(1)]
    X = xyz
    a = 0.5 * np.linalg.norm(np.cross(X[tri[:, 1]] - X[tri[:, 0]], X[tri[:, 2]] - X[tri[:, 0]]), axis=1)
    w = np.zeros(xyz.shape[0])
    for c in range(3):
        np.add.at(w, tri[:, c], a / 3.0)
    w = w[verts]
    if not w.sum() > 0:
        raise ValueError("collision model has no triangle area")
    return w / w.sum()


class CollisionXZ:
    """Soft non-overlap of models in the XZ plane: each model is a circle
    around its area-weighted vertex centroid with the radius fixed at set-up
    (half the XZ diagonal of its bounding box).  For an overlapping pair,
    o = |c_i - c_j| - r_i - r_j < 0, penalty w/2 o^2; the gradient w o u
    (u the unit XZ direction from c_j to c_i) is spread over model i's
    vertices by their weights, and its negative over model j's
    (thermal_params.cpp:585-724)."""

    def __init__(self, xyz, idx, models, weight):
        xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
        idx = np.asarray(idx, np.int64).reshape(-1, 3)
        self.w = float(weight)
        self.models = []
        for verts in models:
            verts = np.asarray(verts, np.int64)
            wts = _vertex_area_weights(xyz, idx, verts)
            lo, hi = xyz[verts].min(0), xyz[verts].max(0)
            r = 0.5 * np.hypot(hi[0] - lo[0], hi[2] - lo[2])
            self.models.append((verts, wts, r))

    def evaluate(self, xyz, grad):
        xyz = np.asarray(xyz, np.float64)
        ctr = [wts @ xyz[v][:, [0, 2]] for v, wts, _ in self.models]
        total = 0.0
        for i in range(len(self.models)):
            for j in range(i + 1, len(self.models)):
                diff = ctr[i] - ctr[j]
                n = np.hypot(*diff)
                o = n - self.models[i][2] - self.models[j][2]
                if o < 0.0 and n > 0.0:
                    total += 0.5 * self.w * o * o
                    g = self.w * o * diff / n
                    for (v, wts, _), sgn in ((self.models[i], 1.0), (self.models[j], -1.0)):
                        np.add.at(grad, (v, 0), sgn * wts * g[0])
                        np.add.at(grad, (v, 2), sgn * wts * g[1])
        return float(total)


class Constraints:
    """Sum of penalty terms; evaluate(xyz, grad) adds d(penalty)/dx to grad
    (V x 3) and returns the penalty (thermal_params.cpp:480-518)."""

    def __init__(self, terms=()):
        self.terms = list(terms)

    def evaluate(self, xyz, grad):
        return float(sum(t.evaluate(xyz, grad) for t in self.terms))


class Objective:
    """L = 1/2 sum_i w_i (x_i - t_i)^2 with w_i = weight_i * area_i, one
    target vector per time step (thermal_objective.cpp:1-78).  A target given
    as NaN is taken from the first solution handed to update_targets (the
    reference's "solve" targets)."""

    def __init__(self, weights, areas, targets, n_steps=1):
        self.w = np.asarray(weights, np.float64) * np.asarray(areas, np.float64)
        t = np.broadcast_to(np.asarray(targets, np.float64), self.w.shape).copy()
        self.targets = [t.copy() for _ in range(n_steps)]

    @property
    def missing(self):
        return sum(int(np.isnan(t).any()) for t in self.targets)

    def update_targets(self, x, step=0):
        if step < len(self.targets):
            t = self.targets[step]
            nan = np.isnan(t)
            t[nan] = np.asarray(x, np.float64)[nan]
        return self.missing > 0

    def eval(self, x, step=0):
        """(value, dL/dx).  Zero while targets are missing or for a step
        without targets, as thermal_objective.cpp:47-50 does."""
        x = np.asarray(x, np.float64)
        if self.missing or step >= len(self.targets):
            return 0.0, np.zeros_like(x)
        d = x - self.targets[step]
        g = self.w * d
        return float(0.5 * (g * d).sum()), g


class GradientDescent:
    """p <- p - step_size g, remembering the parameters with the lowest
    objective handed in (thermal_objective.hpp:19-47)."""

    def __init__(self, step_size=1.0):
        self.step_size = float(step_size)
        self.best_value = np.inf
        self.best_params = None

    def reset(self, step_size=None):
        if step_size is not None:
            self.step_size = float(step_size)

    def _remember(self, p, value):
        if self.best_params is None or value < self.best_value:
            self.best_value = float(value)
            self.best_params = np.array(p, np.float64)

    def step(self, p, grad, value):
        """The objective `value` belongs to p (before the update)."""
        self._remember(p, value)
        return np.asarray(p, np.float64) - self.step_size * np.asarray(grad, np.float64)


class Adam(GradientDescent):
    """Adam (Kingma and Ba, 2015, with the bias corrections folded into the
    step size, as thermal_objective.hpp:49-78 does):
    m <- b1 m + (1-b1) g, v <- b2 v + (1-b2) g^2,
    p <- p - a sqrt(1-b2^t)/(1-b1^t) m / (sqrt(v) + eps)."""

    def __init__(self, step_size=1.0, beta1=0.9, beta2=0.999, eps=1e-8):
        super().__init__(step_size)
        self.b1, self.b2, self.eps = float(beta1), float(beta2), float(eps)
        self.m = self.v = None
        self.t = 0

    def reset(self, step_size=None):
        super().reset(step_size)
        self.m = self.v = None
        self.t = 0

    def step(self, p, grad, value):
        self._remember(p, value)
        p = np.asarray(p, np.float64)
        g = np.asarray(grad, np.float64)
        if self.m is None or self.m.shape != p.shape:
            self.m, self.v, self.t = np.zeros_like(p), np.zeros_like(p), 0
        self.t += 1
        self.m = self.b1 * self.m + (1.0 - self.b1) * g
        self.v = self.b2 * self.v + (1.0 - self.b2) * g * g
        a = self.step_size * np.sqrt(1.0 - self.b2 ** self.t) / (1.0 - self.b1 ** self.t)
        return p - a * self.m / (np.sqrt(self.v) + self.eps)
