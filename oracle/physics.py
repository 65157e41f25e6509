"""Physics assembly restated in numpy FP64 (test infrastructure only).

Follows ``src/implementations/thermal_renderer/thermal_data.cpp``:
triangle areas with the 10*FLT_EPSILON clamp (:56-81), vertex areas A/3,
emissivity 1 - rho_d - rho_s (thermal_scene.cpp:87-97), lumped mass
rho*c*thickness*A/3 from the unclamped area (:216-251) and the cotangent
Laplacian scaled by conductivity*thickness (:253-313).

Pinned by: vertex areas summing to the surface area, mass = areal capacity x
area, and the Laplacian's symmetry, zero row sums and exactness on linear
functions at interior vertices of a planar mesh (tests/test_oracle_physics.py).
"""
import numpy as np

FLT_EPSILON = np.finfo(np.float32).eps


def triangle_areas(xyz, idx):
    p = np.asarray(xyz, np.float32).astype(np.float64)[np.asarray(idx)]
    c = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    raw = 0.5 * np.sqrt((c * c).sum(1))
    return raw, np.maximum(raw, 10.0 * FLT_EPSILON)


def bins(sd, bin_mode):
    """(area, eps*area, mass, source power) per bin."""
    raw, area = triangle_areas(sd.xyz, sd.idx)
    rd = sd.rho_d.astype(np.float32)
    rs = sd.rho_s.astype(np.float32)
    tot = rd + rs
    scale = np.where(tot > 1.0, tot, np.float32(1.0)).astype(np.float32)
    rd = (rd / scale).astype(np.float32)
    rs = (rs / scale).astype(np.float32)
    eps = 1.0 - rd.astype(np.float64) - rs.astype(np.float64)
    rct = sd.density.astype(np.float64) * sd.heat_capacity.astype(np.float64) * \
        sd.thickness.astype(np.float64)
    src = sd.heat_flux.astype(np.float64) * area
    if bin_mode == 1:
        return area, eps * area, raw * rct, src
    nv = sd.nv
    out = []
    for v in (area / 3.0, eps * area / 3.0, raw / 3.0 * rct, src / 3.0):
        acc = np.zeros(nv)
        for c in range(3):
            np.add.at(acc, sd.idx[:, c], v)
        out.append(acc)
    return tuple(out)


def cotan_laplacian(xyz, idx, k_per_tri, nv):
    """Dense cotangent Laplacian (vertex bins), thermal_data.cpp:253-313."""
    X = np.asarray(xyz, np.float32).astype(np.float64)
    L = np.zeros((nv, nv))

    def cot(a, b):
        aa, ab, bb = a @ a, a @ b, b @ b
        den = aa * bb - ab * ab
        return ab / np.sqrt(den) if den > 0 else 0.0

    for t, (i0, i1, i2) in enumerate(np.asarray(idx)):
        k = float(k_per_tri[t])
        if k == 0.0:
            continue
        a, b, c = X[i0], X[i1], X[i2]
        c1, c2, c3 = cot(a - c, b - c), cot(b - a, c - a), cot(c - b, a - b)
        L[i0, i0] += k * 0.5 * (c1 + c3)
        L[i0, i1] -= k * 0.5 * c1
        L[i0, i2] -= k * 0.5 * c3
        L[i1, i0] -= k * 0.5 * c1
        L[i1, i1] += k * 0.5 * (c1 + c2)
        L[i1, i2] -= k * 0.5 * c2
        L[i2, i0] -= k * 0.5 * c3
        L[i2, i1] -= k * 0.5 * c2
        L[i2, i2] += k * 0.5 * (c2 + c3)
    return L
