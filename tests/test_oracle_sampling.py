"""Pins the oracle's samplers, frames, ray offset and ray/triangle test
(ray_tracing_utils.glsl:57-60,178-182,217-220,242-259; rendering_utils.glsl:212-219,261-272)
by distributions, invariants and closed forms."""
import struct

import numpy as np
import pytest


def test_triangle_sampler_uniform_and_nearest_corner_thirds(oracle):
    u = oracle.rnd_stream(12345, 2 * 60000).reshape(-1, 2)
    b = np.array([oracle.sample_triangle(a, c) for a, c in u])
    assert np.all(np.abs(b.sum(1) - 1.0) < 1e-6)
    assert np.all(b >= 0.0)
    # uniform density: the 4 midpoint sub-triangles hold 1/4 of the area each
    # sub-triangle index from which barycentric exceeds 1/2 (corner) or none (centre)
    corner = np.where(b.max(1) > 0.5, b.argmax(1), 3)
    frac = np.bincount(corner, minlength=4) / len(b)
    sig = np.sqrt(0.25 * 0.75 / len(b))
    assert np.all(np.abs(frac - 0.25) < 5 * sig), frac
    # nearest-corner (argmax) regions are exactly thirds of the triangle
    near = np.bincount(b.argmax(1), minlength=3) / len(b)
    sig3 = np.sqrt(2 / 9 / len(b))
    assert np.all(np.abs(near - 1 / 3) < 5 * sig3), near


def test_cosine_hemisphere_moments(oracle):
    u = oracle.rnd_stream(777, 2 * 60000).reshape(-1, 2)
    d = np.array([oracle.sample_cosine_hemisphere(a, c) for a, c in u])
    assert np.all(np.abs(np.linalg.norm(d, axis=1) - 1.0) < 1e-6)
    assert np.all(d[:, 2] >= 0.0)
    n = len(d)
    # cosine-weighted: E[cos] = 2/3, E[cos^2] = 1/2, E[x] = E[y] = 0
    assert abs(d[:, 2].mean() - 2 / 3) < 5 * np.sqrt(1 / 18 / n)
    assert abs((d[:, 2] ** 2).mean() - 0.5) < 5 * np.sqrt(1 / 12 / n)
    assert abs(d[:, 0].mean()) < 5 * np.sqrt(0.25 / n)
    assert abs(d[:, 1].mean()) < 5 * np.sqrt(0.25 / n)


def test_tangent_frame_orthonormal(oracle):
    rng = np.random.default_rng(3)
    for _ in range(500):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        w = oracle.tangent_to_world(v, n)
        assert abs(np.linalg.norm(w) - 1.0) < 1e-5
        assert abs(np.dot(w, n.astype(np.float32)) - v[2]) < 1e-5
    # axis-aligned normals use the branch-free perpendicular of the reference
    for n in ([0, 0, 1], [0, 0, -1], [1, 0, 0], [0, 1, 0]):
        w = oracle.tangent_to_world([0, 0, 1], n)
        assert np.allclose(w, n, atol=1e-7)


def _f2i(f):
    return struct.unpack("<i", struct.pack("<f", f))[0]


def _i2f(i):
    return struct.unpack("<f", struct.pack("<i", i))[0]


def test_offset_ray_integer_ulp_rule(oracle):
    # Ray Tracing Gems ch. 6: p + n/65536 near the origin, else +- int(256 n) ulps
    rng = np.random.default_rng(5)
    for _ in range(300):
        p = (rng.normal(size=3) * rng.choice([0.01, 1.0, 100.0])).astype(np.float32)
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        out = oracle.offset_ray(p, n)
        for k in range(3):
            if abs(p[k]) < 1.0 / 32.0:
                want = np.float32(p[k] + np.float32(np.float32(1.0 / 65536.0) * n[k]))
            else:
                of = int(np.float32(256.0) * n[k])
                want = np.float32(_i2f(_f2i(float(p[k])) + (-of if p[k] < 0 else of)))
            assert out[k] == want


def test_tri_hit_closed_form_and_miss(oracle):
    p0, p1, p2 = [0, 0, 0], [2, 0, 0], [0, 2, 0]
    hit, t, b1, b2 = oracle.tri_hit([0.5, 0.5, 3.0], [0, 0, -1], p0, p1, p2)
    assert hit and t == 3.0 and b1 == 0.25 and b2 == 0.25
    # from below (two-sided)
    hit, t, _, _ = oracle.tri_hit([0.5, 0.5, -1.5], [0, 0, 1], p0, p1, p2)
    assert hit and t == 1.5
    # outside the hypotenuse
    assert not oracle.tri_hit([1.5, 1.5, 1.0], [0, 0, -1], p0, p1, p2)[0]
    # behind the origin / beyond tmax
    assert not oracle.tri_hit([0.5, 0.5, 3.0], [0, 0, 1], p0, p1, p2)[0]
    assert not oracle.tri_hit([0.5, 0.5, 3.0], [0, 0, -1], p0, p1, p2, tmax=2.0)[0]
    # parallel ray in the plane: no hit
    assert not oracle.tri_hit([-1, 0.5, 0], [1, 0, 0], p0, p1, p2)[0]


def test_tri_hit_watertight_on_shared_edges(oracle):
    """Rays aimed exactly at the shared diagonal of a quad (and at its
    vertices) never pass between the two triangles."""
    q = [np.array(v, np.float32) for v in ([0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0])]
    rng = np.random.default_rng(11)
    leaks = 0
    for _ in range(3000):
        s = np.float32(0.01 + 0.98 * rng.random())
        target = np.array([s, s, 0], np.float32)  # on the diagonal v0-v2
        o = (target + rng.normal(size=3) * [0.3, 0.3, 0] + [0, 0, 1 + rng.random()]).astype(np.float32)
        d = (target - o).astype(np.float32)
        d /= np.float32(np.linalg.norm(d))
        h1 = oracle.tri_hit(o, d, q[0], q[1], q[2])[0]
        h2 = oracle.tri_hit(o, d, q[0], q[2], q[3])[0]
        leaks += not (h1 or h2)
    assert leaks == 0
