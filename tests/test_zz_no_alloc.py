"""irtt_trace_rays_noalloc: ray queries staged through a static device
array instead of allocated buffers.  Called through the C ABI; its answers
must equal irtt_trace_rays's bit for bit, on two calls in a row (the second
with fewer rays, so stale answers of the first would show), and it must
refuse more rays than its array holds and a scene without a BVH."""
import ctypes

import numpy as np
import pytest

from conftest import random_rays

pytestmark = pytest.mark.gpu


def _noalloc(irtt, h, rays, tmin=1e-12, tmax=1e12):
    L = irtt.lib()
    f = L.irtt_trace_rays_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float,
                  ctypes.c_float] + [ctypes.c_void_p] * 4
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = rays.shape[0]
    prim = np.full(n, -7, np.int32)
    t = np.full(n, np.nan, np.float32)
    b1 = np.full(n, np.nan, np.float32)
    b2 = np.full(n, np.nan, np.float32)
    rc = f(h, rays.ctypes.data, n, tmin, tmax, prim.ctypes.data, t.ctypes.data, b1.ctypes.data,
           b2.ctypes.data)
    return rc, prim, t, b1, b2


def test_noalloc_matches_trace_rays(irtt):
    from irtt import scenes
    sd = scenes.city_block(seed=3, nx=2, ny=2, target_tris=3000)
    rng = np.random.default_rng(11)
    lo, hi = sd.xyz.min(0), sd.xyz.max(0)
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_bvh()
        L = irtt.lib()
        L.irtt_trace_rays_noalloc_max.restype = ctypes.c_uint32
        cap = L.irtt_trace_rays_noalloc_max()
        assert cap == 1 << 18
        for n in (20000, 3000):   # the second call smaller than the first
            rays = random_rays(rng, n, lo, hi)
            want = s.trace_rays(rays)
            rc, prim, t, b1, b2 = _noalloc(irtt, s._h, rays)
            assert rc == 0, irtt.lib().irtt_last_error()
            print(f"n={n}: hits {np.sum(prim >= 0)}")
            assert np.array_equal(prim, want[0])
            assert np.array_equal(t, want[1])
            assert np.array_equal(b1, want[2]) and np.array_equal(b2, want[3])
        # more rays than the array holds: refused before anything is read
        small = random_rays(rng, 4, lo, hi)
        f = L.irtt_trace_rays_noalloc
        out = [np.empty(4, dt) for dt in (np.int32, np.float32, np.float32, np.float32)]
        rc = f(s._h, small.ctypes.data, cap + 1, 1e-12, 1e12, *[o.ctypes.data for o in out])
        assert rc == irtt.ERR_LIMIT
        assert b"maximum" in L.irtt_last_error()


def test_noalloc_needs_a_bvh(irtt):
    from irtt import scenes
    sd = scenes.parallel_plates(n=4)
    rng = np.random.default_rng(2)
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        rc, *_ = _noalloc(irtt, s._h, random_rays(rng, 8, -1.0, 2.0))
        assert rc == irtt.ERR_STATE
