"""CPU oracle for the radiative-transport hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or the timed CPU
baseline), never as a product code path.

* ``liboracle.so`` (built from ``irtt_oracle.c`` by ``make``) is a plain-C
  restatement of the reference's photon tracer
  (``assets/shader/raytracing_thermal/thermal_tracing_unified.glsl:247-601``),
  its RNG and samplers (``assets/shader/utils/glsl/rendering_utils.glsl:212-272,
  361-392``, ``ray_tracing_utils.glsl:57-60,178-182,217-220,242-259``), a
  watertight ray/triangle test, brute-force and LBVH closest-hit traversal.
* ``solver.py`` is an FP64 numpy restatement of the Newton solve
  (``src/implementations/thermal_renderer/thermal_solver.cpp:13-229``).
* ``closed_forms.py`` holds analytic view factors and the physics-assembly
  restatement used to pin the above.

Pinning (what independently checks each oracle function) is listed in each
module's docstring and exercised by ``tests/test_oracle_*.py``.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-B", "-C", _HERE], check=True,
                   stdout=subprocess.DEVNULL)


# The reference's generated gradient headers compiled into a checker of the
# oracle's gradient formulas (``make ref``; see ``ref_grad_harness.c``).  Built
# only where the reference tree is present; the library lands in ``_ref/``.
REF_SHADERS = "/root/reference/assets/shader/raytracing_thermal"
REF_GPL = "/root/reference/src/implementations/thermal_renderer/gpl"
_REF_PATH = os.path.join(_HERE, "_ref", "libref_grad.so")
_REF_SOLAR_PATH = os.path.join(_HERE, "_ref", "libref_solar.so")
_ref_solar = None
_ref_lib = None


def build_ref():
    """Compile ``_ref/libref_grad.so`` (the reference's gradient headers) and
    ``_ref/libref_solar.so`` (its solar position routine) from the sources
    where they lie; returns False when the reference tree is absent."""
    if not os.path.isdir(REF_SHADERS) or not os.path.isdir(REF_GPL):
        return False
    subprocess.run(["make", "-B", "-C", _HERE, "ref", f"REF_SHADERS={REF_SHADERS}",
                    f"REF_GPL={REF_GPL}"], check=True, stdout=subprocess.DEVNULL)
    return True


def ref_sun_zenith(latitude, longitude, utc, year, month, day, hour, minute=0):
    """The reference's sunAngle (radians; west-positive longitude and time
    zone), or None when neither the library nor the reference tree is here."""
    global _ref_solar
    if _ref_solar is None:
        if not os.path.exists(_REF_SOLAR_PATH) and not build_ref():
            return None
        L = ctypes.CDLL(_REF_SOLAR_PATH)
        L.ref_sun_zenith.restype = ctypes.c_double
        L.ref_sun_zenith.argtypes = [ctypes.c_double, ctypes.c_double] + [ctypes.c_int] * 6
        _ref_solar = L
    return _ref_solar.ref_sun_zenith(latitude, longitude, utc, year, month, day, hour, minute)


def ref_lib():
    """The compiled reference gradient headers, or None when neither the
    library nor the reference tree is here."""
    global _ref_lib
    if _ref_lib is None:
        if not os.path.exists(_REF_PATH) and not build_ref():
            return None
        L = ctypes.CDLL(_REF_PATH)
        d = ctypes.POINTER(ctypes.c_double)
        L.ref_area_gradient.argtypes = [d] * 6
        L.ref_geom_term_grad.restype = ctypes.c_double
        L.ref_geom_term_grad.argtypes = [d, d, ctypes.c_double, d]
        _ref_lib = L
    return _ref_lib


def ref_area_gradient(a, b, c):
    """d area / d{a, b, c} as the reference's symGradArea3D.h computes it (FP64)."""
    arrs = [np.ascontiguousarray(v, np.float64) for v in (a, b, c)]
    out = [np.empty(3) for _ in range(3)]
    d = ctypes.POINTER(ctypes.c_double)
    ref_lib().ref_area_gradient(*[v.ctypes.data_as(d) for v in arrs + out])
    return out


def ref_geom_term_grad(verts6, bi, bj, eps=1e-2):
    """(G, dG/dv (6, 3)) as the reference's symGradGeomTerm3D.h computes them (FP64)."""
    v = np.ascontiguousarray(np.asarray(verts6, np.float64).reshape(18))
    beta = np.array([bi[0], bi[1], bj[0], bj[1]], np.float64)
    out = np.empty(18)
    d = ctypes.POINTER(ctypes.c_double)
    G = ref_lib().ref_geom_term_grad(v.ctypes.data_as(d), beta.ctypes.data_as(d), float(eps),
                                     out.ctypes.data_as(d))
    return G, out.reshape(6, 3)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class Scene(ctypes.Structure):
    _fields_ = [("nv", ctypes.c_uint32), ("nt", ctypes.c_uint32),
                ("xyz", _f32p), ("idx", _u32p), ("rho_d", _f32p),
                ("rho_s", _f32p), ("flags", _u8p)]


class BVH(ctypes.Structure):
    _fields_ = [("n_leaves", ctypes.c_uint32), ("morton", _u32p),
                ("tri_of_leaf", _u32p), ("nodes", _u32p), ("tris", _u32p),
                ("max_depth", ctypes.c_uint32)]


class TravStats(ctypes.Structure):
    _fields_ = [("node_visits", ctypes.c_uint64), ("tri_tests", ctypes.c_uint64)]


class Params(ctypes.Structure):
    _fields_ = [("threads_per_tri", ctypes.c_uint32),
                ("rays_per_thread", ctypes.c_uint32),
                ("batch_seed", ctypes.c_uint32), ("max_depth", ctypes.c_uint32),
                ("bin_mode", ctypes.c_uint32), ("tiny_dist", ctypes.c_float)]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("photons", "rays", "absorbed", "miss", "too_close", "backface",
                 "degenerate", "forced_absorb", "node_visits", "tri_tests")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


def _declare(L):
    L.or_tea16.restype = ctypes.c_uint32
    L.or_tea16.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.or_rnd_stream.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _f32p]
    L.or_sincos_turns.argtypes = [ctypes.c_float, _f32p, _f32p]
    L.or_sample_triangle.argtypes = [ctypes.c_float, ctypes.c_float, _f32p]
    L.or_sample_cosine_hemisphere.argtypes = [ctypes.c_float, ctypes.c_float, _f32p]
    L.or_tangent_to_world.argtypes = [_f32p, _f32p, _f32p]
    L.or_offset_ray.argtypes = [_f32p, _f32p, _f32p]
    L.or_tri_hit.restype = ctypes.c_int
    L.or_tri_hit.argtypes = [_f32p] * 5 + [ctypes.c_float, ctypes.c_float,
                                           _f32p, _f32p, _f32p]
    L.or_triangle_areas.argtypes = [ctypes.POINTER(Scene), _f32p]
    L.or_bvh_build.restype = ctypes.c_int
    L.or_bvh_build.argtypes = [ctypes.POINTER(Scene), ctypes.POINTER(BVH)]
    L.or_bvh_free.argtypes = [ctypes.POINTER(BVH)]
    L.or_trace_bvh.argtypes = [ctypes.POINTER(BVH), ctypes.c_uint32, _f32p,
                               ctypes.c_float, ctypes.c_float, _i32p, _f32p,
                               _f32p, _f32p, ctypes.POINTER(TravStats)]
    L.or_trace_brute.argtypes = [ctypes.POINTER(Scene), ctypes.c_uint32, _f32p,
                                 ctypes.c_float, ctypes.c_float, _i32p, _f32p,
                                 _f32p, _f32p, _u32p]
    L.or_photon_weights.restype = ctypes.c_double
    L.or_photon_weights.argtypes = [ctypes.POINTER(Scene), _f32p,
                                    ctypes.POINTER(Params), _f32p, _u64p]
    L.or_transport.restype = ctypes.c_int
    L.or_transport.argtypes = [ctypes.POINTER(Scene), ctypes.POINTER(Params),
                               ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u64p, _u32p,
                               _u32p, _u32p, ctypes.c_uint64, _u64p, ctypes.c_int,
                               ctypes.POINTER(Stats)]


def _p(a, t):
    return a.ctypes.data_as(t)


class OracleScene:
    """Keeps the numpy arrays alive behind an ``or_scene``."""

    def __init__(self, xyz, idx, rho_d, rho_s, flags):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        self.idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 3)
        nt = self.idx.shape[0]
        self.rho_d = np.ascontiguousarray(np.broadcast_to(rho_d, (nt,)), dtype=np.float32)
        self.rho_s = np.ascontiguousarray(np.broadcast_to(rho_s, (nt,)), dtype=np.float32)
        self.flags = np.ascontiguousarray(np.broadcast_to(flags, (nt,)), dtype=np.uint8)
        self.c = Scene(self.xyz.shape[0], nt, _p(self.xyz, _f32p), _p(self.idx, _u32p),
                       _p(self.rho_d, _f32p), _p(self.rho_s, _f32p),
                       _p(self.flags, _u8p))

    @classmethod
    def from_irtt(cls, s):
        """From an ``irtt.scenes.SceneData``."""
        return cls(s.xyz, s.idx, s.rho_d, s.rho_s, s.flags)

    @property
    def nv(self):
        return self.xyz.shape[0]

    @property
    def nt(self):
        return self.idx.shape[0]


def tea16(v0, v1):
    return int(lib().or_tea16(v0 & 0xffffffff, v1 & 0xffffffff))


def rnd_stream(seed, n):
    out = np.empty(n, np.float32)
    lib().or_rnd_stream(seed & 0xffffffff, n, _p(out, _f32p))
    return out


def sincos_turns(u):
    s = ctypes.c_float()
    c = ctypes.c_float()
    lib().or_sincos_turns(float(u), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def sample_triangle(u0, u1):
    b = np.empty(3, np.float32)
    lib().or_sample_triangle(float(u0), float(u1), _p(b, _f32p))
    return b


def sample_cosine_hemisphere(u0, u1):
    d = np.empty(3, np.float32)
    lib().or_sample_cosine_hemisphere(float(u0), float(u1), _p(d, _f32p))
    return d


def tangent_to_world(v, n):
    v = np.ascontiguousarray(v, np.float32)
    n = np.ascontiguousarray(n, np.float32)
    out = np.empty(3, np.float32)
    lib().or_tangent_to_world(_p(v, _f32p), _p(n, _f32p), _p(out, _f32p))
    return out


def offset_ray(p, n):
    p = np.ascontiguousarray(p, np.float32)
    n = np.ascontiguousarray(n, np.float32)
    out = np.empty(3, np.float32)
    lib().or_offset_ray(_p(p, _f32p), _p(n, _f32p), _p(out, _f32p))
    return out


def tri_hit(o, d, p0, p1, p2, tmin=1e-12, tmax=1e12):
    arrs = [np.ascontiguousarray(a, np.float32) for a in (o, d, p0, p1, p2)]
    t = ctypes.c_float()
    b1 = ctypes.c_float()
    b2 = ctypes.c_float()
    hit = lib().or_tri_hit(*[_p(a, _f32p) for a in arrs], tmin, tmax,
                           ctypes.byref(t), ctypes.byref(b1), ctypes.byref(b2))
    return (bool(hit), t.value, b1.value, b2.value)


def triangle_areas(sc):
    out = np.empty(sc.nt, np.float32)
    lib().or_triangle_areas(ctypes.byref(sc.c), _p(out, _f32p))
    return out


class OracleBVH:
    def __init__(self, sc):
        self._bvh = BVH()
        rc = lib().or_bvh_build(ctypes.byref(sc.c), ctypes.byref(self._bvh))
        if rc != 0:
            raise RuntimeError("or_bvh_build failed")
        n = self._bvh.n_leaves
        self.n_leaves = n
        self.max_depth = self._bvh.max_depth
        if n:
            self.morton = np.ctypeslib.as_array(self._bvh.morton, (n,)).copy()
            self.tri_of_leaf = np.ctypeslib.as_array(self._bvh.tri_of_leaf, (n,)).copy()
            self.tris = np.ctypeslib.as_array(self._bvh.tris, (n * 12,)).copy().reshape(n, 12)
            ni = max(n - 1, 1)
            self.nodes = np.ctypeslib.as_array(self._bvh.nodes, (ni * 16,)).copy().reshape(ni, 16)[: n - 1]

    def trace(self, rays, tmin=1e-12, tmax=1e12):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = rays.shape[0]
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        b1 = np.empty(n, np.float32)
        b2 = np.empty(n, np.float32)
        st = TravStats()
        lib().or_trace_bvh(ctypes.byref(self._bvh), n, _p(rays, _f32p), tmin, tmax,
                           _p(prim, _i32p), _p(t, _f32p), _p(b1, _f32p),
                           _p(b2, _f32p), ctypes.byref(st))
        return prim, t, b1, b2, (int(st.node_visits), int(st.tri_tests))

    def __del__(self):
        try:
            lib().or_bvh_free(ctypes.byref(self._bvh))
        except Exception:
            pass


def trace_brute(sc, rays, tmin=1e-12, tmax=1e12):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = rays.shape[0]
    prim = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    b1 = np.empty(n, np.float32)
    b2 = np.empty(n, np.float32)
    ties = ctypes.c_uint32()
    lib().or_trace_brute(ctypes.byref(sc.c), n, _p(rays, _f32p), tmin, tmax,
                         _p(prim, _i32p), _p(t, _f32p), _p(b1, _f32p),
                         _p(b2, _f32p), ctypes.byref(ties))
    return prim, t, b1, b2, int(ties.value)


def make_params(threads_per_tri, rays_per_thread, batch_seed=1234, max_depth=512,
                bin_mode=0, tiny_dist=1e-3):
    return Params(threads_per_tri, rays_per_thread, batch_seed, max_depth,
                  bin_mode, tiny_dist)


def photon_weights(sc, params):
    area = triangle_areas(sc)
    w = np.empty(sc.nt, np.float32)
    q = np.empty(sc.nt, np.uint64)
    Q = lib().or_photon_weights(ctypes.byref(sc.c), _p(area, _f32p),
                                ctypes.byref(params), _p(w, _f32p), _p(q, _u64p))
    return w, q, Q


def transport(sc, params, tri_begin=0, tri_end=None, dense=True, records=0,
              n_threads=0, tri_stride=1, with_tri=False):
    """Runs the oracle photon tracer.  Returns (acc or None, (abs, src) records
    -- plus each record's source triangle with ``with_tri`` -- or None, stats
    dict)."""
    if tri_end is None:
        tri_end = sc.nt
    nb = sc.nt if params.bin_mode else sc.nv
    acc = np.zeros((nb, nb), np.uint64) if dense else None
    ra = rs = rt = None
    with_tri = bool(with_tri and records)
    if records:
        ra = np.empty(records, np.uint32)
        rs = np.empty(records, np.uint32)
        if with_tri:
            rt = np.empty(records, np.uint32)
    cnt = ctypes.c_uint64()
    st = Stats()
    rc = lib().or_transport(ctypes.byref(sc.c), ctypes.byref(params), tri_begin,
                            tri_end, tri_stride, _p(acc, _u64p) if dense else None,
                            _p(ra, _u32p) if records else None,
                            _p(rs, _u32p) if records else None,
                            _p(rt, _u32p) if with_tri else None, records,
                            ctypes.byref(cnt), n_threads, ctypes.byref(st))
    if rc != 0:
        raise RuntimeError("or_transport failed")
    recs = None
    if records:
        recs = (ra[: cnt.value], rs[: cnt.value])
        if with_tri:
            recs += (rt[: cnt.value],)
    return acc, recs, st.as_dict()


ADJ_AREA = 1
ADJ_CONTINUOUS = 2
ADJ_EDGE = 4
ADJ_REFERENCE_DROPPED = 8
ADJ_EDGE_GAIN = 16
ADJ_HIT_AREA = 32
ADJ_EDGE_PROJECTED = 64
ADJ_COLLIMATED = 128


class AdjointParams(ctypes.Structure):
    _fields_ = [("edge_ratio", ctypes.c_float), ("tiny_eps", ctypes.c_float),
                ("flags", ctypes.c_uint32)]


class AdjointStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("photons", "absorbed", "segments", "continuous", "edge_hits", "occluded",
                 "rays")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


def _declare_adjoint(L):
    _f64p = ctypes.POINTER(ctypes.c_double)
    L.or_area_gradient.argtypes = [_f32p] * 6
    L.or_geom_term_grad.restype = ctypes.c_float
    L.or_geom_term_grad.argtypes = [_f32p] * 6 + [ctypes.c_float] * 5 + [_f32p]
    L.or_adjoint.restype = ctypes.c_int
    L.or_adjoint.argtypes = [ctypes.POINTER(Scene), ctypes.POINTER(Params),
                             ctypes.POINTER(AdjointParams), _f64p, _f64p, _f64p, ctypes.c_int,
                             ctypes.POINTER(AdjointStats)]


def area_gradient(a, b, c):
    arrs = [np.ascontiguousarray(v, np.float32) for v in (a, b, c)]
    out = [np.empty(3, np.float32) for _ in range(3)]
    L = lib()
    _declare_adjoint(L)
    L.or_area_gradient(*[_p(v, _f32p) for v in arrs + out])
    return out


def geom_term_grad(verts6, bi, bj, eps=1e-2):
    vs = [np.ascontiguousarray(v, np.float32) for v in verts6]
    out = np.empty(18, np.float32)
    L = lib()
    _declare_adjoint(L)
    G = L.or_geom_term_grad(*[_p(v, _f32p) for v in vs], float(bi[0]), float(bi[1]),
                            float(bj[0]), float(bj[1]), float(eps), _p(out, _f32p))
    return G, out.reshape(6, 3)


def adjoint(sc, params, lam, x4, edge_ratio=0.05, tiny_eps=1e-2,
            flags=ADJ_AREA | ADJ_CONTINUOUS | ADJ_EDGE, n_threads=0):
    """Adjoint photon trace; returns (grad (nv, 3) float64, stats)."""
    L = lib()
    _declare_adjoint(L)
    lam = np.ascontiguousarray(lam, np.float64)
    x4 = np.ascontiguousarray(x4, np.float64)
    grad = np.zeros(3 * sc.nv)
    ap = AdjointParams(edge_ratio, tiny_eps, flags)
    st = AdjointStats()
    _f64p = ctypes.POINTER(ctypes.c_double)
    rc = L.or_adjoint(ctypes.byref(sc.c), ctypes.byref(params), ctypes.byref(ap),
                      lam.ctypes.data_as(_f64p), x4.ctypes.data_as(_f64p),
                      grad.ctypes.data_as(_f64p), n_threads, ctypes.byref(st))
    if rc != 0:
        raise RuntimeError("or_adjoint failed")
    return grad.reshape(-1, 3), st.as_dict()
