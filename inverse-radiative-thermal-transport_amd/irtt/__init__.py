"""ctypes binding of libirtt.so, the MI355X radiative-thermal core.

The product is the C ABI declared in ``include/irtt.h`` (HIP kernels for
gfx950 behind it).  This module only marshals numpy arrays across that ABI; it
has no compute path of its own and raises when the library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# IRTT_LIB selects another build of the same library (e.g. a variant built
# with different compile-time switches for an A/B measurement)
LIB_PATH = os.environ.get("IRTT_LIB") or os.path.join(_HERE, "libirtt.so")
_lib = None

BIN_VERTEX = 0
BIN_TRIANGLE = 1
MAT_DIFFUSE_EMISSION = 1
MAT_TRACEABLE = 2
JAC_EXACT = 0
JAC_REFERENCE = 1


# status codes of the C ABI (include/irtt.h)
OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_STATE = -3
ERR_NOMEM = -4
ERR_LIMIT = -5
ERR_NOCONV = -6


class IrttError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"irtt error {code}: {msg}")
        self.code = code


class Material(ctypes.Structure):
    _fields_ = [("diffuse_reflectance", ctypes.c_float),
                ("specular_reflectance", ctypes.c_float),
                ("thickness", ctypes.c_float), ("density", ctypes.c_float),
                ("heat_capacity", ctypes.c_float),
                ("heat_conductivity", ctypes.c_float),
                ("heat_flux", ctypes.c_float), ("flags", ctypes.c_uint32)]


MATERIAL_DTYPE = np.dtype([("diffuse_reflectance", "<f4"), ("specular_reflectance", "<f4"),
                           ("thickness", "<f4"), ("density", "<f4"),
                           ("heat_capacity", "<f4"), ("heat_conductivity", "<f4"),
                           ("heat_flux", "<f4"), ("flags", "<u4")])


class SceneInfo(ctypes.Structure):
    _fields_ = [("n_vertices", ctypes.c_uint32), ("n_triangles", ctypes.c_uint32),
                ("n_bins", ctypes.c_uint32), ("n_traceable", ctypes.c_uint32),
                ("bvh_max_depth", ctypes.c_uint32), ("bin_mode", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64)]


class TransportParams(ctypes.Structure):
    _fields_ = [("threads_per_tri", ctypes.c_uint32), ("rays_per_thread", ctypes.c_uint32),
                ("batch_seed", ctypes.c_uint32), ("max_depth", ctypes.c_uint32),
                ("tiny_dist", ctypes.c_float), ("layout", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class TransportStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("photons", "rays", "absorbed", "miss", "too_close", "backface",
                 "degenerate", "forced_absorb")] + [
        ("quantum", ctypes.c_double), ("trace_ms", ctypes.c_float),
        ("finalize_ms", ctypes.c_float), ("bvh_ms", ctypes.c_float),
        ("layout", ctypes.c_uint32), ("node_visits", ctypes.c_uint64),
        ("tri_tests", ctypes.c_uint64), ("nnz", ctypes.c_uint64)] + [
        (n, ctypes.c_double) for n in
        ("emitted_w", "absorbed_w", "escaped_w", "dropped_w", "energy_rel")] + [
        ("energy_exact", ctypes.c_uint32), ("reserved2", ctypes.c_uint32)]

    def as_dict(self):
        return {n: (getattr(self, n)) for n, _ in self._fields_}


class SolverParams(ctypes.Structure):
    _fields_ = [("steady", ctypes.c_uint32), ("jacobian", ctypes.c_uint32),
                ("max_newton", ctypes.c_uint32), ("max_linear", ctypes.c_uint32),
                ("dt", ctypes.c_double), ("ambient_kelvin", ctypes.c_double),
                ("x_tol", ctypes.c_double), ("f_tol", ctypes.c_double),
                ("lin_rtol", ctypes.c_double), ("krylov", ctypes.c_uint32),
                ("gmres_restart", ctypes.c_uint32)]


class SolveStats(ctypes.Structure):
    _fields_ = [("newton_iters", ctypes.c_uint32), ("linear_iters", ctypes.c_uint32),
                ("line_search_halvings", ctypes.c_uint32), ("converged", ctypes.c_uint32),
                ("residual_sq", ctypes.c_double), ("max_dx", ctypes.c_double),
                ("solve_ms", ctypes.c_float), ("reserved", ctypes.c_uint32),
                ("power_in", ctypes.c_double), ("power_out", ctypes.c_double),
                ("balance_rel", ctypes.c_double)]

    def as_dict(self):
        return {n: (getattr(self, n)) for n, _ in self._fields_}


_vp = ctypes.c_void_p


class ServeConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint32), ("reserved0", ctypes.c_uint32),
                ("memory_budget", ctypes.c_uint64)]


class AdjointParams(ctypes.Structure):
    _fields_ = [("edge_ratio", ctypes.c_float), ("tiny_eps", ctypes.c_float),
                ("flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class Request(ctypes.Structure):
    _fields_ = [("xyz", _vp), ("nv", ctypes.c_uint32), ("nt", ctypes.c_uint32), ("tri", _vp),
                ("mat", _vp), ("bin_mode", ctypes.c_uint32), ("operator_kind", ctypes.c_uint32),
                ("kelvin0", _vp), ("fixed", _vp), ("transport", TransportParams),
                ("solver", SolverParams), ("target", _vp), ("objective_weight", ctypes.c_double),
                ("nnz", ctypes.c_uint32), ("n_params", ctypes.c_uint32), ("vertex", _vp),
                ("param", _vp), ("dxdp", _vp), ("kelvin", _vp), ("dp", _vp),
                ("objective", _vp), ("adjoint", AdjointParams), ("source", _vp),
                ("group", ctypes.c_uint64)]


class RequestStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("n_bins", ctypes.c_uint32),
                ("queued_ms", ctypes.c_double), ("run_ms", ctypes.c_double),
                ("latency_ms", ctypes.c_double), ("footprint_bytes", ctypes.c_uint64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def _declare(L):
    L.irtt_version.restype = ctypes.c_int
    L.irtt_last_error.restype = ctypes.c_char_p
    L.irtt_device_info.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)]
    L.irtt_scene_create.argtypes = [_vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp,
                                    ctypes.c_uint32, ctypes.POINTER(_vp)]
    L.irtt_scene_destroy.argtypes = [_vp]
    L.irtt_scene_info_get.argtypes = [_vp, ctypes.POINTER(SceneInfo)]
    L.irtt_scene_set_vertex_state.argtypes = [_vp, _vp, _vp]
    L.irtt_scene_set_bin_state.argtypes = [_vp, _vp, _vp]
    L.irtt_scene_update_vertices.argtypes = [_vp, _vp]
    L.irtt_scene_get_bin_data.argtypes = [_vp, _vp, _vp, _vp]
    L.irtt_build_bvh.argtypes = [_vp]
    L.irtt_get_bvh.argtypes = [_vp, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_uint32),
                               ctypes.POINTER(ctypes.c_uint32)]
    L.irtt_trace_rays.argtypes = [_vp, _vp, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                  _vp, _vp, _vp, _vp]
    L.irtt_build_transport.argtypes = [_vp, ctypes.POINTER(TransportParams),
                                       ctypes.POINTER(TransportStats)]
    L.irtt_get_transport.argtypes = [_vp, _vp]
    L.irtt_get_transport_fixed.argtypes = [_vp, _vp, ctypes.POINTER(ctypes.c_double)]
    L.irtt_get_transport_sparse.argtypes = [_vp, ctypes.POINTER(ctypes.c_uint64), _vp, _vp, _vp]
    L.irtt_get_column_sums.argtypes = [_vp, _vp]
    L.irtt_get_transport_line.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32, _vp]
    L.irtt_solve.argtypes = [_vp, ctypes.POINTER(SolverParams), ctypes.POINTER(SolveStats)]
    L.irtt_get_temperatures.argtypes = [_vp, _vp]
    L.irtt_set_temperatures.argtypes = [_vp, _vp]
    L.irtt_residual.argtypes = [_vp, ctypes.POINTER(SolverParams), _vp, _vp]
    L.irtt_adjoint_solve.argtypes = [_vp, ctypes.POINTER(SolverParams), _vp, _vp, _vp,
                                     ctypes.POINTER(SolveStats)]
    L.irtt_adjoint_trace.argtypes = [_vp, ctypes.POINTER(AdjointParams), ctypes.c_uint32, _vp,
                                     _vp, _vp, ctypes.POINTER(AdjointStats)]
    L.irtt_source_gradient.argtypes = [_vp, _vp, _vp]
    L.irtt_design_gradient.argtypes = [_vp, ctypes.POINTER(SolverParams), _vp,
                                       ctypes.POINTER(AdjointParams), ctypes.c_uint32, _vp, _vp,
                                       _vp, ctypes.c_uint32, _vp, _vp, _vp,
                                       ctypes.POINTER(SolveStats)]
    L.irtt_project_gradient.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, _vp,
                                        ctypes.c_uint32, _vp]
    L.irtt_serve_create.argtypes = [ctypes.POINTER(ServeConfig), ctypes.POINTER(_vp)]
    L.irtt_serve_submit.argtypes = [_vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
    L.irtt_serve_poll.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(RequestStatus)]
    L.irtt_serve_wait.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(RequestStatus)]
    L.irtt_serve_destroy.argtypes = [_vp]
    L.irtt_select_operator.argtypes = [_vp, ctypes.c_uint32]
    L.irtt_scene_set_source.argtypes = [_vp, _vp]
    L.irtt_apply_operator.argtypes = [_vp, ctypes.c_uint32, _vp, _vp, ctypes.c_uint32]
    L.irtt_simulate.argtypes = [_vp, ctypes.POINTER(SolverParams), ctypes.c_uint32, _vp,
                                ctypes.POINTER(SolveStats)]
    L.irtt_adjoint_history.argtypes = [_vp, ctypes.POINTER(SolverParams), ctypes.c_uint32, _vp,
                                       _vp, _vp, ctypes.c_uint32, _vp,
                                       ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(AdjointParams), _vp,
                                       ctypes.POINTER(SolveStats)]
    L.irtt_solve_batch.argtypes = [_vp, ctypes.POINTER(SolverParams), ctypes.c_uint32, _vp, _vp,
                                   ctypes.POINTER(SolveStats)]
    L.irtt_adjoint_solve_batch.argtypes = [_vp, ctypes.POINTER(SolverParams), ctypes.c_uint32,
                                           _vp, _vp, _vp, ctypes.POINTER(SolveStats)]
    L.irtt_compress_transport.argtypes = [_vp, ctypes.POINTER(AcaParams), ctypes.POINTER(AcaStats)]
    L.irtt_get_aca.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
    L.irtt_compress_hierarchical.argtypes = [_vp, ctypes.POINTER(HmatrixParams),
                                             ctypes.POINTER(HmatrixStats)]
    L.irtt_get_hmatrix.argtypes = [_vp, _vp, ctypes.POINTER(ctypes.c_uint32), _vp, _vp, _vp]
    L.irtt_get_hmatrix_factors.argtypes = [_vp, ctypes.c_uint32, _vp, _vp]


OP_DENSE = 0
OP_ACA = 1
OP_DENSE_F16 = 2
OP_SPARSE = 3
OP_HMATRIX = 4
HBLOCK_ZERO = 0
HBLOCK_NOCONV = 1
HBLOCK_REJECTED = 2
HBLOCK_LOWRANK = 3
LAYOUT_DENSE = 0
LAYOUT_SPARSE = 1
TRANSPORT_KEEP_ABSORBERS = 1
TRANSPORT_KEEP_PATHS = 2
ACA_FULL = 0
ACA_PARTIAL = 1
ACA_PLUS = 2
STORE_F32 = 0
STORE_F16 = 1


class AcaParams(ctypes.Structure):
    _fields_ = [("algorithm", ctypes.c_uint32), ("rank", ctypes.c_uint32),
                ("column_sum_correction", ctypes.c_uint32), ("seed", ctypes.c_uint32),
                ("precision", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 3)]


class AcaStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("err_inf", "err_frob", "err_inf_rel", "err_frob_rel")] + [
        ("compress_ms", ctypes.c_float), ("rank", ctypes.c_uint32),
        ("n_corrected", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


class HmatrixParams(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("eta", ctypes.c_double),
                ("max_rank", ctypes.c_uint32), ("leaf_size", ctypes.c_uint32),
                ("precision", ctypes.c_uint32), ("allow_larger", ctypes.c_uint32),
                ("keep_factors", ctypes.c_uint32), ("workspace_mib", ctypes.c_uint32)]


class HmatrixStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in
                ("n_clusters", "n_leaves", "n_far", "n_near", "n_lowrank", "n_zero",
                 "n_noconv", "n_reject", "max_rank_used", "reserved")] + [
        ("mean_rank", ctypes.c_double)] + [(n, ctypes.c_uint64) for n in
                                           ("nnz", "nnz_near", "nnz_lowrank", "bytes_sparse",
                                            "bytes_hmatrix", "bytes_lowrank", "build_bytes")] + [
        (n, ctypes.c_double) for n in ("err_frob_rel", "err_frob_rel_stored", "fro_far_rel")] + [
        ("build_ms", ctypes.c_float), ("reserved2", ctypes.c_float)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if not n.startswith("reserved")}


ADJ_AREA = 1
ADJ_CONTINUOUS = 2
ADJ_EDGE = 4
ADJ_REFERENCE_DROPPED = 8
ADJ_EDGE_GAIN = 16
ADJ_HIT_AREA = 32
ADJ_EDGE_PROJECTED = 64
ADJ_COLLIMATED = 128


class AdjointStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("photons", "absorbed", "segments", "continuous", "edge_hits", "occluded",
                 "rays")] + [("trace_ms", ctypes.c_float), ("kept_paths", ctypes.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def project_gradient(grad_x, vertex, param, dxdp, n_params):
    """dp[param[k]] += dxdp[k] . grad_x[vertex[k]] (irtt_project_gradient)."""
    g = np.ascontiguousarray(grad_x, np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(vertex, np.uint32)
    p = np.ascontiguousarray(param, np.uint32)
    d = np.ascontiguousarray(dxdp, np.float32).reshape(-1, 3)
    dp = np.zeros(n_params)
    _check(lib().irtt_project_gradient(_ptr(g), g.shape[0], v.size, _ptr(v), _ptr(p), _ptr(d),
                                       n_params, _ptr(dp)))
    return dp


def lib():
    """Loads libirtt.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise IrttError(-3, f"{LIB_PATH} is missing: run __graft_entry__.build()")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().irtt_last_error()
        raise IrttError(rc, msg.decode() if msg else "")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def device_info():
    buf = ctypes.create_string_buffer(256)
    cu = ctypes.c_int()
    _check(lib().irtt_device_info(buf, 256, ctypes.byref(cu)))
    return buf.value.decode(), cu.value


def make_materials(nt, rho_d=0.0, rho_s=0.0, thickness=1.0, density=1.0, heat_capacity=1.0,
                   conductivity=0.0, heat_flux=0.0, flags=MAT_DIFFUSE_EMISSION | MAT_TRACEABLE):
    m = np.zeros(nt, MATERIAL_DTYPE)
    m["diffuse_reflectance"] = rho_d
    m["specular_reflectance"] = rho_s
    m["thickness"] = thickness
    m["density"] = density
    m["heat_capacity"] = heat_capacity
    m["heat_conductivity"] = conductivity
    m["heat_flux"] = heat_flux
    m["flags"] = flags
    return m


def transport_params(threads_per_tri=32768, rays_per_thread=8, batch_seed=1234, max_depth=512,
                     tiny_dist=1e-3, layout=0, flags=0):
    p = TransportParams()
    p.threads_per_tri = threads_per_tri
    p.rays_per_thread = rays_per_thread
    p.batch_seed = batch_seed
    p.max_depth = max_depth
    p.tiny_dist = tiny_dist
    p.layout = layout
    p.flags = flags
    return p


KRYLOV_BICGSTAB = 0
KRYLOV_GMRES = 1


def solver_params(steady=True, jacobian=JAC_EXACT, max_newton=100, max_linear=300, dt=300.0,
                  ambient_kelvin=-1.0, x_tol=1e-5, f_tol=1e-12, lin_rtol=1e-10, krylov=0,
                  gmres_restart=0):
    p = SolverParams()
    p.krylov = krylov
    p.gmres_restart = gmres_restart
    p.steady = 1 if steady else 0
    p.jacobian = jacobian
    p.max_newton = max_newton
    p.max_linear = max_linear
    p.dt = dt
    p.ambient_kelvin = ambient_kelvin
    p.x_tol = x_tol
    p.f_tol = f_tol
    p.lin_rtol = lin_rtol
    return p


class Scene:
    """One scene/design resident in HBM (an ``irtt_scene`` handle)."""

    def __init__(self, xyz, idx, materials, bin_mode=BIN_VERTEX):
        self._h = None
        self.xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        self.idx = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
        mats = np.ascontiguousarray(materials, MATERIAL_DTYPE)
        if mats.shape[0] != self.idx.shape[0]:
            raise ValueError("one material per triangle")
        h = ctypes.c_void_p()
        _check(lib().irtt_scene_create(_ptr(self.xyz), self.xyz.shape[0], _ptr(self.idx),
                                       self.idx.shape[0], _ptr(mats), bin_mode,
                                       ctypes.byref(h)))
        self._h = h
        self.info = self.get_info()

    @classmethod
    def from_data(cls, sd, bin_mode=BIN_VERTEX):
        """From an ``irtt.scenes.SceneData``."""
        s = cls(sd.xyz, sd.idx, sd.materials(), bin_mode)
        s.set_vertex_state(sd.kelvin, sd.fixed)
        return s

    def close(self):
        if self._h is not None and self._h.value:
            lib().irtt_scene_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def get_info(self):
        i = SceneInfo()
        _check(lib().irtt_scene_info_get(self._h, ctypes.byref(i)))
        return {n: int(getattr(i, n)) for n, _ in i._fields_}

    @property
    def n_bins(self):
        return self.info["n_bins"]

    def set_vertex_state(self, kelvin, fixed=None):
        k = np.ascontiguousarray(kelvin, np.float32)
        f = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
        _check(lib().irtt_scene_set_vertex_state(self._h, _ptr(k), _ptr(f)))

    def set_bin_state(self, kelvin, fixed=None):
        k = np.ascontiguousarray(kelvin, np.float32)
        f = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
        _check(lib().irtt_scene_set_bin_state(self._h, _ptr(k), _ptr(f)))

    def update_vertices(self, xyz):
        x = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        _check(lib().irtt_scene_update_vertices(self._h, _ptr(x)))
        self.xyz = x

    def bin_data(self):
        n = self.n_bins
        a, e, m = np.empty(n), np.empty(n), np.empty(n)
        _check(lib().irtt_scene_get_bin_data(self._h, _ptr(a), _ptr(e), _ptr(m)))
        return a, e, m

    def build_bvh(self):
        _check(lib().irtt_build_bvh(self._h))
        self.info = self.get_info()

    def get_bvh(self):
        n = ctypes.c_uint32()
        d = ctypes.c_uint32()
        _check(lib().irtt_get_bvh(self._h, None, None, None, ctypes.byref(n), ctypes.byref(d)))
        n = n.value
        nodes = np.zeros((max(n - 1, 0), 16), np.uint32)
        tris = np.zeros((n, 12), np.uint32)
        morton = np.zeros(n, np.uint32)
        _check(lib().irtt_get_bvh(self._h, _ptr(nodes) if n > 1 else None, _ptr(tris),
                                  _ptr(morton), None, None))
        return nodes, tris, morton, d.value

    def trace_rays(self, rays, tmin=1e-12, tmax=1e12):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = r.shape[0]
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        b1 = np.empty(n, np.float32)
        b2 = np.empty(n, np.float32)
        _check(lib().irtt_trace_rays(self._h, _ptr(r), n, tmin, tmax, _ptr(prim), _ptr(t),
                                     _ptr(b1), _ptr(b2)))
        return prim, t, b1, b2

    def build_transport(self, params=None, **kw):
        p = params if params is not None else transport_params(**kw)
        st = TransportStats()
        _check(lib().irtt_build_transport(self._h, ctypes.byref(p), ctypes.byref(st)))
        self.info = self.get_info()
        return st.as_dict()

    def transport(self):
        n = self.n_bins
        T = np.empty((n, n), np.float32)
        _check(lib().irtt_get_transport(self._h, _ptr(T)))
        return T

    def transport_fixed(self):
        n = self.n_bins
        acc = np.empty((n, n), np.uint64)
        q = ctypes.c_double()
        _check(lib().irtt_get_transport_fixed(self._h, _ptr(acc), ctypes.byref(q)))
        return acc, q.value

    def transport_sparse(self):
        """The operator in compressed rows: (rowptr, col, val) (irtt_get_transport_sparse)."""
        nnz = ctypes.c_uint64()
        _check(lib().irtt_get_transport_sparse(self._h, ctypes.byref(nnz), None, None, None))
        rp = np.empty(self.n_bins + 1, np.uint32)
        col = np.empty(nnz.value, np.uint32)
        val = np.empty(nnz.value, np.float32)
        _check(lib().irtt_get_transport_sparse(self._h, ctypes.byref(nnz), _ptr(rp), _ptr(col),
                                               _ptr(val)))
        return rp, col, val

    def transport_line(self, b, column=True):
        """T[:, b] (column=True: where b's emission is absorbed) or T[b, :]
        (where the power b absorbs comes from), FP32 (irtt_get_transport_line)."""
        out = np.empty(self.n_bins, np.float32)
        _check(lib().irtt_get_transport_line(self._h, int(b), int(bool(column)), _ptr(out)))
        return out

    def column_sums(self):
        c = np.empty(self.n_bins)
        _check(lib().irtt_get_column_sums(self._h, _ptr(c)))
        return c

    def set_source(self, source):
        """External absorbed power per bin (W), irtt_scene_set_source."""
        src = np.ascontiguousarray(source, np.float64)
        if src.shape != (self.n_bins,):
            raise ValueError("source must have n_bins entries")
        _check(lib().irtt_scene_set_source(self._h, _ptr(src)))

    def select_operator(self, op):
        """OP_DENSE, OP_DENSE_F16 or OP_ACA (dense layout), OP_SPARSE (sparse
        layout) for the solves that follow."""
        _check(lib().irtt_select_operator(self._h, op))

    def simulate(self, steps, params=None, **kw):
        """steps implicit-Euler steps (one solve when steady); returns the
        history (steps, n_bins) and stats (irtt_simulate)."""
        p = params if params is not None else solver_params(**kw)
        H = np.empty((steps, self.n_bins))
        st = SolveStats()
        _check(lib().irtt_simulate(self._h, ctypes.byref(p), steps, _ptr(H), ctypes.byref(st)))
        return H, st.as_dict()

    def adjoint_history(self, history, weights, targets, params=None, gradient=False,
                        edge_ratio=0.05, tiny_eps=1e-2, flags=0, **kw):
        """Backward adjoint over a history (irtt_adjoint_history).  Returns
        (lambda_history, objective, grad_x or None, stats)."""
        p = params if params is not None else solver_params(**kw)
        Hh = np.ascontiguousarray(np.atleast_2d(history), np.float64)
        w = np.ascontiguousarray(weights, np.float64)
        tg = np.ascontiguousarray(targets, np.float64)
        stride = self.n_bins if tg.ndim == 2 else 0
        lam = np.empty_like(Hh)
        obj = ctypes.c_double()
        g = np.empty(3 * self.info["n_vertices"]) if gradient else None
        ap = AdjointParams(edge_ratio, tiny_eps, flags, 0)
        st = SolveStats()
        _check(lib().irtt_adjoint_history(self._h, ctypes.byref(p), Hh.shape[0], _ptr(Hh), _ptr(w),
                                          _ptr(tg), stride, _ptr(lam), ctypes.byref(obj),
                                          ctypes.byref(ap), _ptr(g), ctypes.byref(st)))
        return lam, obj.value, (g.reshape(-1, 3) if gradient else None), st.as_dict()

    def solve_batch(self, sources=None, n_rhs=None, params=None, **kw):
        """Steady solves for the rows of sources (n_rhs, n_bins) in lockstep
        (irtt_solve_batch); returns (kelvin (n_rhs, n_bins), stats)."""
        p = params if params is not None else solver_params(**kw)
        if sources is not None:
            src = np.ascontiguousarray(np.atleast_2d(sources), np.float64)
            R = src.shape[0]
        else:
            src, R = None, n_rhs
        X = np.empty((R, self.n_bins))
        st = SolveStats()
        _check(lib().irtt_solve_batch(self._h, ctypes.byref(p), R, _ptr(src), _ptr(X),
                                      ctypes.byref(st)))
        return X, st.as_dict()

    def adjoint_solve_batch(self, x, rhs, params=None, **kw):
        """lambda_c with J(x_c)^T lambda_c = rhs_c for the rows of x and rhs."""
        p = params if params is not None else solver_params(**kw)
        x = np.ascontiguousarray(np.atleast_2d(x), np.float64)
        r = np.ascontiguousarray(np.atleast_2d(rhs), np.float64)
        lam = np.empty_like(r)
        st = SolveStats()
        _check(lib().irtt_adjoint_solve_batch(self._h, ctypes.byref(p), r.shape[0], _ptr(x),
                                              _ptr(r), _ptr(lam), ctypes.byref(st)))
        return lam, st.as_dict()

    def apply_operator(self, X, transposed=False):
        """T X (or T^T X) for the rows of X (n_rhs, n_bins) with the selected
        operator form (irtt_apply_operator)."""
        X = np.ascontiguousarray(np.atleast_2d(X), np.float64)
        if X.shape[1] != self.n_bins:
            raise ValueError("X must be (n_rhs, n_bins)")
        Y = np.empty_like(X)
        _check(lib().irtt_apply_operator(self._h, X.shape[0], _ptr(X), _ptr(Y), int(transposed)))
        return Y

    def compress(self, algorithm=ACA_PARTIAL, rank=50, column_sum_correction=False, seed=1,
                 precision=STORE_F32):
        """ACA of the operator's photon part; selects OP_ACA (irtt_compress_transport)."""
        p = AcaParams(algorithm, rank, int(bool(column_sum_correction)), seed, precision)
        st = AcaStats()
        _check(lib().irtt_compress_transport(self._h, ctypes.byref(p), ctypes.byref(st)))
        self._aca_rank = st.rank
        return st.as_dict()

    def aca_factors(self):
        """(Ut (k x n), V (k x n), C (n), pivot rows (k), pivot columns (k));
        Ut holds the columns of U."""
        n, k = self.n_bins, self._aca_rank
        U = np.empty((n, k))
        V = np.empty((k, n))
        C = np.empty(n)
        pr = np.empty(k, np.int32)
        pc = np.empty(k, np.int32)
        _check(lib().irtt_get_aca(self._h, _ptr(U), _ptr(V), _ptr(C), _ptr(pr), _ptr(pc)))
        return np.ascontiguousarray(U.T), V, C, pr, pc

    def compress_hierarchical(self, tol=1e-4, eta=1.0, max_rank=32, leaf_size=64,
                              precision=STORE_F16, allow_larger=False, keep_factors=False,
                              workspace_mib=0):
        """H-matrix form of the sparse-layout operator; selects OP_HMATRIX
        (irtt_compress_hierarchical)."""
        p = HmatrixParams(tol, eta, max_rank, leaf_size, precision, int(bool(allow_larger)),
                          int(bool(keep_factors)), int(workspace_mib))
        st = HmatrixStats()
        _check(lib().irtt_compress_hierarchical(self._h, ctypes.byref(p), ctypes.byref(st)))
        return st.as_dict()

    def hmatrix_blocks(self):
        """(perm, blocks (n_far x 6: r0 r1 c0 c1 rank status), err_rel,
        pivots (n_far x 64: 32 rows then 32 columns, -1 past the rank))."""
        nf = ctypes.c_uint32(0)
        _check(lib().irtt_get_hmatrix(self._h, None, ctypes.byref(nf), None, None, None))
        perm = np.empty(self.n_bins, np.uint32)
        blocks = np.empty((nf.value, 6), np.uint32)
        err = np.empty(nf.value)
        piv = np.empty((nf.value, 64), np.int32)
        _check(lib().irtt_get_hmatrix(self._h, _ptr(perm), ctypes.byref(nf), _ptr(blocks),
                                      _ptr(err), _ptr(piv)))
        return perm, blocks, err, piv

    def hmatrix_factors(self, b, blocks):
        """FP64 factors of far block b as built: (U rank x m, V rank x n)."""
        r0, r1, c0, c1, k, _ = (int(x) for x in blocks[b])
        U = np.zeros((k, r1 - r0))
        V = np.zeros((k, c1 - c0))
        _check(lib().irtt_get_hmatrix_factors(self._h, b, _ptr(U), _ptr(V)))
        return U, V

    def solve(self, params=None, **kw):
        p = params if params is not None else solver_params(**kw)
        st = SolveStats()
        _check(lib().irtt_solve(self._h, ctypes.byref(p), ctypes.byref(st)))
        return st.as_dict()

    def temperatures(self):
        x = np.empty(self.n_bins)
        _check(lib().irtt_get_temperatures(self._h, _ptr(x)))
        return x

    def set_temperatures(self, x):
        x = np.ascontiguousarray(x, np.float64)
        _check(lib().irtt_set_temperatures(self._h, _ptr(x)))

    def residual(self, x, params=None, **kw):
        p = params if params is not None else solver_params(**kw)
        x = np.ascontiguousarray(x, np.float64)
        r = np.empty(self.n_bins)
        _check(lib().irtt_residual(self._h, ctypes.byref(p), _ptr(x), _ptr(r)))
        return r

    def adjoint_solve(self, rhs, x=None, params=None, **kw):
        """lambda with J(x)^T lambda = rhs; x = None linearises at the current state."""
        p = params if params is not None else solver_params(**kw)
        r = np.ascontiguousarray(rhs, np.float64)
        xx = None if x is None else np.ascontiguousarray(x, np.float64)
        lam = np.empty(self.n_bins)
        st = SolveStats()
        _check(lib().irtt_adjoint_solve(self._h, ctypes.byref(p), _ptr(xx), _ptr(r), _ptr(lam),
                                        ctypes.byref(st)))
        return lam, st.as_dict()

    def adjoint_trace(self, lam, x4, edge_ratio=0.05, tiny_eps=1e-2,
                      flags=ADJ_AREA | ADJ_CONTINUOUS | ADJ_EDGE):
        """Vertex gradient (n_vertices, 3) = -sum_h lam_h^T (dT/dv) x4_h over
        the history rows of lam / x4 (shape (n_hist, n_bins) or (n_bins,));
        x4 is the emission factor (x^4 for temperatures x)."""
        lam = np.ascontiguousarray(np.atleast_2d(lam), np.float64)
        x = np.ascontiguousarray(np.atleast_2d(x4), np.float64)
        if lam.shape != x.shape or lam.shape[1] != self.n_bins:
            raise ValueError("lambda and x4 must be (n_hist, n_bins)")
        g = np.empty(3 * self.info["n_vertices"])
        ap = AdjointParams(edge_ratio, tiny_eps, flags, 0)
        st = AdjointStats()
        _check(lib().irtt_adjoint_trace(self._h, ctypes.byref(ap), lam.shape[0], _ptr(lam), _ptr(x),
                                        _ptr(g), ctypes.byref(st)))
        return g.reshape(-1, 3), st.as_dict()

    def source_gradient(self, lam):
        """lambda^T db/dvertex (n_vertices, 3) of the material-flux source
        (irtt_source_gradient)."""
        lam = np.ascontiguousarray(lam, np.float64)
        if lam.shape != (self.n_bins,):
            raise ValueError("lambda must have n_bins entries")
        g = np.zeros(3 * self.info["n_vertices"])
        _check(lib().irtt_source_gradient(self._h, _ptr(lam), _ptr(g)))
        return g.reshape(-1, 3)

    def design_gradient(self, dLdx, params=None, solver=None, edge_ratio=0.05, tiny_eps=1e-2,
                        flags=ADJ_AREA | ADJ_CONTINUOUS | ADJ_EDGE):
        """Gradient at the current solved state (irtt_design_gradient).
        params: (vertex, param, dxdp, n_params) as for project_gradient.
        Returns (dp or None, grad_x (n_vertices, 3), lambda, stats)."""
        sp = solver if solver is not None else solver_params()
        g = np.ascontiguousarray(dLdx, np.float64)
        if g.shape != (self.n_bins,):
            raise ValueError("dLdx must have n_bins entries")
        gx = np.empty(3 * self.info["n_vertices"])
        lam = np.empty(self.n_bins)
        ap = AdjointParams(edge_ratio, tiny_eps, flags, 0)
        st = SolveStats()
        dp = None
        args = (0, None, None, None, 0, None)
        if params is not None:
            v, p, d, n_params = params
            v = np.ascontiguousarray(v, np.uint32)
            p = np.ascontiguousarray(p, np.uint32)
            d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
            dp = np.zeros(n_params)
            args = (v.size, _ptr(v), _ptr(p), _ptr(d), n_params, _ptr(dp))
        _check(lib().irtt_design_gradient(self._h, ctypes.byref(sp), _ptr(g), ctypes.byref(ap),
                                          *args, _ptr(gx), _ptr(lam), ctypes.byref(st)))
        return dp, gx.reshape(-1, 3), lam, st.as_dict()


REQ_QUEUED = 0
REQ_RUNNING = 1
REQ_DONE = 2


class Server:
    """Request loop over design variants (irtt_serve_*): each request is a
    scene + transport + solve (+ adjoint and dL/dp when a target is given);
    requests run concurrently on the device as slots and HBM allow."""

    def __init__(self, slots=4, memory_budget=0):
        h = ctypes.c_void_p()
        cfg = ServeConfig(slots, 0, memory_budget)
        _check(lib().irtt_serve_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, sd, bin_mode=BIN_VERTEX, operator=OP_DENSE, transport=None, solver=None,
               kelvin0=None, fixed=None, target=None, objective_weight=1.0, params=None,
               adjoint_flags=0, source=None, group=0):
        """sd: SceneData.  params: (vertex, param, dxdp, n_params) for dL/dp.
        adjoint_flags: ADJ_* terms of the trace (0: area | continuous | edge).
        source: per-bin external power replacing the materials' flux.  group:
        nonzero to share one scene, build and batched solves with queued
        requests of the same group and identical inputs but source and target.
        Returns a ticket; outputs come back from wait()."""
        nb = sd.nv if bin_mode == BIN_VERTEX else sd.nt
        xyz = np.ascontiguousarray(sd.xyz, np.float32)
        tri = np.ascontiguousarray(sd.idx, np.uint32)
        mat = sd.materials()
        arrays = [xyz, tri, mat]
        r = Request()
        r.xyz, r.nv, r.nt, r.tri, r.mat = _ptr(xyz), sd.nv, sd.nt, _ptr(tri), _ptr(mat)
        r.bin_mode, r.operator_kind = bin_mode, operator
        if kelvin0 is not None:
            k0 = np.ascontiguousarray(kelvin0, np.float32)
            arrays.append(k0)
            r.kelvin0 = _ptr(k0)
        if fixed is not None:
            fx = np.ascontiguousarray(fixed, np.uint8)
            arrays.append(fx)
            r.fixed = _ptr(fx)
        r.transport = transport if transport is not None else transport_params()
        r.solver = solver if solver is not None else solver_params()
        r.adjoint = AdjointParams(0.0, 0.0, adjoint_flags, 0)
        r.group = group
        if source is not None:
            src = np.ascontiguousarray(source, np.float64)
            if src.shape != (nb,):
                raise ValueError("source must have n_bins entries")
            arrays.append(src)
            r.source = _ptr(src)
        kelvin = np.zeros(nb)
        objective = np.zeros(1)
        dp = None
        r.kelvin, r.objective = _ptr(kelvin), _ptr(objective)
        if target is not None:
            t = np.ascontiguousarray(np.broadcast_to(target, (nb,)), np.float64)
            arrays.append(t)
            r.target, r.objective_weight = _ptr(t), objective_weight
            if params is not None:
                v, p, d, n_params = params
                v = np.ascontiguousarray(v, np.uint32)
                p = np.ascontiguousarray(p, np.uint32)
                d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
                arrays += [v, p, d]
                dp = np.zeros(n_params)
                r.nnz, r.n_params = v.size, n_params
                r.vertex, r.param, r.dxdp, r.dp = _ptr(v), _ptr(p), _ptr(d), _ptr(dp)
        tk = ctypes.c_uint64()
        _check(lib().irtt_serve_submit(self._h, ctypes.byref(r), ctypes.byref(tk)))
        self._keep[tk.value] = {"kelvin": kelvin, "dp": dp, "objective": objective}
        return tk.value

    def poll(self, ticket):
        st = RequestStatus()
        _check(lib().irtt_serve_poll(self._h, ticket, ctypes.byref(st)))
        return st.as_dict()

    def wait(self, ticket):
        """(status dict, {"kelvin", "dp", "objective"}); raises if the request failed."""
        st = RequestStatus()
        _check(lib().irtt_serve_wait(self._h, ticket, ctypes.byref(st)))
        if st.state < 0:
            raise IrttError(st.state, f"request {ticket} failed with status {st.state}")
        out = self._keep.pop(ticket)
        out = {"kelvin": out["kelvin"], "dp": out["dp"], "objective": float(out["objective"][0])}
        return st.as_dict(), out

    def close(self):
        if self._h:
            _check(lib().irtt_serve_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
