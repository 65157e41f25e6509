"""The reference's external C ABI (thermal_api.hpp:10-62) served by libirtt.so:
results equal the irtt_* path on the same scene, sky values map to kelvin as
thermal_common.cpp:12-18 does, time stepping advances the clock in hours."""
import ctypes

import numpy as np
import pytest

from test_abi import ObjectProperties

pytestmark = pytest.mark.gpu

SIGMA = 5.670374419e-8


def _objects():
    from irtt import scenes
    sd = scenes.parallel_plates(n=6, emissivity=0.9)
    na, nva = sd.meta["plate_a_tris"], sd.meta["plate_a_verts"]
    props = (ObjectProperties * 2)()
    for i, (vo, vc, io, ic, k, fixed) in enumerate(
            [(0, nva, 0, na, 400.0, True), (nva, sd.nv - nva, na, sd.nt - na, 0.0, False)]):
        p = props[i]
        p.vertex_offset, p.vertex_count, p.indices_offset, p.indices_count = vo, vc, io, ic
        p.kelvin, p.temperature_fixed = k, fixed
        p.thickness, p.density, p.heat_capacity, p.heat_conductivity = 0.01, 2000.0, 900.0, 0.0
        p.diffuse_reflectance, p.specular_reflectance = 0.1, 0.0
        p.diffuse_emission, p.traceable = True, True
    local = sd.idx.copy()
    local[na:] -= nva
    return sd, props, local.astype(np.uint32)


def _lib(irtt):
    L = ctypes.CDLL(irtt.LIB_PATH)
    L.load.argtypes = [ctypes.c_bool]
    L.simulate.argtypes = [ctypes.c_float, ctypes.c_uint, ctypes.POINTER(ctypes.c_float)]
    L.set_minimum_sky_kelvin.argtypes = [ctypes.c_float]
    L.set_steady_state.argtypes = [ctypes.c_bool]
    return L


def _vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_compat_steady_equals_direct_path(irtt):
    L = _lib(irtt)
    sd, props, local = _objects()
    xyz = np.ascontiguousarray(sd.xyz.ravel(), np.float32)
    rgb = np.zeros_like(xyz)
    assert L.load(False) == 0
    assert L.load_geometry(_vp(xyz), _vp(rgb), xyz.size, _vp(local), local.size, props, 2) == 0
    empty = np.zeros(3, np.float32)
    eidx = np.zeros(4, np.uint32)
    assert L.load_sky(_vp(empty), 0, _vp(eidx), _vp(empty), 0) == 0
    assert L.set_ray_batch_count(64, 8) == 0
    assert L.set_steady_state(True) == 0
    t = ctypes.c_float(5.0)
    assert L.simulate(1.0, 1, ctypes.byref(t)) == 0, irtt.lib().irtt_last_error()
    assert t.value == 5.0  # steady state does not advance the clock
    out = np.empty(3 * sd.nv, np.float32)
    assert L.get_vertex_temperatures(_vp(out), out.size) == 0
    kel = out[0::3]
    assert np.array_equal(out[0::3], out[1::3]) and np.array_equal(out[0::3], out[2::3])
    W = np.empty(3 * sd.nv, np.float32)
    assert L.get_vertex_values(_vp(W), W.size, 1) == 0
    L.unload()
    # the same scene through irtt_*
    mats = sd.materials()
    mats["thickness"], mats["density"], mats["heat_capacity"] = 0.01, 2000.0, 900.0
    mats["heat_flux"] = 0.0
    nva = sd.meta["plate_a_verts"]
    kelvin = np.zeros(sd.nv, np.float32)
    kelvin[:nva] = 400.0
    fixed = np.zeros(sd.nv, np.uint8)
    fixed[:nva] = 1
    with irtt.Scene(sd.xyz, sd.idx, mats, irtt.BIN_VERTEX) as s:
        s.set_vertex_state(kelvin, fixed)
        s.build_transport(threads_per_tri=64, rays_per_thread=8, batch_seed=1234)
        s.solve(irtt.solver_params(ambient_kelvin=-1.0, x_tol=1e-5, f_tol=1e-12, max_linear=400))
        x = s.temperatures()
        area, _, _ = s.bin_data()
    assert np.array_equal(kel, x.astype(np.float32))
    assert np.all(kel[:nva] == 400.0)
    assert np.allclose(W[0::3], SIGMA * area * x ** 4, rtol=1e-5)
    print(f"plate B mean {kel[nva:].mean():.3f} K")


def test_compat_sky_and_transient_clock(irtt):
    L = _lib(irtt)
    sd, props, local = _objects()
    xyz = np.ascontiguousarray(sd.xyz.ravel(), np.float32)
    assert L.load(False) == 0
    assert L.load_geometry(_vp(xyz), _vp(np.zeros_like(xyz)), xyz.size, _vp(local), local.size,
                           props, 2) == 0
    # four sky quads above the plates
    sv = np.array([[-2, -2, 3], [3, -2, 3], [3, 3, 3], [-2, 3, 3], [0.5, 0.5, 3.5]], np.float32)
    sky_xyz = np.ascontiguousarray(np.concatenate([sv, sv + [0, 0, 1]]).ravel(), np.float32)
    quads = np.array([0, 1, 2, 3, 5, 6, 7, 8, 0, 1, 6, 5, 2, 3, 8, 7], np.uint32)
    values = np.array([300.0, 0.0, 50.0, 400.0], np.float32)
    assert L.load_sky(_vp(sky_xyz), 10, _vp(quads), _vp(values), 4) == 0, \
        irtt.lib().irtt_last_error()
    nv_all = sd.nv + 10
    out = np.empty(3 * nv_all, np.float32)
    assert L.get_vertex_temperatures(_vp(out), out.size) == 0
    used = np.unique(quads)  # sky vertices 4 and 9 belong to no quad and stay at 0 K
    sky_k = out[0::3][sd.nv:][used]
    want = {max((v / SIGMA) ** 0.25, 111.575) for v in values}
    assert set(np.round(sky_k, 2)) <= {round(w, 2) for w in want}
    # update_sky_values uses the minimum set by set_minimum_sky_kelvin
    assert L.set_minimum_sky_kelvin(150.0) == 0
    assert L.update_sky_values(_vp(values), 4) == 0
    assert L.get_vertex_temperatures(_vp(out), out.size) == 0
    assert out[0::3][sd.nv:][used].min() >= 150.0 - 1e-3
    # transient stepping advances the clock by steps * hours
    assert L.set_ray_batch_count(32, 4) == 0
    assert L.set_steady_state(False) == 0
    t = ctypes.c_float(2.0)
    assert L.simulate(0.25, 4, ctypes.byref(t)) == 0, irtt.lib().irtt_last_error()
    assert abs(t.value - 3.0) < 1e-6
    assert L.get_vertex_temperatures(_vp(out), 3 * sd.nv) == 0
    assert np.all(np.isfinite(out[: 3 * sd.nv]))
    assert L.reset_simulation() == 0
    assert L.unload_scene() == 0
    assert L.unload() == 0
