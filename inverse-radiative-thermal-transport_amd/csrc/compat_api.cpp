// compat_api.cpp -- the reference's external C ABI (include/thermal_api_compat.h)
// on top of the irtt_* entry points.
//
// Restates thermal_api.cpp:47-536 (object/sky ingest, sky re-projection),
// ThermalRenderer::thermalTimestep / getKelvin / getValue / setKelvin
// (thermal_renderer.cpp:402-660), ThermalData::SkyData (thermal_data.hpp:19-176)
// and sky_value_to_kelvin (thermal_common.cpp:12-18).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/thermal_api_compat.h"
#include "irtt_internal.hpp"

static_assert(sizeof(ObjectProperties) == 52, "ObjectProperties must match the reference layout");

namespace {

constexpr float kSigmaRaw = 5.670374419e-8f;   // thermal_common.hpp:39
constexpr float kMinimalSkyKelvin = 111.575f;   // thermal_data.hpp:29

struct Obj {
    std::vector<float> xyz;       // 3 per vertex
    std::vector<uint32_t> tri;    // local indices, 3 per triangle
    ObjectProperties p;
};

struct State {
    bool loaded = false;
    std::vector<Obj> objs;
    std::vector<float> sky_xyz;
    std::vector<uint32_t> sky_tri;
    std::vector<float> sky_values;
    float aabb_min[3] = {0, 0, 0}, aabb_max[3] = {0, 0, 0};
    float sphere[4] = {0, 0, 0, 0};
    uint32_t threads_per_tri = 32768, rays_per_thread = 8;  // thermal_common.hpp:93-94
    float sky_min_kelvin = 0.0f;
    bool steady = true;                                      // thermal_solver.hpp:36
    double step_s = 300.0;                                   // thermal_solver.hpp:35
    double sim_time_s = 0.0;
    irtt_scene scene = nullptr;
    bool transport_built = false;
    uint32_t nv = 0, sky_vertex_offset = 0, sky_vertex_count = 0;
    std::vector<float> kelvin0;
    std::vector<uint8_t> fixed;
    std::vector<double> vertex_area;
};

State g;

// sky_value_to_kelvin, thermal_common.cpp:12-18 (value in W/m^2)
float sky_value_to_kelvin(float value, float min_k)
{
    float k = (float)std::pow((double)(value / kSigmaRaw), 1.0 / 4.0);
    return std::max(k, min_k);
}

void destroy_scene()
{
    if (g.scene) irtt_scene_destroy(g.scene);
    g.scene = nullptr;
    g.transport_built = false;
}

// ThermalRenderer::computeSceneAABB / computeSceneBoundingSphere,
// thermal_renderer.cpp:173-235: traceable objects only; the max corner starts
// at FLT_MIN as in the reference
void compute_scene_bounds()
{
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {FLT_MIN, FLT_MIN, FLT_MIN};
    for (const Obj &o : g.objs) {
        if (!o.p.traceable) continue;
        for (size_t v = 0; v + 2 < o.xyz.size(); v += 3)
            for (int k = 0; k < 3; k++) {
                mn[k] = std::min(mn[k], o.xyz[v + k]);
                mx[k] = std::max(mx[k], o.xyz[v + k]);
            }
    }
    std::memcpy(g.aabb_min, mn, sizeof(mn));
    std::memcpy(g.aabb_max, mx, sizeof(mx));
    float h[3] = {(mx[0] - mn[0]) * 0.5f, (mx[1] - mn[1]) * 0.5f, (mx[2] - mn[2]) * 0.5f};
    float len = std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    if (len > 0) {
        for (int k = 0; k < 3; k++) g.sphere[k] = mn[k] + h[k];
        g.sphere[3] = len;
    } else {
        g.sphere[0] = g.sphere[1] = g.sphere[2] = g.sphere[3] = 0.0f;
    }
}

int fail(int code, const std::string &msg)
{
    irtt::set_last_error(msg);
    return code;
}

// Builds the irtt scene from the objects and the sky (vertex bins, object
// order then sky, as ThermalData::load numbers them).
int build_scene()
{
    destroy_scene();
    std::vector<float> xyz;
    std::vector<uint32_t> tri;
    std::vector<irtt_material> mat;
    std::vector<float> kelvin;
    std::vector<uint8_t> fixed;
    uint32_t voff = 0;
    for (const Obj &o : g.objs) {
        uint32_t nv = (uint32_t)(o.xyz.size() / 3);
        xyz.insert(xyz.end(), o.xyz.begin(), o.xyz.end());
        irtt_material m{};
        m.diffuse_reflectance = o.p.diffuse_reflectance;
        m.specular_reflectance = o.p.specular_reflectance;
        m.thickness = o.p.thickness;
        m.density = o.p.density;
        m.heat_capacity = o.p.heat_capacity;
        m.heat_conductivity = o.p.heat_conductivity;
        m.heat_flux = 0.0f;
        m.flags = (o.p.diffuse_emission ? IRTT_MAT_DIFFUSE_EMISSION : 0u) |
                  (o.p.traceable ? IRTT_MAT_TRACEABLE : 0u);
        for (size_t t = 0; t < o.tri.size(); t++) tri.push_back(o.tri[t] + voff);
        mat.insert(mat.end(), o.tri.size() / 3, m);
        kelvin.insert(kelvin.end(), nv, o.p.kelvin);
        fixed.insert(fixed.end(), nv, o.p.temperature_fixed ? 1 : 0);
        voff += nv;
    }
    g.sky_vertex_offset = voff;
    g.sky_vertex_count = (uint32_t)(g.sky_xyz.size() / 3);
    if (g.sky_vertex_count) {
        xyz.insert(xyz.end(), g.sky_xyz.begin(), g.sky_xyz.end());
        irtt_material m{};  // thermal_api.cpp:279-291
        m.thickness = 1.0f;
        m.density = 1.0f;
        m.heat_capacity = 1.0f;
        for (size_t t = 0; t < g.sky_tri.size(); t++) tri.push_back(g.sky_tri[t] + voff);
        mat.insert(mat.end(), g.sky_tri.size() / 3, m);
        kelvin.insert(kelvin.end(), g.sky_vertex_count, 0.0f);
        fixed.insert(fixed.end(), g.sky_vertex_count, 1);
        // initial sky temperatures, ThermalData::load -> SkyData::setHourByTime
        // (averaged, clipped at MINIMAL_KELVIN), thermal_data.cpp:104-143
        for (size_t q = 0; q < g.sky_values.size(); q++) {
            float kk = sky_value_to_kelvin(g.sky_values[q], kMinimalSkyKelvin);
            for (int c = 0; c < 6; c++) kelvin[voff + g.sky_tri[6 * q + c]] = kk;
        }
    }
    if (tri.empty()) return fail(IRTT_ERR_STATE, "no triangles loaded");
    g.nv = (uint32_t)(xyz.size() / 3);
    int rc = irtt_scene_create(xyz.data(), g.nv, tri.data(), (uint32_t)(tri.size() / 3), mat.data(),
                               IRTT_BIN_VERTEX, &g.scene);
    if (rc) return fail(rc, irtt_last_error());
    g.kelvin0 = kelvin;
    g.fixed = fixed;
    rc = irtt_scene_set_vertex_state(g.scene, g.kelvin0.data(), g.fixed.data());
    if (rc) return fail(rc, irtt_last_error());
    g.vertex_area.assign(g.nv, 0.0);
    rc = irtt_scene_get_bin_data(g.scene, g.vertex_area.data(), nullptr, nullptr);
    if (rc) return fail(rc, irtt_last_error());
    g.sim_time_s = 0.0;
    return 0;
}

}  // namespace

extern "C" {

int load(bool)
{
    char name[128];
    int cus = 0;
    int rc = irtt_device_info(name, sizeof(name), &cus);
    if (rc) return fail(rc, irtt_last_error());
    g = State{};
    g.loaded = true;
    return 0;
}

int unload(void)
{
    destroy_scene();
    g = State{};
    return 0;
}

int test_load_scene(void)
{
    // thermal_api.cpp:47-85: one quad, two triangles, default properties
    Obj o;
    o.xyz = {-1, 1, 0, 1, 1, 0, 1, -1, 0, -1, -1, 0};
    o.tri = {0, 1, 3, 1, 2, 3};
    std::memset(&o.p, 0, sizeof(o.p));
    g.objs.push_back(o);
    return 0;
}

int test_lib(int _value) { return _value * 2; }

int test_fill_array(float *_arr, int _size)
{
    if (!_arr && _size > 0) return fail(IRTT_ERR_ARG, "null array");
    for (int i = 0; i < _size; i++) _arr[i] = _arr[i] * float(i * 3);
    return 0;
}

int load_geometry(float *_vertices, float *_vertex_colors, unsigned int _total_vertex_count,
                  unsigned int *_indices, unsigned int _total_indices_count,
                  ObjectProperties *_object_properties, unsigned int _object_count)
{
    (void)_vertex_colors;
    (void)_total_indices_count;
    if (!g.loaded) return fail(IRTT_ERR_STATE, "load() has not been called");
    if (_object_count && (!_vertices || !_indices || !_object_properties))
        return fail(IRTT_ERR_ARG, "null pointer");
    for (unsigned int oi = 0; oi < _object_count; oi++) {
        const ObjectProperties &p = _object_properties[oi];
        uint64_t vs = (uint64_t)p.vertex_offset * 3, ve = vs + (uint64_t)p.vertex_count * 3;
        if (ve > _total_vertex_count)
            return fail(IRTT_ERR_ARG, "object " + std::to_string(oi) + " vertex range exceeds the array");
        Obj o;
        o.p = p;
        o.xyz.assign(_vertices + vs, _vertices + ve);
        uint64_t is = (uint64_t)p.indices_offset * 3, ie = is + (uint64_t)p.indices_count * 3;
        o.tri.assign(_indices + is, _indices + ie);
        for (uint32_t v : o.tri)
            if (v >= p.vertex_count)
                return fail(IRTT_ERR_ARG, "object " + std::to_string(oi) + " index out of range");
        g.objs.push_back(std::move(o));
    }
    compute_scene_bounds();
    return 0;
}

int load_sky(float *_vertices, unsigned int _vertex_count, unsigned int *_indices, float *_values,
             unsigned int _quad_count)
{
    if (!g.loaded) return fail(IRTT_ERR_STATE, "load() has not been called");
    if (_quad_count && (!_vertices || !_indices || !_values)) return fail(IRTT_ERR_ARG, "null pointer");
    std::vector<float> v(_vertices, _vertices + 3ull * _vertex_count);
    std::vector<uint32_t> tri;
    tri.reserve(6ull * _quad_count);
    const float cx = g.sphere[0], cy = g.sphere[1], cz = g.sphere[2], r = g.sphere[3];
    const float scaling = std::sqrt(r * r + r * r) * 1.2f;
    for (unsigned int q = 0; q < _quad_count; q++) {
        uint32_t ia = _indices[4 * q], ib = _indices[4 * q + 1], ic = _indices[4 * q + 2],
                 id = _indices[4 * q + 3];
        if (ia >= _vertex_count || ib >= _vertex_count || ic >= _vertex_count || id >= _vertex_count)
            return fail(IRTT_ERR_ARG, "sky quad index out of range");
        // split (c,b,d), (b,a,d), thermal_api.cpp:206-214
        tri.insert(tri.end(), {ic, ib, id, ib, ia, id});
        // re-project the quad's corners around the scene bounding sphere,
        // in quad order (shared corners move again), thermal_api.cpp:216-248
        auto P = [&](uint32_t i, int k) { return v[3 * i + k]; };
        float qc[3], ab[3], ac[3], ad[3];
        for (int k = 0; k < 3; k++) {
            qc[k] = (P(ia, k) + P(ib, k) + P(ic, k) + P(id, k)) * 0.25f;
            ab[k] = P(ib, k) - P(ia, k);
            ac[k] = P(ic, k) - P(ia, k);
            ad[k] = P(id, k) - P(ia, k);
        }
        float n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2],
                      ab[0] * ac[1] - ab[1] * ac[0]};
        float nl = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; k++) n[k] /= nl;
        float cs[3] = {cx - qc[0], cy - qc[1], cz - qc[2]};
        float pd = n[0] * cs[0] + n[1] * cs[1] + n[2] * cs[2];
        float proj[3] = {cx - n[0] * pd, cy - n[1] * pd, cz - n[2] * pd};
        for (uint32_t ind : {ia, ib, ic, id}) {
            float d[3] = {P(ind, 0) - qc[0], P(ind, 1) - qc[1], P(ind, 2) - qc[2]};
            float dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            for (int k = 0; k < 3; k++) v[3 * ind + k] = proj[k] + d[k] / dl * scaling;
        }
    }
    g.sky_xyz = v;
    g.sky_tri = tri;
    g.sky_values.assign(_values, _values + _quad_count);
    // loading the sky loads the scene and resets the simulation, thermal_api.cpp:433-440
    return build_scene();
}

int unload_scene(void)
{
    destroy_scene();
    g.objs.clear();
    g.sky_xyz.clear();
    g.sky_tri.clear();
    g.sky_values.clear();
    return 0;
}

int update_sky_values(float *_values, unsigned int _quad_count)
{
    if (!g.scene) return fail(IRTT_ERR_STATE, "no scene loaded");
    if (_quad_count > g.sky_tri.size() / 6) return fail(IRTT_ERR_ARG, "more values than sky quads");
    if (_quad_count && !_values) return fail(IRTT_ERR_ARG, "null pointer");
    // ThermalRenderer::setKelvin(sky), thermal_renderer.cpp:429-497
    std::vector<double> x(g.nv);
    int rc = irtt_get_temperatures(g.scene, x.data());
    if (rc) return fail(rc, irtt_last_error());
    for (unsigned int q = 0; q < _quad_count; q++) {
        float kk = sky_value_to_kelvin(_values[q], g.sky_min_kelvin);
        g.sky_values[q] = _values[q];
        for (int c = 0; c < 6; c++) g.kelvin0[g.sky_vertex_offset + g.sky_tri[6 * q + c]] = kk;
    }
    rc = irtt_scene_set_vertex_state(g.scene, g.kelvin0.data(), g.fixed.data());
    if (rc) return fail(rc, irtt_last_error());
    for (uint32_t i = 0; i < g.sky_vertex_count; i++)
        x[g.sky_vertex_offset + i] = g.kelvin0[g.sky_vertex_offset + i];
    rc = irtt_set_temperatures(g.scene, x.data());
    return rc ? fail(rc, irtt_last_error()) : 0;
}

int reset_simulation(void)
{
    if (!g.scene) return 0;
    g.sim_time_s = 0.0;
    int rc = irtt_scene_set_vertex_state(g.scene, g.kelvin0.data(), g.fixed.data());
    return rc ? fail(rc, irtt_last_error()) : 0;
}

int set_ray_batch_count(unsigned int _ray_count, unsigned int _batch_count)
{
    if (_ray_count == 0 || _batch_count == 0) return fail(IRTT_ERR_ARG, "counts must be positive");
    g.threads_per_tri = _ray_count;
    g.rays_per_thread = _batch_count;
    g.transport_built = false;
    return 0;
}

int set_minimum_sky_kelvin(float _min)
{
    g.sky_min_kelvin = _min;
    return 0;
}

int set_steady_state(bool _enabled)
{
    g.steady = _enabled;
    return 0;
}

int simulate(float _step_size_hours, unsigned int _time_step_count, float *_time_hours)
{
    if (!g.scene) return fail(IRTT_ERR_STATE, "no scene loaded (load_sky loads it)");
    if (!_time_hours) return fail(IRTT_ERR_ARG, "null time pointer");
    g.step_s = (double)_step_size_hours * 3600.0;   // ThermalRenderer::setTimeStep
    g.sim_time_s = (double)(*_time_hours) * 3600.0; // ThermalRenderer::setTime
    if (!g.transport_built) {
        irtt_transport_params tp{};
        tp.threads_per_tri = g.threads_per_tri;
        tp.rays_per_thread = g.rays_per_thread;
        tp.batch_seed = 1234;  // constRandSeed / randSeed, thermal_common.hpp:126-127
        tp.max_depth = 0;
        tp.tiny_dist = 1e-3f;
        // the reference's dense matrix overflows its indexing above 46 340
        // bins (thermal_transport.cpp:143); larger scenes use the sparse
        // layout, whose entries equal the dense ones
        irtt_scene_info info{};
        if (irtt_scene_info_get(g.scene, &info) == IRTT_OK && info.n_bins > 46340u)
            tp.layout = IRTT_LAYOUT_SPARSE;
        int rc = irtt_build_transport(g.scene, &tp, nullptr);
        if (rc) return fail(rc, irtt_last_error());
        g.transport_built = true;
    }
    irtt_solver_params sp{};
    sp.steady = g.steady ? 1u : 0u;
    sp.jacobian = IRTT_JAC_EXACT;
    sp.max_newton = 100;
    sp.max_linear = 400;
    sp.dt = g.step_s > 0 ? g.step_s : 300.0;
    sp.ambient_kelvin = -1.0;
    sp.x_tol = 1e-5;
    sp.f_tol = 1e-12;
    sp.lin_rtol = 1e-10;
    for (unsigned int i = 0; i < _time_step_count; i++) {
        int rc = irtt_solve(g.scene, &sp, nullptr);
        if (rc) return fail(rc, irtt_last_error());
        if (!g.steady) g.sim_time_s += sp.dt;
    }
    *_time_hours = (float)(g.sim_time_s / 3600.0);
    return 0;
}

int get_vertex_values(float *_vertex_temperatures, unsigned int _total_vertex_count, unsigned int _type)
{
    if (!g.scene) return fail(IRTT_ERR_STATE, "no scene loaded");
    if (_total_vertex_count && !_vertex_temperatures) return fail(IRTT_ERR_ARG, "null pointer");
    if (_total_vertex_count / 3 > g.nv) return fail(IRTT_ERR_ARG, "count exceeds the vertex count");
    std::vector<double> x(g.nv);
    int rc = irtt_get_temperatures(g.scene, x.data());
    if (rc) return fail(rc, irtt_last_error());
    for (unsigned int i = 0; i + 2 < _total_vertex_count; i += 3) {
        float val = (float)x[i / 3];
        if (_type == 1)
            val = (float)(kSigmaRaw * g.vertex_area[i / 3] * std::pow((double)val, 4.0));
        _vertex_temperatures[i] = val;
        _vertex_temperatures[i + 1] = val;
        _vertex_temperatures[i + 2] = val;
    }
    return 0;
}

int get_vertex_temperatures(float *_vertex_temperatures, unsigned int _total_vertex_count)
{
    return get_vertex_values(_vertex_temperatures, _total_vertex_count, 0);
}

}  // extern "C"
