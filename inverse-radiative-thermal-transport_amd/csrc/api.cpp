// api.cpp -- C ABI entry points of libirtt.so (declared in include/irtt.h)
// and the host-side scene assembly.
//
// Scene assembly restates ThermalScene::initObject (thermal_scene.cpp:76-108),
// ThermalData::updateAreaVectorsForMesh / load / computeLumpedThermalMass /
// computeConductiveLaplacian (thermal_data.cpp:56-81,83-173,216-313) for
// per-triangle materials and the two bin modes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>

#include <memory>
#include <mutex>
#include <unordered_map>
#include <new>
#include <string>

#include "irtt_internal.hpp"

using namespace irtt;

namespace {

thread_local std::string g_err;

template <class F>
int guard(F &&f)
{
    try {
        f();
        return IRTT_OK;
    } catch (const Error &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_err = "host allocation failed";
        return IRTT_ERR_NOMEM;
    } catch (const std::exception &e) {
        g_err = e.what();
        return IRTT_ERR_STATE;
    }
}

Scene *S(irtt_scene h)
{
    if (!h) throw Error(IRTT_ERR_ARG, "null scene handle");
    Scene *s = reinterpret_cast<Scene *>(h);
    t_alloc_stream = s->stream;
    return s;
}

// Scene buffers come from the device's stream-ordered pool (hipMallocAsync on
// the scene's stream).  By default the pool hands freed memory back to the
// driver at every synchronisation, so a request that creates, builds and
// destroys a scene maps and unmaps all its buffers again; the pool is told to
// keep up to kPoolKeep bytes of freed memory for reuse instead.  Once per
// device.
#ifndef IRTT_POOL_KEEP_GIB
#define IRTT_POOL_KEEP_GIB 8
#endif
constexpr uint64_t kPoolKeep = (uint64_t)IRTT_POOL_KEEP_GIB << 30;

void configure_pool()
{
    static std::mutex mu;
    static std::vector<int> done;
    int dev = 0;
    IRTT_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if (kPoolKeep == 0 || std::find(done.begin(), done.end(), dev) != done.end()) return;
    done.push_back(dev);
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess || !pool) {
        (void)hipGetLastError();  // no pool: allocations fall back to hipMalloc
        return;
    }
    uint64_t keep = kPoolKeep;
    if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess)
        (void)hipGetLastError();
}

// cotangent of the angle between a and b, thermal_data.cpp:253-258; a
// degenerate corner contributes nothing instead of inf/NaN
double cotan(const double a[3], const double b[3])
{
    double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    double bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    double den = aa * bb - ab * ab;
    if (!(den > 0.0)) return 0.0;
    return ab / std::sqrt(den);
}

// Cotangent Laplacian scaled by conductivity * thickness (thermal_data.cpp:259-313)
// into s.lap (CSR, columns ascending).  Each triangle adds nine terms to its
// vertices' rows.  The terms are bucketed by row in triangle order (counting
// sort), sorted by column within a row (stable) and summed in that order from
// 0.0, which gives every entry the same sum as accumulating the triangles one
// after another.
static void build_laplacian(Scene &s)
{
    const uint32_t nt = s.nt, nv = s.nv;
    std::vector<uint32_t> start(nv + 1, 0);
    for (uint32_t t = 0; t < nt; t++) {
        if ((double)s.mat[t].heat_conductivity * (double)s.mat[t].thickness == 0.0) continue;
        for (int c = 0; c < 3; c++) start[s.idx[3 * t + c] + 1] += 3;
    }
    for (uint32_t i = 0; i < nv; i++) start[i + 1] += start[i];
    std::vector<std::pair<uint32_t, double>> ent(start[nv]);
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (uint32_t t = 0; t < nt; t++) {
        const irtt_material &m = s.mat[t];
        double k = (double)m.heat_conductivity * (double)m.thickness;
        if (k == 0.0) continue;
        const uint32_t v[3] = {s.idx[3 * t], s.idx[3 * t + 1], s.idx[3 * t + 2]};
        double a[3], b[3], c[3];
        for (int q = 0; q < 3; q++) {
            a[q] = s.xyz[3 * v[0] + q];
            b[q] = s.xyz[3 * v[1] + q];
            c[q] = s.xyz[3 * v[2] + q];
        }
        double ac[3], bc[3], ba[3], ca[3], cb[3], ab[3];
        for (int q = 0; q < 3; q++) {
            ac[q] = a[q] - c[q]; bc[q] = b[q] - c[q];
            ba[q] = b[q] - a[q]; ca[q] = c[q] - a[q];
            cb[q] = c[q] - b[q]; ab[q] = a[q] - b[q];
        }
        const double cot1 = cotan(ac, bc), cot2 = cotan(ba, ca), cot3 = cotan(cb, ab);
        // w[r][c]: the term of vertex r's row in vertex c's column
        const double w[3][3] = {{k * 0.5 * (cot1 + cot3), k * 0.5 * (-cot1), k * 0.5 * (-cot3)},
                                {k * 0.5 * (-cot1), k * 0.5 * (cot1 + cot2), k * 0.5 * (-cot2)},
                                {k * 0.5 * (-cot3), k * 0.5 * (-cot2), k * 0.5 * (cot2 + cot3)}};
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) ent[fill[v[r]]++] = {v[cc], w[r][cc]};
    }
    s.lap.n = nv;
    s.lap.rowptr.assign(nv + 1, 0);
    s.lap.col.clear();
    s.lap.val.clear();
    s.lap.col.reserve(ent.size());
    s.lap.val.reserve(ent.size());
    for (uint32_t i = 0; i < nv; i++) {
        auto b0 = ent.begin() + start[i], b1 = ent.begin() + start[i + 1];
        // insertion sort by column: stable, and a row holds a few dozen terms
        for (auto it = b0 + (b0 != b1); it < b1; ++it) {
            const auto e = *it;
            auto j = it;
            for (; j != b0 && (j - 1)->first > e.first; --j) *j = *(j - 1);
            *j = e;
        }
        for (auto it = b0; it != b1;) {
            const uint32_t col = it->first;
            double sum = 0.0;
            for (; it != b1 && it->first == col; ++it) sum += it->second;
            s.lap.col.push_back(col);
            s.lap.val.push_back(sum);
        }
        s.lap.rowptr[i + 1] = (uint32_t)s.lap.col.size();
    }
}

void assemble(Scene &s)
{
    const uint32_t nt = s.nt, nv = s.nv;
    s.assembly_gen++;
    s.rho_d.resize(nt);
    s.rho_s.resize(nt);
    s.eps.resize(nt);
    s.tri_area.resize(nt);
    s.flags.resize(nt);
    std::vector<double> area_d(nt), area_raw(nt);
    for (uint32_t t = 0; t < nt; t++) {
        const irtt_material &m = s.mat[t];
        float rd = m.diffuse_reflectance, rs = m.specular_reflectance;
        // thermal_scene.cpp:87-97: normalise when total reflectance exceeds 1
        float tot = rd + rs;
        if (tot > 1.0f) {
            rd /= tot;
            rs /= tot;
        }
        s.rho_d[t] = rd;
        s.rho_s[t] = rs;
        s.eps[t] = 1.0f - rd - rs;  // thermal_tracing_unified.glsl:309
        s.flags[t] = (uint8_t)(m.flags & 3u);
        const float *p0 = &s.xyz[3 * s.idx[3 * t]];
        const float *p1 = &s.xyz[3 * s.idx[3 * t + 1]];
        const float *p2 = &s.xyz[3 * s.idx[3 * t + 2]];
        double e1[3], e2[3];
        for (int k = 0; k < 3; k++) {
            e1[k] = (double)p1[k] - (double)p0[k];
            e2[k] = (double)p2[k] - (double)p0[k];
        }
        double cx = e1[1] * e2[2] - e1[2] * e2[1];
        double cy = e1[2] * e2[0] - e1[0] * e2[2];
        double cz = e1[0] * e2[1] - e1[1] * e2[0];
        double a = std::sqrt(cx * cx + cy * cy + cz * cz) * 0.5;
        area_raw[t] = a;
        // thermal_data.cpp:67-75: clamp tiny elements
        if (a < 10.0 * FLT_EPSILON) a = 10.0 * FLT_EPSILON;
        area_d[t] = a;
        s.tri_area[t] = (float)a;
    }
    const uint32_t nb = s.bin_mode == IRTT_BIN_TRIANGLE ? nt : nv;
    s.nb = nb;
    s.ld = (nb + 63u) & ~63u;
    s.bin_area.assign(nb, 0.0);
    s.bin_eps_area.assign(nb, 0.0);
    s.bin_mass.assign(nb, 0.0);
    s.bin_source.assign(nb, 0.0);
    s.source_explicit = false;
    bool any_k = false;
    for (uint32_t t = 0; t < nt; t++) {
        const irtt_material &m = s.mat[t];
        double eps = 1.0 - (double)s.rho_d[t] - (double)s.rho_s[t];
        double rct = (double)m.density * (double)m.heat_capacity * (double)m.thickness;
        double src = (double)m.heat_flux;
        if (m.heat_conductivity != 0.0f) any_k = true;
        if (s.bin_mode == IRTT_BIN_TRIANGLE) {
            s.bin_area[t] += area_d[t];
            s.bin_eps_area[t] += eps * area_d[t];
            s.bin_mass[t] += area_raw[t] * rct;
            s.bin_source[t] += src * area_d[t];
        } else {
            for (int c = 0; c < 3; c++) {
                uint32_t v = s.idx[3 * t + c];
                s.bin_area[v] += area_d[t] / 3.0;
                s.bin_eps_area[v] += eps * area_d[t] / 3.0;
                s.bin_mass[v] += area_raw[t] / 3.0 * rct;
                s.bin_source[v] += src * area_d[t] / 3.0;
            }
        }
    }
    s.lap = Csr{};
    if (any_k) {
        if (s.bin_mode == IRTT_BIN_TRIANGLE)
            throw Error(IRTT_ERR_ARG, "heat conduction needs vertex bins (IRTT_BIN_VERTEX)");
        // cotangent Laplacian scaled by conductivity * thickness, thermal_data.cpp:259-313
        build_laplacian(s);
    }
    s.d_lap_rowptr.release();
    s.d_lap_col.release();
    s.d_lap_val.release();
}

void upload_geometry(Scene &s)
{
    hipStream_t st = s.stream;
    s.d_xyz.alloc(s.xyz.size());
    s.d_idx.alloc(s.idx.size());
    s.d_rho_d.alloc(s.nt);
    s.d_rho_s.alloc(s.nt);
    s.d_flags.alloc(s.nt);
    IRTT_HIP(hipMemcpyAsync(s.d_xyz.p, s.xyz.data(), s.xyz.size() * 4, hipMemcpyHostToDevice, st));
    IRTT_HIP(hipMemcpyAsync(s.d_idx.p, s.idx.data(), s.idx.size() * 4, hipMemcpyHostToDevice, st));
    IRTT_HIP(hipMemcpyAsync(s.d_rho_d.p, s.rho_d.data(), s.nt * 4, hipMemcpyHostToDevice, st));
    IRTT_HIP(hipMemcpyAsync(s.d_rho_s.p, s.rho_s.data(), s.nt * 4, hipMemcpyHostToDevice, st));
    IRTT_HIP(hipMemcpyAsync(s.d_flags.p, s.flags.data(), s.nt, hipMemcpyHostToDevice, st));
    std::vector<uint32_t> tr;
    tr.reserve(s.nt);
    for (uint32_t t = 0; t < s.nt; t++)
        if (s.flags[t] & IRTT_MAT_TRACEABLE) tr.push_back(t);
    s.d_traceable.release();
    if (!tr.empty()) {
        s.d_traceable.alloc(tr.size());
        IRTT_HIP(hipMemcpyAsync(s.d_traceable.p, tr.data(), tr.size() * 4, hipMemcpyHostToDevice, st));
    }
    IRTT_HIP(hipStreamSynchronize(st));
    s.bvh.valid = false;
    s.has_transport = false;
    op_reset(s);
}

}  // namespace

// ---- process-wide device memory cache and stream pool ----------------------
#ifndef IRTT_CACHE_GIB
#define IRTT_CACHE_GIB 96
#endif
namespace {
struct CachedBlock {
    void *p;
    size_t bytes;
    hipEvent_t ev;   // recorded on the freeing stream
    int dev;
    uint64_t age;
    bool async;      // allocated stream-ordered (hipMallocAsync)
};
std::mutex g_cache_mu;
std::vector<CachedBlock> g_cache;
std::unordered_map<void *, bool> g_async;               // live block -> stream-ordered
std::vector<std::pair<int, hipEvent_t>> g_events;      // idle events per device
std::vector<std::pair<int, hipStream_t>> g_streams;    // idle streams per device
size_t g_cache_bytes = 0;
uint64_t g_cache_clock = 0;
// counters: cache hits, fresh allocations, their bytes, nanoseconds in the
// allocator, blocks handed back by trims, their bytes
uint64_t g_ctr[6] = {0, 0, 0, 0, 0, 0};
constexpr size_t kCacheMax = (size_t)IRTT_CACHE_GIB << 30;

size_t round_bytes(size_t b)
{
    constexpr size_t g = 2u << 20;
    return b < (1u << 20) ? (b + 255) / 256 * 256 : (b + g - 1) / g * g;
}

hipEvent_t take_event_locked(int dev)
{
    for (size_t i = g_events.size(); i-- > 0;)
        if (g_events[i].first == dev) {
            hipEvent_t e = g_events[i].second;
            g_events.erase(g_events.begin() + i);
            return e;
        }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}

// takes the oldest cached blocks out of the cache until at most `keep` bytes
// stay cached (the caller holds the lock); release_blocks hands them back
std::vector<CachedBlock> trim_locked(size_t keep)
{
    std::vector<CachedBlock> out;
    if (g_cache_bytes <= keep) return out;
    std::vector<size_t> order(g_cache.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::sort(order.begin(), order.end(),
              [](size_t a, size_t b) { return g_cache[a].age < g_cache[b].age; });
    std::vector<uint8_t> take(g_cache.size(), 0);
    for (size_t i : order) {
        if (g_cache_bytes <= keep) break;
        take[i] = 1;
        g_cache_bytes -= g_cache[i].bytes;
    }
    std::vector<CachedBlock> stay;
    for (size_t i = 0; i < g_cache.size(); i++) (take[i] ? out : stay).push_back(g_cache[i]);
    for (const CachedBlock &b : out) {
        g_ctr[4]++;
        g_ctr[5] += b.bytes;
    }
    g_cache.swap(stay);
    return out;
}

// Returns blocks to the device without holding the lock.  A stream-ordered
// block is freed on `st` after the event of its last use (hipFreeAsync: the
// host does not wait and the device is not synchronised, so other scenes'
// streams keep running); a plain hipMalloc block needs hipFree.
void release_blocks(std::vector<CachedBlock> &v, hipStream_t st)
{
    for (CachedBlock &b : v) {
        if (b.async && st) {
            if (b.ev) (void)hipStreamWaitEvent(st, b.ev, 0);
            if (hipFreeAsync(b.p, st) != hipSuccess) (void)hipGetLastError();
        } else {
            if (b.ev) (void)hipEventSynchronize(b.ev);
            if (hipFree(b.p) != hipSuccess) (void)hipGetLastError();
        }
    }
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (CachedBlock &b : v)
        if (b.ev) g_events.emplace_back(b.dev, b.ev);
    v.clear();
}
}  // namespace

void *irtt::dev_alloc(size_t bytes, hipStream_t stream)
{
    int dev = 0;
    IRTT_HIP(hipGetDevice(&dev));
    const size_t want = round_bytes(bytes);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        size_t best = g_cache.size();
        for (size_t i = 0; i < g_cache.size(); i++) {
            const CachedBlock &b = g_cache[i];
            if (b.dev != dev || b.bytes < want || b.bytes > want + want / 4 + (32u << 20)) continue;
            if (best == g_cache.size() || b.bytes < g_cache[best].bytes) best = i;
        }
        if (best < g_cache.size()) {
            CachedBlock b = g_cache[best];
            g_cache.erase(g_cache.begin() + best);
            g_cache_bytes -= b.bytes;
            g_async[b.p] = b.async;
            g_ctr[0]++;
            if (b.ev) {
                // the block's last use on its old stream finishes first
                (void)hipStreamWaitEvent(stream, b.ev, 0);
                g_events.emplace_back(dev, b.ev);
            }
            return b.p;
        }
    }
    for (int attempt = 0; attempt < 3; attempt++) {
        void *p = nullptr;
        const auto t0 = std::chrono::steady_clock::now();
        hipError_t e = stream ? hipMallocAsync(&p, want, stream) : hipMalloc(&p, want);
        const uint64_t ns = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                                std::chrono::steady_clock::now() - t0).count();
        if (e == hipSuccess) {
            std::lock_guard<std::mutex> lk(g_cache_mu);
            g_async[p] = stream != nullptr;
            g_ctr[1]++;
            g_ctr[2] += want;
            g_ctr[3] += ns;
            return p;
        }
        (void)hipGetLastError();
        if (attempt == 0) {
            // out of memory: give the cached blocks back and try once more
            std::vector<CachedBlock> v;
            {
                std::lock_guard<std::mutex> lk(g_cache_mu);
                v = trim_locked(0);
            }
            if (!v.empty()) {
                release_blocks(v, stream);
                continue;
            }
            attempt = 1;
        }
        if (attempt == 1) {
            // the pool's unused reservations (returned blocks of other sizes)
            // go back to the device once every queued free has run, so that a
            // request larger than any of them can be reserved anew
            hipMemPool_t pool = nullptr;
            if (hipDeviceSynchronize() == hipSuccess &&
                hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool &&
                hipMemPoolTrimTo(pool, 0) == hipSuccess)
                continue;
            (void)hipGetLastError();
        }
        throw Error(IRTT_ERR_NOMEM, "device allocation of " + std::to_string(bytes) +
                                        " bytes failed: " + hipGetErrorString(e));
    }
    throw Error(IRTT_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
}

void irtt::dev_free(void *p, size_t bytes, hipStream_t stream)
{
    if (!p) return;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    std::vector<CachedBlock> v;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        bool async = false;
        auto it = g_async.find(p);
        if (it != g_async.end()) {
            async = it->second;
            g_async.erase(it);
        }
        hipEvent_t ev = take_event_locked(dev);
        if (ev && hipEventRecord(ev, stream) != hipSuccess) {
            (void)hipGetLastError();
            g_events.emplace_back(dev, ev);
            ev = nullptr;
        }
        if (!ev) (void)hipStreamSynchronize(stream);  // no event: wait for the last use here
        g_cache.push_back(CachedBlock{p, round_bytes(bytes), ev, dev, ++g_cache_clock, async});
        g_cache_bytes += round_bytes(bytes);
        // over the cap: the oldest blocks go back to the device, down to 3/4
        if (g_cache_bytes > kCacheMax) v = trim_locked(kCacheMax / 4 * 3);
    }
    if (!v.empty()) release_blocks(v, stream);
}

size_t irtt::dev_available()
{
    size_t fr = 0, tot = 0;
    IRTT_HIP(hipMemGetInfo(&fr, &tot));
    size_t pooled = 0;
    int dev = 0;
    hipMemPool_t pool = nullptr;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess &&
        pool) {
        uint64_t reserved = 0, used = 0;
        if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &reserved) ==
                hipSuccess &&
            hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &used) == hipSuccess &&
            reserved > used)
            pooled = (size_t)(reserved - used);
    }
    (void)hipGetLastError();
    return fr + pooled + dev_cache_bytes();
}

void irtt::dev_counters(uint64_t *out)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (int i = 0; i < 6; i++) out[i] = g_ctr[i];
}

size_t irtt::dev_cache_bytes()
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    return g_cache_bytes;
}

hipStream_t irtt::stream_take()
{
    int dev = 0;
    IRTT_HIP(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        for (size_t i = g_streams.size(); i-- > 0;)
            if (g_streams[i].first == dev) {
                hipStream_t st = g_streams[i].second;
                g_streams.erase(g_streams.begin() + i);
                return st;
            }
    }
    hipStream_t st = nullptr;
    IRTT_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    return st;
}

void irtt::stream_give(hipStream_t st)
{
    if (!st) return;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    (void)hipStreamSynchronize(st);
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_streams.emplace_back(dev, st);
}

// process-wide pool of pinned host blocks (never freed: freeing pinned
// memory synchronises the device)
namespace {
std::mutex g_pin_mu;
std::vector<std::pair<void *, size_t>> g_pin_free;
}  // namespace

void *irtt::pinned_take(size_t bytes, size_t &got)
{
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        size_t best = g_pin_free.size();
        for (size_t i = 0; i < g_pin_free.size(); i++)
            if (g_pin_free[i].second >= bytes &&
                (best == g_pin_free.size() || g_pin_free[i].second < g_pin_free[best].second))
                best = i;
        if (best < g_pin_free.size()) {
            void *p = g_pin_free[best].first;
            got = g_pin_free[best].second;
            g_pin_free.erase(g_pin_free.begin() + best);
            return p;
        }
    }
    void *p = nullptr;
    IRTT_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    got = bytes;
    return p;
}

void irtt::pinned_give(void *p, size_t bytes)
{
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_pin_free.emplace_back(p, bytes);
}

void irtt::set_last_error(const std::string &msg) { g_err = msg; }

thread_local hipStream_t irtt::t_alloc_stream = nullptr;

uint64_t Scene::device_bytes() const
{
    return d_xyz.bytes() + d_idx.bytes() + d_rho_d.bytes() + d_rho_s.bytes() + d_flags.bytes() +
           d_traceable.bytes() + bvh.nodes.bytes() + bvh.tris.bytes() + bvh.morton.bytes() +
           d_q.bytes() + d_acc.bytes() + d_touch.bytes() + d_T.bytes() + d_diag.bytes() + d_colacc.bytes() +
           d_tdiag.bytes() + d_h16.bytes() + d_h16s.bytes() + d_h16f.bytes() + d_h16ft.bytes() +
           d_gpart.bytes() + d_gpart64.bytes() + aca.d_ut.bytes() +
           aca.d_v.bytes() + aca.d_uf.bytes() + aca.d_vf.bytes() + aca.d_uh.bytes() +
           aca.d_vh.bytes() + d_stats.bytes() + d_vec.bytes() + d_red.bytes() + d_lap_rowptr.bytes() +
           d_lap_col.bytes() + d_lap_val.bytes() + sp.rowptr.bytes() + sp.col.bytes() +
           sp.val.bytes() + sp.colptr.bytes() + sp.row.bytes() + sp.cval.bytes() +
           sp.d_xt.bytes() + d_absorber.bytes() + paths.bytes();
}

extern "C" {

int irtt_version(void) { return 1; }

const char *irtt_last_error(void) { return g_err.c_str(); }

int irtt_device_info(char *name, size_t name_len, int *cu_count)
{
    return guard([&] {
        int n = 0;
        IRTT_HIP(hipGetDeviceCount(&n));
        if (n < 1) throw Error(IRTT_ERR_HIP, "no HIP device");
        hipDeviceProp_t prop;
        IRTT_HIP(hipGetDeviceProperties(&prop, 0));
        if (name && name_len) {
            std::string nm = std::string(prop.name) + " (" + prop.gcnArchName + ")";
            std::strncpy(name, nm.c_str(), name_len - 1);
            name[name_len - 1] = 0;
        }
        if (cu_count) *cu_count = prop.multiProcessorCount;
    });
}

int irtt_scene_create(const float *xyz, uint32_t nv, const uint32_t *tri, uint32_t nt,
                      const irtt_material *mat, uint32_t bin_mode, irtt_scene *out)
{
    return guard([&] {
        if (!out) throw Error(IRTT_ERR_ARG, "out is null");
        *out = nullptr;
        if (!xyz || !tri || !mat || nv == 0 || nt == 0)
            throw Error(IRTT_ERR_ARG, "empty geometry or null pointer");
        if (bin_mode > IRTT_BIN_TRIANGLE) throw Error(IRTT_ERR_ARG, "unknown bin mode");
        configure_pool();
        for (uint64_t i = 0; i < 3ull * nt; i++)
            if (tri[i] >= nv)
                throw Error(IRTT_ERR_ARG, "triangle index " + std::to_string(tri[i]) +
                                              " out of range (nv = " + std::to_string(nv) + ")");
        for (uint64_t i = 0; i < 3ull * nv; i++)
            if (!std::isfinite(xyz[i])) throw Error(IRTT_ERR_ARG, "non-finite vertex coordinate");
        auto s = std::make_unique<Scene>();
        s->nv = nv;
        s->nt = nt;
        s->bin_mode = bin_mode;
        s->xyz.assign(xyz, xyz + 3ull * nv);
        s->idx.assign(tri, tri + 3ull * nt);
        s->mat.assign(mat, mat + nt);
        assemble(*s);
        s->bin_kelvin0.assign(s->nb, 0.0);
        s->bin_fixed.assign(s->nb, 0);
        s->x.assign(s->nb, 0.0);
        s->stream = stream_take();
        IRTT_HIP(hipEventCreate(&s->ev0));
        IRTT_HIP(hipEventCreate(&s->ev1));
        IRTT_HIP(hipEventCreate(&s->ev2));
        t_alloc_stream = s->stream;
        try {
            upload_geometry(*s);
        } catch (...) {
            hipStream_t st = s->stream;
            hipEvent_t e[3] = {s->ev0, s->ev1, s->ev2};
            s.reset();
            (void)hipStreamSynchronize(st);
            for (hipEvent_t x : e) (void)hipEventDestroy(x);
            stream_give(st);
            t_alloc_stream = nullptr;
            throw;
        }
        *out = reinterpret_cast<irtt_scene>(s.release());
    });
}

int irtt_scene_destroy(irtt_scene h)
{
    return guard([&] {
        if (!h) return;
        Scene *s = S(h);
        t_alloc_stream = nullptr;
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        hipStream_t st = s->stream;
        hipEvent_t e[3] = {s->ev0, s->ev1, s->ev2};
        delete s;  // buffers go back to the pool in the scene's stream order
        if (st) (void)hipStreamSynchronize(st);
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
        stream_give(st);
    });
}

int irtt_scene_info_get(irtt_scene h, irtt_scene_info *info)
{
    return guard([&] {
        Scene *s = S(h);
        if (!info) throw Error(IRTT_ERR_ARG, "info is null");
        info->n_vertices = s->nv;
        info->n_triangles = s->nt;
        info->n_bins = s->nb;
        info->n_traceable = (uint32_t)s->d_traceable.n;
        info->bvh_max_depth = s->bvh.valid ? s->bvh.max_depth : 0;
        info->bin_mode = s->bin_mode;
        info->device_bytes = s->device_bytes();
    });
}

int irtt_scene_set_vertex_state(irtt_scene h, const float *kelvin, const uint8_t *fixed)
{
    return guard([&] {
        Scene *s = S(h);
        if (!kelvin) throw Error(IRTT_ERR_ARG, "kelvin is null");
        if (s->bin_mode == IRTT_BIN_VERTEX) {
            for (uint32_t v = 0; v < s->nv; v++) {
                s->bin_kelvin0[v] = kelvin[v];
                s->bin_fixed[v] = fixed ? (fixed[v] ? 1 : 0) : 0;
            }
        } else {
            for (uint32_t t = 0; t < s->nt; t++) {
                uint32_t a = s->idx[3 * t], b = s->idx[3 * t + 1], c = s->idx[3 * t + 2];
                s->bin_kelvin0[t] = ((double)kelvin[a] + kelvin[b] + kelvin[c]) / 3.0;
                s->bin_fixed[t] = fixed ? ((fixed[a] && fixed[b] && fixed[c]) ? 1 : 0) : 0;
            }
        }
        // ThermalData::reset, thermal_data.cpp:196-199
        s->x = s->bin_kelvin0;
    });
}

int irtt_scene_set_bin_state(irtt_scene h, const float *kelvin, const uint8_t *fixed)
{
    return guard([&] {
        Scene *s = S(h);
        if (!kelvin) throw Error(IRTT_ERR_ARG, "kelvin is null");
        for (uint32_t i = 0; i < s->nb; i++) {
            s->bin_kelvin0[i] = kelvin[i];
            s->bin_fixed[i] = fixed ? (fixed[i] ? 1 : 0) : 0;
        }
        s->x = s->bin_kelvin0;
    });
}

int irtt_scene_set_source(irtt_scene h, const double *source)
{
    return guard([&] {
        Scene *s = S(h);
        if (!source) throw Error(IRTT_ERR_ARG, "source is null");
        for (uint32_t i = 0; i < s->nb; i++) {
            if (!std::isfinite(source[i])) throw Error(IRTT_ERR_ARG, "non-finite source");
            s->bin_source[i] = source[i];
        }
        s->source_explicit = true;
    });
}

int irtt_scene_update_vertices(irtt_scene h, const float *xyz)
{
    return guard([&] {
        Scene *s = S(h);
        if (!xyz) throw Error(IRTT_ERR_ARG, "xyz is null");
        for (uint64_t i = 0; i < 3ull * s->nv; i++)
            if (!std::isfinite(xyz[i])) throw Error(IRTT_ERR_ARG, "non-finite vertex coordinate");
        s->xyz.assign(xyz, xyz + 3ull * s->nv);
        assemble(*s);
        upload_geometry(*s);
    });
}

int irtt_scene_get_bin_data(irtt_scene h, double *area, double *eps_area, double *mass)
{
    return guard([&] {
        Scene *s = S(h);
        if (area) std::copy(s->bin_area.begin(), s->bin_area.end(), area);
        if (eps_area) std::copy(s->bin_eps_area.begin(), s->bin_eps_area.end(), eps_area);
        if (mass) std::copy(s->bin_mass.begin(), s->bin_mass.end(), mass);
    });
}

int irtt_build_bvh(irtt_scene h)
{
    return guard([&] { bvh_build(*S(h)); });
}

int irtt_get_bvh(irtt_scene h, uint32_t *nodes, uint32_t *tris, uint32_t *morton,
                 uint32_t *n_leaves, uint32_t *max_depth)
{
    return guard([&] {
        Scene *s = S(h);
        if (!s->bvh.valid) bvh_build(*s);
        const uint32_t n = s->bvh.n_leaves;
        if (n_leaves) *n_leaves = n;
        if (max_depth) *max_depth = s->bvh.max_depth;
        if (n == 0) return;
        if (nodes && n > 1)
            IRTT_HIP(hipMemcpy(nodes, s->bvh.nodes.p, (size_t)(n - 1) * 64, hipMemcpyDeviceToHost));
        if (tris) IRTT_HIP(hipMemcpy(tris, s->bvh.tris.p, (size_t)n * 48, hipMemcpyDeviceToHost));
        if (morton)
            IRTT_HIP(hipMemcpy(morton, s->bvh.morton.p, (size_t)n * 4, hipMemcpyDeviceToHost));
    });
}

int irtt_trace_rays(irtt_scene h, const float *rays, uint32_t n, float tmin, float tmax,
                    int32_t *prim, float *t, float *b1, float *b2)
{
    return guard([&] {
        Scene *s = S(h);
        if (n && (!rays || !prim || !t || !b1 || !b2)) throw Error(IRTT_ERR_ARG, "null pointer");
        trace_rays(*s, rays, n, tmin, tmax, prim, t, b1, b2);
    });
}

int irtt_trace_rays_noalloc(irtt_scene h, const float *rays, uint32_t n, float tmin, float tmax,
                            int32_t *prim, float *t, float *b1, float *b2)
{
    return guard([&] {
        Scene *s = S(h);
        if (n && (!rays || !prim || !t || !b1 || !b2)) throw Error(IRTT_ERR_ARG, "null pointer");
        trace_rays_noalloc(*s, rays, n, tmin, tmax, prim, t, b1, b2);
    });
}

uint32_t irtt_trace_rays_noalloc_max(void) { return trace_rays_noalloc_max(); }

int irtt_build_transport(irtt_scene h, const irtt_transport_params *p, irtt_transport_stats *st)
{
    return guard([&] {
        Scene *s = S(h);
        if (!p) throw Error(IRTT_ERR_ARG, "params is null");
        if (p->layout == IRTT_LAYOUT_DENSE) {
            // dense operator: nb*ld fixed-point (8 B) + FP32 (4 B) + chunk marks
            size_t need = (size_t)s->nb * s->ld * 12 + (((size_t)s->nb * s->ld) >> kTouchShift);
            const size_t fr = dev_available();
            size_t held = s->d_acc.bytes() + s->d_touch.bytes() + s->d_T.bytes();
            if (need > fr + held)
                throw Error(IRTT_ERR_LIMIT, "dense operator needs " + std::to_string(need) +
                                                " bytes, " + std::to_string(fr + held) +
                                                " available");
        }
        irtt_transport_stats tmp{};
        transport_build(*s, *p, tmp);
        if (st) *st = tmp;
    });
}

int irtt_select_operator(irtt_scene h, uint32_t op)
{
    return guard([&] { op_select(*S(h), op); });
}

int irtt_apply_operator(irtt_scene h, uint32_t n_rhs, const double *X, double *Y,
                        uint32_t transposed)
{
    return guard([&] {
        Scene *s = S(h);
        if (!s->has_transport) throw Error(IRTT_ERR_STATE, "transport operator not built");
        if (n_rhs == 0) return;
        if (!X || !Y) throw Error(IRTT_ERR_ARG, "null pointer");
        const size_t nb = s->nb, ld = s->ld;
        DBuf<double> dx, dy;
        dx.alloc((size_t)n_rhs * ld);
        dy.alloc((size_t)n_rhs * ld);
        hipStream_t st = s->stream;
        IRTT_HIP(hipMemsetAsync(dx.p, 0, dx.bytes(), st));
        IRTT_HIP(hipMemcpy2DAsync(dx.p, ld * 8, X, nb * 8, nb * 8, n_rhs, hipMemcpyHostToDevice, st));
        op_apply_batch(*s, n_rhs, dx.p, dy.p, transposed != 0);
        IRTT_HIP(hipGetLastError());
        IRTT_HIP(hipMemcpy2DAsync(Y, nb * 8, dy.p, ld * 8, nb * 8, n_rhs, hipMemcpyDeviceToHost, st));
        IRTT_HIP(hipStreamSynchronize(st));
    });
}

int irtt_simulate(irtt_scene h, const irtt_solver_params *p, uint32_t steps, double *history,
                  irtt_solve_stats *st)
{
    return guard([&] {
        Scene *s = S(h);
        if (!p || !history) throw Error(IRTT_ERR_ARG, "null pointer");
        irtt_solve_stats tmp{};
        simulate_history(*s, *p, steps, history, tmp);
        if (st) *st = tmp;
    });
}

int irtt_adjoint_history(irtt_scene h, const irtt_solver_params *p, uint32_t steps,
                         const double *history, const double *weights, const double *targets,
                         uint32_t target_stride, double *lambda_history, double *objective,
                         const irtt_adjoint_params *ap, double *grad_x, irtt_solve_stats *st)
{
    return guard([&] {
        Scene *s = S(h);
        if (!p || !history || !weights || !targets || !lambda_history)
            throw Error(IRTT_ERR_ARG, "null pointer");
        if (target_stride != 0 && target_stride != s->nb)
            throw Error(IRTT_ERR_ARG, "target_stride must be 0 or n_bins");
        irtt_solve_stats tmp{};
        adjoint_history(*s, *p, steps, history, weights, targets, target_stride, lambda_history,
                        objective, ap, grad_x, tmp);
        if (st) *st = tmp;
    });
}

int irtt_solve_batch(irtt_scene h, const irtt_solver_params *p, uint32_t n_rhs,
                     const double *source, double *kelvin, irtt_solve_stats *st)
{
    return guard([&] {
        Scene *s = S(h);
        if (!p || !kelvin) throw Error(IRTT_ERR_ARG, "null pointer");
        if (n_rhs == 0) return;
        irtt_solve_stats tmp{};
        solve_batch(*s, *p, n_rhs, source, kelvin, tmp);
        if (st) *st = tmp;
    });
}

int irtt_adjoint_solve_batch(irtt_scene h, const irtt_solver_params *p, uint32_t n_rhs,
                             const double *x, const double *rhs, double *lambda,
                             irtt_solve_stats *st)
{
    return guard([&] {
        Scene *s = S(h);
        if (!p || !x || !rhs || !lambda) throw Error(IRTT_ERR_ARG, "null pointer");
        if (n_rhs == 0) return;
        irtt_solve_stats tmp{};
        adjoint_solve_batch(*s, *p, n_rhs, x, rhs, lambda, tmp);
        if (st) *st = tmp;
    });
}

int irtt_compress_transport(irtt_scene h, const irtt_aca_params *p, irtt_aca_stats *st)
{
    return guard([&] {
        Scene *s = S(h);
        if (!p) throw Error(IRTT_ERR_ARG, "params is null");
        irtt_aca_stats tmp{};
        aca_compress(*s, *p, tmp);
        if (st) *st = tmp;
    });
}

int irtt_compress_hierarchical(irtt_scene h, const irtt_hmatrix_params *p,
                               irtt_hmatrix_stats *st)
{
    return guard([&] {
        Scene *s = S(h);
        if (!p) throw Error(IRTT_ERR_ARG, "params is null");
        irtt_hmatrix_stats tmp{};
        hm_compress(*s, *p, tmp);
        s->op_kind = IRTT_OP_HMATRIX;
        if (st) *st = tmp;
    });
}

int irtt_get_hmatrix(irtt_scene h, uint32_t *perm, uint32_t *n_far, uint32_t *blocks,
                     double *err_rel, int32_t *pivots)
{
    return guard([&] { hm_get(*S(h), perm, n_far, blocks, err_rel, pivots); });
}

int irtt_get_hmatrix_factors(irtt_scene h, uint32_t b, double *U, double *V)
{
    return guard([&] { hm_get_factors(*S(h), b, U, V); });
}

int irtt_get_aca(irtt_scene h, double *U, double *V, double *cdiag, int32_t *piv_rows,
                 int32_t *piv_cols)
{
    return guard([&] { aca_get(*S(h), U, V, cdiag, piv_rows, piv_cols); });
}

int irtt_get_transport(irtt_scene h, float *dense)
{
    return guard([&] {
        Scene *s = S(h);
        if (!s->has_transport) throw Error(IRTT_ERR_STATE, "transport operator not built");
        if (!dense) throw Error(IRTT_ERR_ARG, "dense is null");
        const size_t nb = s->nb;
        if (s->layout == IRTT_LAYOUT_SPARSE) {
            std::vector<uint32_t> rp(nb + 1), col(s->sp.nnz);
            std::vector<float> val(s->sp.nnz);
            sp_download(*s, rp.data(), col.data(), val.data());
            std::fill(dense, dense + nb * nb, 0.0f);
            for (size_t i = 0; i < nb; i++)
                for (uint32_t k = rp[i]; k < rp[i + 1]; k++) dense[i * nb + col[k]] = val[k];
            return;
        }
        IRTT_HIP(hipMemcpy2D(dense, nb * 4, s->d_T.p, (size_t)s->ld * 4, nb * 4, nb,
                             hipMemcpyDeviceToHost));
    });
}

int irtt_get_transport_sparse(irtt_scene h, uint64_t *nnz, uint32_t *rowptr, uint32_t *col,
                              float *val)
{
    return guard([&] {
        Scene *s = S(h);
        if (!s->has_transport) throw Error(IRTT_ERR_STATE, "transport operator not built");
        if (!nnz) throw Error(IRTT_ERR_ARG, "nnz is null");
        const bool fill = rowptr && col && val;
        const size_t nb = s->nb;
        if (s->layout == IRTT_LAYOUT_SPARSE) {
            *nnz = s->sp.nnz;
            if (fill) sp_download(*s, rowptr, col, val);
            return;
        }
        std::vector<float> d(nb * nb);
        IRTT_HIP(hipMemcpy2D(d.data(), nb * 4, s->d_T.p, (size_t)s->ld * 4, nb * 4, nb,
                             hipMemcpyDeviceToHost));
        uint64_t n = 0;
        for (size_t i = 0; i < nb; i++) {
            if (fill) rowptr[i] = (uint32_t)n;
            for (size_t j = 0; j < nb; j++) {
                const float v = d[i * nb + j];
                if (v == 0.0f && i != j) continue;
                if (fill) {
                    col[n] = (uint32_t)j;
                    val[n] = v;
                }
                n++;
            }
        }
        if (fill) rowptr[nb] = (uint32_t)n;
        *nnz = n;
    });
}

int irtt_get_transport_fixed(irtt_scene h, uint64_t *acc, double *quantum)
{
    return guard([&] {
        Scene *s = S(h);
        if (!s->has_transport) throw Error(IRTT_ERR_STATE, "transport operator not built");
        if (acc && s->layout != IRTT_LAYOUT_DENSE)
            throw Error(IRTT_ERR_STATE, "fixed-point accumulators are kept by the dense layout only");
        if (quantum) *quantum = s->quantum;
        if (acc)
            IRTT_HIP(hipMemcpy2D(acc, (size_t)s->nb * 8, s->d_acc.p, (size_t)s->ld * 8,
                                 (size_t)s->nb * 8, s->nb, hipMemcpyDeviceToHost));
    });
}

int irtt_get_column_sums(irtt_scene h, double *colsum)
{
    return guard([&] {
        Scene *s = S(h);
        if (!s->has_transport) throw Error(IRTT_ERR_STATE, "transport operator not built");
        if (!colsum) throw Error(IRTT_ERR_ARG, "colsum is null");
        std::copy(s->colsum.begin(), s->colsum.end(), colsum);
    });
}

int irtt_get_transport_line(irtt_scene h, uint32_t bin, uint32_t column, float *out)
{
    return guard([&] {
        Scene *s = S(h);
        if (!s->has_transport) throw Error(IRTT_ERR_STATE, "transport operator not built");
        if (!out) throw Error(IRTT_ERR_ARG, "out is null");
        const uint32_t nb = s->nb;
        if (bin >= nb) throw Error(IRTT_ERR_ARG, "bin out of range");
        hipStream_t st = s->stream;
        if (s->layout == IRTT_LAYOUT_SPARSE) {
            const Sparse &p = s->sp;
            const uint32_t *ptr = column ? p.colptr.p : p.rowptr.p;
            const uint32_t *idx = column ? p.row.p : p.col.p;
            const float *val = column ? p.cval.p : p.val.p;
            uint32_t be[2];
            IRTT_HIP(hipMemcpyAsync(be, ptr + bin, 8, hipMemcpyDeviceToHost, st));
            IRTT_HIP(hipStreamSynchronize(st));
            const uint32_t n = be[1] - be[0];
            std::vector<uint32_t> ix(n);
            std::vector<float> v(n);
            if (n) {
                IRTT_HIP(hipMemcpyAsync(ix.data(), idx + be[0], (size_t)n * 4, hipMemcpyDeviceToHost, st));
                IRTT_HIP(hipMemcpyAsync(v.data(), val + be[0], (size_t)n * 4, hipMemcpyDeviceToHost, st));
                IRTT_HIP(hipStreamSynchronize(st));
            }
            std::fill(out, out + nb, 0.0f);
            for (uint32_t k = 0; k < n; k++) out[ix[k]] = v[k];
            return;
        }
        if (!column) {
            IRTT_HIP(hipMemcpyAsync(out, s->d_T.p + (size_t)bin * s->ld, (size_t)nb * 4,
                                    hipMemcpyDeviceToHost, st));
        } else {
            IRTT_HIP(hipMemcpy2DAsync(out, sizeof(float), s->d_T.p + bin, (size_t)s->ld * 4,
                                      sizeof(float), nb, hipMemcpyDeviceToHost, st));
        }
        IRTT_HIP(hipStreamSynchronize(st));
    });
}

int irtt_solve(irtt_scene h, const irtt_solver_params *p, irtt_solve_stats *st)
{
    return guard([&] {
        Scene *s = S(h);
        if (!p) throw Error(IRTT_ERR_ARG, "params is null");
        irtt_solve_stats tmp{};
        solve(*s, *p, tmp);
        if (st) *st = tmp;
    });
}

int irtt_get_temperatures(irtt_scene h, double *kelvin)
{
    return guard([&] {
        Scene *s = S(h);
        if (!kelvin) throw Error(IRTT_ERR_ARG, "kelvin is null");
        std::copy(s->x.begin(), s->x.end(), kelvin);
    });
}

int irtt_set_temperatures(irtt_scene h, const double *kelvin)
{
    return guard([&] {
        Scene *s = S(h);
        if (!kelvin) throw Error(IRTT_ERR_ARG, "kelvin is null");
        s->x.assign(kelvin, kelvin + s->nb);
    });
}

int irtt_residual(irtt_scene h, const irtt_solver_params *p, const double *x, double *r)
{
    return guard([&] {
        Scene *s = S(h);
        if (!p || !x || !r) throw Error(IRTT_ERR_ARG, "null pointer");
        residual_host(*s, *p, x, r);
    });
}

int irtt_adjoint_solve(irtt_scene h, const irtt_solver_params *p, const double *x,
                       const double *rhs, double *lambda, irtt_solve_stats *st)
{
    return guard([&] {
        Scene *s = S(h);
        if (!p || !rhs || !lambda) throw Error(IRTT_ERR_ARG, "null pointer");
        irtt_solve_stats tmp{};
        adjoint_solve(*s, *p, x, rhs, lambda, tmp);
        if (st) *st = tmp;
    });
}

int irtt_adjoint_trace(irtt_scene h, const irtt_adjoint_params *ap, uint32_t n_hist,
                       const double *lambda, const double *x, double *grad_x,
                       irtt_adjoint_stats *st)
{
    return guard([&] {
        Scene *s = S(h);
        if (!lambda || !x || !grad_x) throw Error(IRTT_ERR_ARG, "null pointer");
        irtt_adjoint_params def{};
        irtt_adjoint_stats tmp{};
        adjoint_trace(*s, ap ? *ap : def, n_hist, lambda, x, grad_x, tmp);
        if (st) *st = tmp;
    });
}

int irtt_project_gradient(const double *grad_x, uint32_t n_vertices, uint32_t nnz,
                          const uint32_t *vertex, const uint32_t *param, const float *dxdp,
                          uint32_t n_params, double *dp)
{
    return guard([&] {
        if (nnz && (!grad_x || !vertex || !param || !dxdp || !dp))
            throw Error(IRTT_ERR_ARG, "null pointer");
        for (uint32_t k = 0; k < nnz; k++) {
            if (vertex[k] >= n_vertices || param[k] >= n_params)
                throw Error(IRTT_ERR_ARG, "projection entry " + std::to_string(k) + " out of range");
            const double *g = grad_x + 3 * (size_t)vertex[k];
            dp[param[k]] += (double)dxdp[3 * k] * g[0] + (double)dxdp[3 * k + 1] * g[1] +
                            (double)dxdp[3 * k + 2] * g[2];
        }
    });
}

int irtt_source_gradient(irtt_scene h, const double *lambda, double *grad_x)
{
    return guard([&] {
        Scene *s = S(h);
        if (!lambda || !grad_x) throw Error(IRTT_ERR_ARG, "null pointer");
        source_gradient(*s, lambda, grad_x);
    });
}

int irtt_design_gradient(irtt_scene h, const irtt_solver_params *p, const double *dLdx,
                         const irtt_adjoint_params *ap, uint32_t nnz, const uint32_t *vertex,
                         const uint32_t *param, const float *dxdp, uint32_t n_params,
                         double *dp, double *grad_x, double *lambda, irtt_solve_stats *st)
{
    return guard([&] {
        Scene *s = S(h);
        if (!p || !dLdx) throw Error(IRTT_ERR_ARG, "null pointer");
        if (dp && nnz && (!vertex || !param || !dxdp))
            throw Error(IRTT_ERR_ARG, "null projection arrays");
        irtt_adjoint_params a = ap ? *ap : irtt_adjoint_params{0.05f, 1e-2f, 0u, 0u};
        irtt_solve_stats tmp{};
        design_gradient(*s, *p, dLdx, a, nnz, vertex, param, dxdp, n_params, dp, grad_x, lambda,
                        tmp);
        if (st) *st = tmp;
    });
}

}  // extern "C"
