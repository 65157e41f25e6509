// trace.hip -- closest-hit traversal of the LBVH and the forward photon
// transport kernel.
//
// Replaces the reference's ray-tracing pipeline launch
// (src/implementations/thermal_renderer/thermal_transport.cpp:599-738) and its
// raygen program (assets/shader/raytracing_thermal/thermal_tracing_unified.glsl:
// 247-601, TRACING_MODE 1).  CDNA4 has no ray-tracing units, so traversal is a
// per-lane stack walk over 64-byte two-child nodes with the stack in LDS.
// Photon hits are accumulated as exact fixed-point integers (one 64-bit
// integer add per absorbed photon, weight q_t of its source triangle), so the
// operator is bit-reproducible and independent of launch order.  The dense
// layout adds into an n_bins x ld array; the sparse layout into a hash table
// keyed by (absorbing bin, emitting bin) (sparse_table.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "device_math.hpp"
#include "irtt_internal.hpp"
#include "path_records.hpp"
#include "sparse_table.hpp"
#include "traverse.hpp"

namespace irtt {
namespace {

template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_trace_rays(
    const float4 *__restrict__ nodes, const float4 *__restrict__ tris, uint32_t n_leaves,
    const float *__restrict__ rays, uint32_t n, float tmin, float tmax, int32_t *prim, float *tout,
    float *b1o, float *b2o)
{
    __shared__ uint32_t stk_n[STACK * kTraceBlock];
    __shared__ float stk_t[stack_t_entries(STACK) * kTraceBlock];
    uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    const float *r = rays + 6 * (size_t)i;
    uint32_t nv = 0, nt = 0;
    Hit h = trace_closest<STACK>(nodes, tris, n_leaves, V3{r[0], r[1], r[2]},
                                 V3{r[3], r[4], r[5]}, tmin, tmax, stk_n + threadIdx.x,
                                 stk_t + threadIdx.x, nv, nt);
    prim[i] = h.prim;
    tout[i] = h.t;
    b1o[i] = h.b1;
    b2o[i] = h.b2;
}

// statistic slots
enum { S_RAYS = 0, S_ABS, S_MISS, S_CLOSE, S_BACK, S_DEG, S_FORCED, S_COUNT };

struct PhotonArgs {
    const float *xyz;
    const uint32_t *idx;
    const float *rho_d, *rho_s;
    const uint8_t *flags;
    const unsigned long long *q;
    const float4 *nodes, *tris;
    uint32_t n_leaves;
    uint32_t nt, nb, ld, bin_mode;
    uint32_t threads_per_tri, rays_per_thread, per_tri, batch_seed, max_depth;
    float tiny_dist;
    uint64_t photon_begin, photon_end;
    PhotonSink sink;
    uint32_t *absorber;  // per photon, or nullptr
    unsigned long long *stats;
};

__device__ __forceinline__ V3 load_v(const float *xyz, uint32_t v)
{
    return V3{xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2]};
}

// waves per SIMD that the LDS stack admits (160 KiB of LDS per CU, four
// SIMDs), capped at 6: the register budget is held to that occupancy
constexpr int photon_waves(int stack)
{
    return (160 * 1024 / (stack * 4 * kTraceBlock)) / 4 < 6
               ? (160 * 1024 / (stack * 4 * kTraceBlock)) / 4
               : 6;
}

// Photon g's first ray (getEmissionRay, thermal_tracing_unified.glsl:125-161,
// seed at :291): origin, direction, RNG state after it, emitting bin and
// triangle.  False for a degenerate (zero-area) emitter.
__device__ __forceinline__ bool emit_photon(const PhotonArgs &a, uint64_t g, uint32_t &seed, V3 &o,
                                            V3 &d, uint32_t &src, uint32_t &tri, float &eb1,
                                            float &eb2, int &ecorner)
{
    uint32_t t = (uint32_t)(g / a.per_tri);
    uint32_t k = (uint32_t)(g - (uint64_t)t * a.per_tri);
    uint32_t x = k / a.rays_per_thread, nn = k - x * a.rays_per_thread;
    seed = tea16(a.nt * x + t, a.batch_seed + nn);
    uint32_t i0 = a.idx[3 * t], i1 = a.idx[3 * t + 1], i2 = a.idx[3 * t + 2];
    V3 p0 = load_v(a.xyz, i0), p1 = load_v(a.xyz, i1), p2 = load_v(a.xyz, i2);
    float u0 = rnd(seed), u1 = rnd(seed);
    (void)rnd(seed);
    float su = sqrt_rn(u0);
    float b0 = 1.0f - su, b1 = su * (1.0f - u1), b2 = su * u1;
    V3 pos = bary_point(p0, p1, p2, b0, b1, b2);
    float len;
    V3 nrm = normalize3(cross3(p1 - p0, p2 - p0), len);
    o = offset_ray(pos, nrm);
    d = nrm;
    if (a.flags[t] & IRTT_MAT_DIFFUSE_EMISSION) {
        float r0 = rnd(seed), r1 = rnd(seed);
        (void)rnd(seed);
        d = tangent_to_world(cosine_hemisphere(r0, r1), nrm);
    }
    if (!(len > 0.0f)) return false;
    int sc = max_corner(b0, b1, b2);
    src = a.bin_mode ? t : (sc == 0 ? i0 : (sc == 1 ? i1 : i2));
    tri = t;
    eb1 = b1;
    eb2 = b2;
    ecorner = sc;
    return true;
}

// The answered ray segment `depth` (1 = emission ray) of photon g: either ends
// the photon and returns the counter slot of how (S_MISS, S_CLOSE, S_BACK,
// S_ABS, or S_FORCED at max_depth; absorptions add the photon's weight qt to
// (absorbing bin, src)), or turns o, d, seed into its next ray (getNextRay,
// thermal_tracing_unified.glsl:163-197) and returns -1.  With PATHS, seg
// tells whether the adjoint's path replay takes the segment (the same
// tests, except that a hit with a non-finite normal ends the replayed path).
template <bool PATHS>
__device__ __forceinline__ int shade_segment(const PhotonArgs &a, const Hit &h, uint64_t g,
                                             uint32_t src, unsigned long long qt, uint32_t depth,
                                             uint32_t &seed, V3 &o, V3 &d, bool &seg)
{
    if constexpr (PATHS) seg = false;
    if (h.prim < 0) return S_MISS;
    if (h.t < a.tiny_dist) return S_CLOSE;
    uint32_t hp = (uint32_t)h.prim;
    uint32_t j0 = a.idx[3 * hp], j1 = a.idx[3 * hp + 1], j2 = a.idx[3 * hp + 2];
    V3 q0 = load_v(a.xyz, j0), q1 = load_v(a.xyz, j1), q2 = load_v(a.xyz, j2);
    V3 hn = normalize3(cross3(q1 - q0, q2 - q0));
    const float hd = dot3(hn, d);
    if (hd > -1e-3f) return S_BACK;
    if constexpr (PATHS) seg = hd <= -1e-3f;
    float hb0 = 1.0f - h.b1 - h.b2;
    V3 hpos = bary_point(q0, q1, q2, hb0, h.b1, h.b2);
    o = offset_ray(hpos, hn);
    float u = rnd(seed);
    float rd = a.rho_d[hp], rs = a.rho_s[hp];
    bool absorbed = false;
    if (u >= rd + rs) {
        absorbed = true;
    } else if (u >= rd) {
        V3 in = normalize3(d);
        float dn = 2.0f * dot3(hn, in);
        d = V3{in.x - dn * hn.x, in.y - dn * hn.y, in.z - dn * hn.z};
    } else {
        float r0 = rnd(seed), r1 = rnd(seed);
        (void)rnd(seed);
        d = tangent_to_world(cosine_hemisphere(r0, r1), hn);
    }
    if (!absorbed && depth != a.max_depth) return -1;
    // thermal_tracing_unified.glsl:578-583: row = absorber, column = source
    int hc = max_corner(hb0, h.b1, h.b2);
    uint32_t ab = a.bin_mode ? hp : (hc == 0 ? j0 : (hc == 1 ? j1 : j2));
    if (a.sink.tab)
        tab_add(a.sink.tab, a.sink.lg, tab_key(ab, src, a.sink.key_bits), qt, a.sink.overflow);
    else if (a.sink.acc) {
        const uint64_t o = (uint64_t)ab * a.ld + src;
        atomicAdd(a.sink.acc + o, qt);
        a.sink.touch[o >> kTouchShift] = 1;  // every writer stores the same byte
    }
    if (a.absorber) a.absorber[g] = ab;
    return absorbed ? S_ABS : S_FORCED;
}

// counts one answered segment and its outcome (constant indices only, so the
// counters stay in registers)
__device__ __forceinline__ void count_segment(uint32_t *cnt, int outcome)
{
    cnt[S_RAYS]++;
#pragma unroll
    for (int k = S_ABS; k < S_COUNT; k++) cnt[k] += outcome == k ? 1u : 0u;
}

// wave sums of the photon counters into stats (one wave per block); esc /
// drop: fixed-point weight of the photons that escaped / were dropped
__device__ __forceinline__ void flush_counts(unsigned long long *stats, const uint32_t *cnt,
                                             uint32_t nvis, uint32_t ntri, unsigned long long esc,
                                             unsigned long long drop)
{
    for (int off = 32; off > 0; off >>= 1) {
        esc += __shfl_down(esc, off, 64);
        drop += __shfl_down(drop, off, 64);
    }
    if (threadIdx.x == 0 && esc) atomicAdd(stats + S_COUNT + 2, esc);
    if (threadIdx.x == 0 && drop) atomicAdd(stats + S_COUNT + 3, drop);
#pragma unroll
    for (int s = 0; s < S_COUNT; s++) {
        uint32_t v = cnt[s];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (threadIdx.x == 0 && v) atomicAdd(stats + s, (unsigned long long)v);
    }
    unsigned long long nv64 = nvis, nt64 = ntri;
    for (int off = 32; off > 0; off >>= 1) {
        nv64 += __shfl_down(nv64, off, 64);
        nt64 += __shfl_down(nt64, off, 64);
    }
    if (threadIdx.x == 0 && (nv64 || nt64)) {
        atomicAdd(stats + S_COUNT, nv64);
        atomicAdd(stats + S_COUNT + 1, nt64);
    }
}

// One kernel for whole photon paths.  Each loop iteration traces one ray
// segment per lane; a lane whose photon ends starts its next photon in the
// following iteration.  A lane takes photons g, g + stride, ...; lanes of a
// wave start on consecutive photons of the same source triangle (coherent
// first-bounce rays).
template <int STACK>
__global__ __launch_bounds__(kTraceBlock, photon_waves(STACK)) void k_photons(PhotonArgs a)
{
    __shared__ uint32_t stk_n[STACK * kTraceBlock];
    __shared__ float stk_t[stack_t_entries(STACK) * kTraceBlock];
    uint32_t *sn = stk_n + threadIdx.x;
    float *stt = stk_t + threadIdx.x;
    uint32_t cnt[S_COUNT] = {0, 0, 0, 0, 0, 0, 0};
    unsigned long long q_esc = 0, q_drop = 0;  // energy accounting
    uint32_t nvis = 0, ntri = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kTraceBlock;
    uint64_t g = a.photon_begin + (uint64_t)blockIdx.x * kTraceBlock + threadIdx.x;
    bool live = false;
    uint32_t seed = 0, src = 0, depth = 0;
    unsigned long long qt = 0;
    V3 o{0.0f, 0.0f, 0.0f}, d{0.0f, 0.0f, 1.0f};
    for (;;) {
        if (!live) {
            // this lane's next photon; degenerate emitters are skipped
            while (g < a.photon_end) {
                uint32_t tri;
                float eb1, eb2;
                int ec;
                if (emit_photon(a, g, seed, o, d, src, tri, eb1, eb2, ec)) {
                    qt = a.q[tri];
                    depth = 0;
                    live = true;
                    break;
                }
                cnt[S_DEG]++;
                g += stride;
            }
            if (!live) break;
        }
        ++depth;
        Hit h;
        {
            Query q;
            if (!query_begin(q, a.tris, a.n_leaves, o, d, 1e-12f, 1e12f, ntri)) q.cur = kQueryDone;
            query_run_ref(q, a.nodes, a.tris, 1e-12f, sn, nvis, ntri);
            h = query_end(q, 1e12f);
        }
        bool unused;
        const int outcome = shade_segment<false>(a, h, g, src, qt, depth, seed, o, d, unused);
        count_segment(cnt, outcome);
        q_esc += outcome == S_MISS ? qt : 0ull;
        q_drop += (outcome == S_CLOSE || outcome == S_BACK) ? qt : 0ull;
        if (outcome >= 0) {
            live = false;
            g += stride;
        }
    }
    flush_counts(a.stats, cnt, nvis, ntri, q_esc, q_drop);
}

// Where a transport build keeps its photons' paths (IRTT_TRANSPORT_KEEP_PATHS).
struct PathRecs {
    RecBuf segs, edges;
    unsigned long long *first_dropped;  // smallest photon with a record that did not fit
};

// waves per SIMD of the path-keeping photon kernel: what its LDS stack
// admits, at most 5 (its records' state takes about 90 VGPRs)
constexpr int path_waves(int stack)
{
    return (160 * 1024 / (stack * 4 * kTraceBlock)) / 4 < 5
               ? (160 * 1024 / (stack * 4 * kTraceBlock)) / 4
               : 5;
}

// k_photons that also keeps the paths: per accepted segment the segment
// record, per hit within kKeptEdgeRatio of an edge the raw edge record, as
// the adjoint trace's path replay writes them for the photons it replays
// (the absorbing bin, known only at the photon's end, is taken from the
// per-photon record by the adjoint trace).  Every lane stays in the loop
// until its wave is done, since the record slots are taken by the whole wave
// together.
template <int STACK>
__global__ __launch_bounds__(kTraceBlock, path_waves(STACK)) void k_photons_paths(PhotonArgs a,
                                                                                  PathRecs pr)
{
    __shared__ uint32_t stk_n[STACK * kTraceBlock];
    __shared__ float stk_t[stack_t_entries(STACK) * kTraceBlock];
    uint32_t *sn = stk_n + threadIdx.x;
    float *stt = stk_t + threadIdx.x;
    (void)stt;
    uint32_t cnt[S_COUNT] = {0, 0, 0, 0, 0, 0, 0};
    unsigned long long q_esc = 0, q_drop = 0;
    uint32_t nvis = 0, ntri = 0;
    uint32_t s_chunk = kNoChunk, s_fill = 0, e_chunk = kNoChunk, e_fill = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kTraceBlock;
    uint64_t g = a.photon_begin + (uint64_t)blockIdx.x * kTraceBlock + threadIdx.x;
    bool live = false, rec = false;
    uint32_t seed = 0, src = 0, depth = 0, t = 0, prev_tri = 0;
    int prev_max = 0;
    float pb1 = 0.0f, pb2 = 0.0f;
    unsigned long long qt = 0;
    V3 o{0.0f, 0.0f, 0.0f}, d{0.0f, 0.0f, 1.0f};
    for (;;) {
        if (!live) {
            while (g < a.photon_end) {
                uint32_t tri;
                float eb1, eb2;
                int ec;
                if (emit_photon(a, g, seed, o, d, src, tri, eb1, eb2, ec)) {
                    qt = a.q[tri];
                    depth = 0;
                    live = true;
                    rec = true;
                    t = tri;
                    prev_tri = tri;
                    pb1 = eb1;
                    pb2 = eb2;
                    prev_max = ec;
                    break;
                }
                cnt[S_DEG]++;
                g += stride;
            }
        }
        if (__ballot(live) == 0) break;
        Query q{};
        q.cur = kQueryDone;
        if (live) {
            ++depth;
            if (!query_begin(q, a.tris, a.n_leaves, o, d, 1e-12f, 1e12f, ntri)) q.cur = kQueryDone;
        }
        query_run_ref(q, a.nodes, a.tris, 1e-12f, sn, nvis, ntri);
        Hit h{-1, 0.0f, 0.0f, 0.0f};
        bool seg = false;
        const V3 din = d;
        const uint64_t gp = g;
        if (live) {
            h = query_end(q, 1e12f);
            bool replayed;
            const int outcome = shade_segment<true>(a, h, g, src, qt, depth, seed, o, d, replayed);
            count_segment(cnt, outcome);
            q_esc += outcome == S_MISS ? qt : 0ull;
            q_drop += (outcome == S_CLOSE || outcome == S_BACK) ? qt : 0ull;
            seg = rec && replayed;
            if (!replayed) rec = false;
            if (outcome >= 0) {
                live = false;
                g += stride;
            }
        }
        const uint32_t hp = (uint32_t)h.prim;
        {
            bool ok;
            const uint64_t slot = chunk_slot(seg, s_chunk, s_fill, pr.segs, ok);
            if (seg) {
                if (ok) {
                    SegRec *dst = static_cast<SegRec *>(pr.segs.p) + slot;
                    dst->r0 = make_uint4(prev_tri, hp, src, 0xffffffffu);
                    dst->r1 = make_float4(pb1, pb2, h.b1, h.b2);
                    dst->r2 = make_uint4((uint32_t)gp, t, depth == 1 ? 1u : 0u, 0u);
                } else {
                    atomicMin(pr.first_dropped, (unsigned long long)gp);
                }
            }
        }
        const float hb0 = 1.0f - h.b1 - h.b2;
        bool edge = false;
        if (seg) {
            const int opp = min_corner(hb0, h.b1, h.b2);
            edge = (opp == 0 ? hb0 : (opp == 1 ? h.b1 : h.b2)) <= kKeptEdgeRatio;
        }
        {
            bool ok;
            const uint64_t slot = chunk_slot(edge, e_chunk, e_fill, pr.edges, ok);
            if (edge) {
                if (ok) {
                    EdgeRec *dst = static_cast<EdgeRec *>(pr.edges.p) + slot;
                    dst->r0 = make_float4(din.x, din.y, din.z, h.t);
                    dst->r1 = make_float4(h.b1, h.b2, __uint_as_float(hp), __uint_as_float(t));
                    dst->r2 = make_float4(__uint_as_float(prev_tri),
                                          __uint_as_float((uint32_t)prev_max), __uint_as_float(src),
                                          __uint_as_float(0xffffffffu));
                    dst->r3 = make_uint4(0u, 0u, 0u, (uint32_t)gp);
                } else {
                    atomicMin(pr.first_dropped, (unsigned long long)gp);
                }
            }
        }
        if (seg) {
            prev_tri = hp;
            pb1 = h.b1;
            pb2 = h.b2;
            prev_max = max_corner(hb0, h.b1, h.b2);
        }
    }
    chunk_close(s_chunk, s_fill, pr.segs);
    chunk_close(e_chunk, e_fill, pr.edges);
    flush_counts(a.stats, cnt, nvis, ntri, q_esc, q_drop);
}

// photon kernel grid: at most this many waves per CU, lanes striding over
// the photons.  Smaller per-wave shares even out the end of the launch
// (city block, 2^26 photons: 24 -> 13.3 ms, 64 -> 13.0-13.1, 96 -> 12.7,
// 256 -> 12.09-12.14, 512 -> 12.10-12.13, 1024 -> 14.4)
#ifndef IRTT_PHOTON_WAVES_PER_CU
#define IRTT_PHOTON_WAVES_PER_CU 256
#endif

// T[i][j] = (i == j ? diag_i : 0) - acc[i][j] * Q in FP64, stored FP32, in two
// passes.  k_fill_T writes every entry as if its accumulator were zero
// (FP32(diag_i) on the diagonal, +0 elsewhere: what the expression gives for
// acc = 0) with stores alone; k_finalize then recomputes the entries of the
// chunks the photon kernel marked (every nonzero accumulator lies in one),
// adds their integer column sums, and gathers the stored diagonal into tdiag.
constexpr int kFillRows = 16;   // rows per k_fill_T block (1024 columns wide)
__global__ __launch_bounds__(256) void k_fill_T(const double *__restrict__ diag, uint32_t nb,
                                                uint32_t ld, float *__restrict__ T,
                                                float *__restrict__ tdiag)
{
    const uint32_t j = blockIdx.x * 1024 + threadIdx.x * 4;  // ld is a multiple of 64
    if (j >= ld) return;
    const uint32_t r0 = blockIdx.y * kFillRows, r1 = min(r0 + kFillRows, nb);
    for (uint32_t i = r0; i < r1; i++) {
        float4 f = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (i - j < 4u) {
            const float d = (float)(diag[i] - 0.0);
            tdiag[i] = d;
            if (i == j) f.x = d;
            else if (i == j + 1) f.y = d;
            else if (i == j + 2) f.z = d;
            else f.w = d;
        }
        *reinterpret_cast<float4 *>(T + (size_t)i * ld + j) = f;
    }
}

constexpr int kFinRows = 256;                          // rows per tile
constexpr int kFinCols = 1024;                         // columns per tile
constexpr int kFinChunks = kFinCols >> kTouchShift;    // chunks per tile row (64)
constexpr int kFinPer = 4;                             // chunk quarters per thread and round
static_assert(kFinRows == 256 && kFinChunks == 64, "list entries are (row << 6) | chunk");
__global__ __launch_bounds__(256) void k_finalize(const unsigned long long *__restrict__ acc,
                                                  const uint8_t *__restrict__ touch,
                                                  const double *__restrict__ diag, uint32_t nb,
                                                  uint32_t ld, double Q, float *__restrict__ T,
                                                  unsigned long long *colacc,
                                                  float *__restrict__ tdiag)
{
    // the tile's marked chunks listed in LDS ((row << 6) | chunk, in any
    // order: the entries and the integer column sums do not depend on it);
    // then four threads per chunk, kFinPer chunks per thread in flight
    __shared__ uint16_t list[kFinRows * kFinChunks];
    __shared__ unsigned long long csum[kFinCols];
    __shared__ uint32_t n_list;
    const uint32_t r0 = blockIdx.y * kFinRows, r1 = min(r0 + kFinRows, nb);
    const uint32_t cpr = ld >> kTouchShift;  // a multiple of 4
    const uint32_t c0 = blockIdx.x * kFinChunks;
    const uint32_t j0 = blockIdx.x * kFinCols;
    for (uint32_t k = threadIdx.x; k < kFinCols; k += blockDim.x) csum[k] = 0ull;
    if (threadIdx.x == 0) n_list = 0;
    __syncthreads();
    {
        const uint32_t r = threadIdx.x, i = r0 + r;
        const uint32_t *t4 = reinterpret_cast<const uint32_t *>(touch);
        uint32_t mw[kFinChunks / 4];
#pragma unroll
        for (uint32_t w = 0; w < kFinChunks / 4; w++) {
            const uint32_t c = c0 + 4 * w;
            mw[w] = (i < r1 && c < cpr) ? t4[((size_t)i * cpr + c) >> 2] : 0u;
        }
#pragma unroll
        for (uint32_t w = 0; w < kFinChunks / 4; w++) {
            if (!mw[w]) continue;
#pragma unroll
            for (uint32_t b = 0; b < 4; b++)
                if ((mw[w] >> (8 * b)) & 0xffu)
                    list[atomicAdd(&n_list, 1u)] = (uint16_t)((r << 6) | (4 * w + b));
        }
    }
    __syncthreads();
    const uint32_t n = n_list;
    const uint32_t qc = (threadIdx.x & 3) * 4;  // this thread's four columns of a chunk
    for (uint32_t e0 = threadIdx.x >> 2; e0 < n; e0 += (blockDim.x / 4) * kFinPer) {
        ulonglong2 va[kFinPer], vb[kFinPer];
        uint32_t ent[kFinPer];
#pragma unroll
        for (int k = 0; k < kFinPer; k++) {
            const uint32_t e = e0 + k * (blockDim.x / 4);
            ent[k] = e < n ? (uint32_t)list[e] : ~0u;
            va[k] = vb[k] = make_ulonglong2(0ull, 0ull);
            if (ent[k] != ~0u) {
                const uint32_t i = r0 + (ent[k] >> 6), j = j0 + ((ent[k] & 63u) << kTouchShift) + qc;
                const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(acc + (size_t)i * ld + j);
                va[k] = p[0];
                vb[k] = p[1];
            }
        }
#pragma unroll
        for (int k = 0; k < kFinPer; k++) {
            if (ent[k] == ~0u) continue;
            const uint32_t i = r0 + (ent[k] >> 6), jt = ((ent[k] & 63u) << kTouchShift) + qc;
            const uint32_t j = j0 + jt;
            const unsigned long long v[4] = {va[k].x, va[k].y, vb[k].x, vb[k].y};
            float f[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                if (v[c]) atomicAdd(&csum[jt + c], v[c]);
                double val = (i == j + c ? diag[i] : 0.0) - (double)v[c] * Q;
                f[c] = j + c < nb ? (float)val : 0.0f;
                if (i == j + c) tdiag[i] = f[c];
            }
            *reinterpret_cast<float4 *>(T + (size_t)i * ld + j) = make_float4(f[0], f[1], f[2], f[3]);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < kFinCols; k += blockDim.x)
        if (csum[k] && j0 + k < nb) atomicAdd(colacc + j0 + k, csum[k]);
}

// Zeroes the accumulator chunks the last build marked, and their marks:
// afterwards every accumulator and mark is zero.  With T given, the same
// chunks of T (and of the stored diagonal) are set back to what k_fill_T
// writes, so T again holds +0 off the diagonal and FP32(diag_i) on it
// everywhere.  A block takes 1024 words of marks (4096 chunks) at a time,
// lists the marked chunks in LDS, and clears each with eight lanes (128
// contiguous bytes of accumulators, 64 of T).  n4 = chunks / 4 words of marks
// (the chunk count nb * ld / 16 is a multiple of 4 since ld is a multiple of
// 64, and below 2^32: dense operators stay far below 2^36 entries).
constexpr int kClearWords = 1024;  // words of marks per block and round (4 per thread)
static_assert(kTouchShift == 4, "eight lanes of 16 B (accumulators) / 8 B (T) per chunk");
__global__ __launch_bounds__(256) void k_clear_marked(unsigned long long *__restrict__ acc,
                                                      uint32_t *__restrict__ touch4, size_t n4,
                                                      float *__restrict__ T,
                                                      const double *__restrict__ diag,
                                                      float *__restrict__ tdiag, uint32_t ld)
{
    __shared__ uint32_t list[kClearWords * 4];
    __shared__ uint32_t n_list;
    const uint32_t cpr = ld >> kTouchShift;
    for (size_t base = (size_t)blockIdx.x * kClearWords; base < n4;
         base += (size_t)gridDim.x * kClearWords) {
        if (threadIdx.x == 0) n_list = 0;
        __syncthreads();
        const size_t w0 = base + threadIdx.x * 4;
        uint32_t mw[4];
        if (w0 + 4 <= n4) {
            const uint4 m = *reinterpret_cast<const uint4 *>(touch4 + w0);
            mw[0] = m.x;
            mw[1] = m.y;
            mw[2] = m.z;
            mw[3] = m.w;
        } else {
#pragma unroll
            for (int w = 0; w < 4; w++) mw[w] = w0 + w < n4 ? touch4[w0 + w] : 0u;
        }
#pragma unroll
        for (int w = 0; w < 4; w++) {
            if (!mw[w]) continue;
#pragma unroll
            for (int b = 0; b < 4; b++)
                if ((mw[w] >> (8 * b)) & 0xffu)
                    list[atomicAdd(&n_list, 1u)] = (uint32_t)((w0 + w) * 4 + b);
            touch4[w0 + w] = 0u;
        }
        __syncthreads();
        const uint32_t n = n_list;
        const uint32_t l = threadIdx.x & 7;
        for (uint32_t e = threadIdx.x >> 3; e < n; e += blockDim.x / 8) {
            const uint32_t k = list[e];
            reinterpret_cast<ulonglong2 *>(acc + ((size_t)k << kTouchShift))[l] =
                make_ulonglong2(0ull, 0ull);
            if (T) {
                const uint32_t row = k / cpr, col = ((k - row * cpr) << kTouchShift) + 2 * l;
                float2 f = make_float2(0.0f, 0.0f);
                if (row - col < 2u) {
                    const float d = (float)(diag[row] - 0.0);
                    if (row == col) f.x = d;
                    else f.y = d;
                    tdiag[row] = d;
                }
                *reinterpret_cast<float2 *>(T + (size_t)row * ld + col) = f;
            }
        }
        __syncthreads();  // the list is rebuilt next round
    }
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Energy conservation of a build, one block: every photon of triangle t
// carries q_t, and ends absorbed (in the operator's column sums), escaped or
// dropped, so  per * sum_t q_t = sum_j colacc_j + escaped + dropped  exactly
// in integers; in FP64 (W per unit emission factor): emitted
// sum_i sigma eps_i A_i against Q times each part.  out: emitted, absorbed,
// escaped, dropped (W), |emitted - parts| / emitted, integer identity (1/0).
constexpr int kEnergyThreads = 1024;
__global__ __launch_bounds__(kEnergyThreads) void k_energy(
    const unsigned long long *__restrict__ q, uint32_t nt, uint32_t per,
    const unsigned long long *__restrict__ colacc, const double *__restrict__ diag, uint32_t nb,
    const unsigned long long *__restrict__ stats, double Q, double *__restrict__ out)
{
    __shared__ unsigned long long su[kEnergyThreads], sa[kEnergyThreads];
    __shared__ double sd[kEnergyThreads];
    // four independent partial sums per thread (loads in flight), added in a
    // fixed order: the FP64 total does not depend on timing
    unsigned long long uq[4] = {0, 0, 0, 0}, ua[4] = {0, 0, 0, 0};
    double em[4] = {0.0, 0.0, 0.0, 0.0};
    constexpr uint32_t B = kEnergyThreads;
    for (uint32_t t0 = threadIdx.x; t0 < nt; t0 += 4 * B) {
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (t0 + k * B < nt) uq[k] += q[t0 + k * B];
    }
    for (uint32_t j0 = threadIdx.x; j0 < nb; j0 += 4 * B) {
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (j0 + k * B < nb) {
                ua[k] += colacc[j0 + k * B];
                em[k] += diag[j0 + k * B];
            }
    }
    su[threadIdx.x] = (uq[0] + uq[1]) + (uq[2] + uq[3]);
    sa[threadIdx.x] = (ua[0] + ua[1]) + (ua[2] + ua[3]);
    sd[threadIdx.x] = (em[0] + em[1]) + (em[2] + em[3]);
    __syncthreads();
    for (int off = kEnergyThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            su[threadIdx.x] += su[threadIdx.x + off];
            sa[threadIdx.x] += sa[threadIdx.x + off];
            sd[threadIdx.x] += sd[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const unsigned long long emitted = su[0] * (unsigned long long)per;
        const unsigned long long esc = stats[S_COUNT + 2], drop = stats[S_COUNT + 3];
        const double e = sd[0], ab = (double)sa[0] * Q, es = (double)esc * Q,
                     dr = (double)drop * Q;
        out[0] = e;
        out[1] = ab;
        out[2] = es;
        out[3] = dr;
        out[4] = e > 0.0 ? fabs(e - (ab + es + dr)) / e : 0.0;
        out[5] = emitted == sa[0] + esc + drop ? 1.0 : 0.0;
    }
}

}  // namespace

int stack_size_for(uint32_t depth)
{
    if (depth + 1 <= 32) return 32;
    if (depth + 1 <= 64) return 64;
    if (depth + 1 <= 128) return 128;
    throw Error(IRTT_ERR_LIMIT, "LBVH depth " + std::to_string(depth) +
                                    " exceeds the traversal stack (128 entries)");
}

void trace_rays(Scene &s, const float *rays, uint32_t n, float tmin, float tmax, int32_t *prim,
                float *t, float *b1, float *b2)
{
    if (!s.bvh.valid) bvh_build(s);
    if (n == 0) return;
    DBuf<float> d_rays, d_t, d_b1, d_b2;
    DBuf<int32_t> d_prim;
    d_rays.alloc(6 * (size_t)n);
    d_t.alloc(n); d_b1.alloc(n); d_b2.alloc(n); d_prim.alloc(n);
    hipStream_t st = s.stream;
    IRTT_HIP(hipMemcpyAsync(d_rays.p, rays, d_rays.bytes(), hipMemcpyHostToDevice, st));
    int S = stack_size_for(s.bvh.max_depth);
    dim3 g(cdiv(n, kTraceBlock)), b(kTraceBlock);
    const float4 *nd = s.bvh.nodes.p, *tr = s.bvh.tris.p;
    if (S == 32)
        k_trace_rays<32><<<g, b, 0, st>>>(nd, tr, s.bvh.n_leaves, d_rays.p, n, tmin, tmax, d_prim.p,
                                          d_t.p, d_b1.p, d_b2.p);
    else if (S == 64)
        k_trace_rays<64><<<g, b, 0, st>>>(nd, tr, s.bvh.n_leaves, d_rays.p, n, tmin, tmax, d_prim.p,
                                          d_t.p, d_b1.p, d_b2.p);
    else
        k_trace_rays<128><<<g, b, 0, st>>>(nd, tr, s.bvh.n_leaves, d_rays.p, n, tmin, tmax,
                                           d_prim.p, d_t.p, d_b1.p, d_b2.p);
    IRTT_HIP(hipGetLastError());
    IRTT_HIP(hipMemcpyAsync(prim, d_prim.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    IRTT_HIP(hipMemcpyAsync(t, d_t.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
    IRTT_HIP(hipMemcpyAsync(b1, d_b1.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
    IRTT_HIP(hipMemcpyAsync(b2, d_b2.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
    IRTT_HIP(hipStreamSynchronize(st));
}

// ---- ray queries without allocation -------------------------------------------
// Staging for irtt_trace_rays_noalloc: the rays of one call (6 floats each)
// and their answers (prim, t, b1, b2: 4 words each), for up to
// kNoAllocRays rays, in one statically sized device array (10 MiB; each
// device holds its own copy until the process ends).  Host code never uses
// the array's name as an address: it asks the runtime for the device address
// of this device's copy once and keeps it.
constexpr uint32_t kNoAllocRays = 1u << 18;
static __device__ __align__(256) float scratch[10 * (size_t)kNoAllocRays];

namespace {
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void *g_scratch_dev[kMaxDevices] = {};

// device address of the current device's scratch (looked up on its first use)
float *scratch_address()
{
    int dev = 0;
    IRTT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(IRTT_ERR_LIMIT, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(IRTT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                          hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    return static_cast<float *>(g_scratch_dev[dev]);
}
}  // namespace

uint32_t trace_rays_noalloc_max() { return kNoAllocRays; }

// irtt_trace_rays through the static staging array: no device allocation.
// The BVH must exist (its build allocates).  Calls that overlap on one device
// share the array and race.
void trace_rays_noalloc(Scene &s, const float *rays, uint32_t n, float tmin, float tmax,
                        int32_t *prim, float *t, float *b1, float *b2)
{
    if (!s.bvh.valid)
        throw Error(IRTT_ERR_STATE, "irtt_trace_rays_noalloc: build the BVH first "
                                    "(irtt_build_bvh); this entry point allocates nothing");
    if (n > kNoAllocRays)
        throw Error(IRTT_ERR_LIMIT, "irtt_trace_rays_noalloc: " + std::to_string(n) +
                                        " rays exceed the maximum of " +
                                        std::to_string(kNoAllocRays) + " per call");
    if (n == 0) return;
    float *base = scratch_address();
    float *d_rays = base;
    int32_t *d_prim = reinterpret_cast<int32_t *>(base + 6 * (size_t)kNoAllocRays);
    float *d_t = base + 7 * (size_t)kNoAllocRays;
    float *d_b1 = base + 8 * (size_t)kNoAllocRays;
    float *d_b2 = base + 9 * (size_t)kNoAllocRays;
    hipStream_t st = s.stream;
    IRTT_HIP(hipMemcpyAsync(d_rays, rays, 6 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, st));
    const int S = stack_size_for(s.bvh.max_depth);
    dim3 g(cdiv(n, kTraceBlock)), b(kTraceBlock);
    const float4 *nd = s.bvh.nodes.p, *tr = s.bvh.tris.p;
    if (S == 32)
        k_trace_rays<32><<<g, b, 0, st>>>(nd, tr, s.bvh.n_leaves, d_rays, n, tmin, tmax, d_prim, d_t,
                                          d_b1, d_b2);
    else if (S == 64)
        k_trace_rays<64><<<g, b, 0, st>>>(nd, tr, s.bvh.n_leaves, d_rays, n, tmin, tmax, d_prim, d_t,
                                          d_b1, d_b2);
    else
        k_trace_rays<128><<<g, b, 0, st>>>(nd, tr, s.bvh.n_leaves, d_rays, n, tmin, tmax, d_prim,
                                           d_t, d_b1, d_b2);
    IRTT_HIP(hipGetLastError());
    IRTT_HIP(hipMemcpyAsync(prim, d_prim, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    IRTT_HIP(hipMemcpyAsync(t, d_t, n * sizeof(float), hipMemcpyDeviceToHost, st));
    IRTT_HIP(hipMemcpyAsync(b1, d_b1, n * sizeof(float), hipMemcpyDeviceToHost, st));
    IRTT_HIP(hipMemcpyAsync(b2, d_b2, n * sizeof(float), hipMemcpyDeviceToHost, st));
    IRTT_HIP(hipStreamSynchronize(st));
}

static uint32_t ceil_log2_u64(uint64_t x)
{
    uint32_t l = 0;
    while (l < 64 && ((uint64_t)1 << l) < x) l++;
    return l;
}

// Waves of the path-keeping photon kernel: at most 128 per CU, at least 16
// photons per lane.
static uint64_t path_waves_total(int ncu, uint64_t total)
{
    return std::max<uint64_t>(
        1, std::min<uint64_t>((uint64_t)ncu * 128, (total + 16 * kTraceBlock - 1) / (16 * kTraceBlock)));
}

// Launches the photon kernel over every photon of the scene on the scene's
// stream (timed: between events ev0 and ev1): bins into sink (none: absorbers
// only), records absorbing bins into absorber (or not), counts into stats,
// writes path records into pr (or not).
static void launch_photons(Scene &s, const irtt_transport_params &p, uint32_t per,
                           uint32_t max_depth, uint64_t total, const PhotonSink &sink,
                           uint32_t *absorber, unsigned long long *stats, bool timed,
                           const PathRecs *pr = nullptr)
{
    hipStream_t stm = s.stream;
    PhotonArgs a{};
    a.xyz = s.d_xyz.p;
    a.idx = s.d_idx.p;
    a.rho_d = s.d_rho_d.p;
    a.rho_s = s.d_rho_s.p;
    a.flags = s.d_flags.p;
    a.q = reinterpret_cast<const unsigned long long *>(s.d_q.p);
    a.nodes = s.bvh.nodes.p;
    a.tris = s.bvh.tris.p;
    a.n_leaves = s.bvh.n_leaves;
    a.nt = s.nt;
    a.nb = s.nb;
    a.ld = s.ld;
    a.bin_mode = s.bin_mode;
    a.threads_per_tri = p.threads_per_tri;
    a.rays_per_thread = p.rays_per_thread;
    a.per_tri = per;
    a.batch_seed = p.batch_seed;
    a.max_depth = max_depth;
    a.tiny_dist = p.tiny_dist;
    a.photon_begin = 0;
    a.photon_end = total;
    a.sink = sink;
    a.absorber = absorber;
    a.stats = stats;

    int dev = 0, ncu = 256;
    IRTT_HIP(hipGetDevice(&dev));
    IRTT_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // The photon kernel's occupancy is bound by its LDS stack: 24 entries
    // (6 KiB per wave) admit 24 waves per CU against 20 for 32 entries, for
    // trees up to 23 levels (the 51.8k-triangle city block has 21).  Measured
    // on that block: 17.2 against 18.9 ms.  (A 16-entry LDS stack spilling
    // deeper entries to global memory, with 64 VGPRs for 32 waves, measured
    // 17.4 ms: the register spills cost what the occupancy gains.)
    const int S = s.bvh.max_depth + 1 <= 24 ? 24 : stack_size_for(s.bvh.max_depth);
    // enough waves to cover every CU several times; lanes loop over photons
    uint64_t want = cdiv(total, kTraceBlock);
    uint32_t grid = (uint32_t)std::min<uint64_t>(want, (uint64_t)ncu * IRTT_PHOTON_WAVES_PER_CU);
    if (grid == 0) grid = 1;
    if (timed) IRTT_HIP(hipEventRecord(s.ev0, stm));
    if (pr) {
        // the path replay's grid (each wave's last record chunks are partly
        // empty: fewer, longer waves)
        grid = (uint32_t)path_waves_total(ncu, total);
        if (S == 24) k_photons_paths<24><<<grid, kTraceBlock, 0, stm>>>(a, *pr);
        else if (S == 32) k_photons_paths<32><<<grid, kTraceBlock, 0, stm>>>(a, *pr);
        else if (S == 64) k_photons_paths<64><<<grid, kTraceBlock, 0, stm>>>(a, *pr);
        else k_photons_paths<128><<<grid, kTraceBlock, 0, stm>>>(a, *pr);
    } else if (S == 24) k_photons<24><<<grid, kTraceBlock, 0, stm>>>(a);
    else if (S == 32) k_photons<32><<<grid, kTraceBlock, 0, stm>>>(a);
    else if (S == 64) k_photons<64><<<grid, kTraceBlock, 0, stm>>>(a);
    else k_photons<128><<<grid, kTraceBlock, 0, stm>>>(a);
    IRTT_HIP(hipGetLastError());
    if (timed) IRTT_HIP(hipEventRecord(s.ev1, stm));
}

uint64_t kept_path_bytes(uint64_t photons)
{
    const uint64_t waves = path_waves_total(256, photons);
    const uint64_t sch = (photons + kChunk - 1) / kChunk + waves + 1;
    const uint64_t ech = (photons / 2 + kChunk - 1) / kChunk + waves + 1;
    return sch * kChunk * sizeof(SegRec) + ech * kChunk * sizeof(EdgeRec) + (sch + ech) * 4 + 24;
}

// A forward pass that records each photon's absorbing bin and bins nothing:
// the adjoint trace's pass 0 when the build did not keep its record.
void absorber_pass(Scene &s, uint32_t *absorber, uint64_t &rays, uint64_t &segments)
{
    const irtt_transport_params &p = s.tparams;
    const uint32_t per = p.threads_per_tri * p.rays_per_thread;
    const uint32_t max_depth =
        (p.max_depth == 0 || p.max_depth > kMaxDepthDefault) ? kMaxDepthDefault : p.max_depth;
    const uint64_t total = (uint64_t)per * s.nt;
    hipStream_t stm = s.stream;
    DBuf<unsigned long long> d_st;
    d_st.alloc(16);
    IRTT_HIP(hipMemsetAsync(d_st.p, 0, d_st.bytes(), stm));
    IRTT_HIP(hipMemsetAsync(absorber, 0xff, total * sizeof(uint32_t), stm));
    launch_photons(s, p, per, max_depth, total, PhotonSink{}, absorber, d_st.p, false);
    unsigned long long hs[16];
    IRTT_HIP(hipMemcpyAsync(hs, d_st.p, sizeof(hs), hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipStreamSynchronize(stm));
    rays = hs[S_RAYS];
    segments = hs[S_RAYS] - hs[S_MISS] - hs[S_CLOSE] - hs[S_BACK];
}

void transport_build(Scene &s, const irtt_transport_params &p, irtt_transport_stats &st)
{
    if (p.threads_per_tri == 0 || p.rays_per_thread == 0)
        throw Error(IRTT_ERR_ARG, "threads_per_tri and rays_per_thread must be positive");
    uint64_t per64 = (uint64_t)p.threads_per_tri * p.rays_per_thread;
    if (per64 > 0xffffffffull)
        throw Error(IRTT_ERR_LIMIT, "photons per triangle exceed 2^32");
    const uint32_t per = (uint32_t)per64;
    const uint32_t max_depth =
        (p.max_depth == 0 || p.max_depth > kMaxDepthDefault) ? kMaxDepthDefault : p.max_depth;
    if (p.layout != IRTT_LAYOUT_DENSE && p.layout != IRTT_LAYOUT_SPARSE)
        throw Error(IRTT_ERR_ARG, "unknown transport layout");
    const bool sparse = p.layout == IRTT_LAYOUT_SPARSE;
    if (!s.bvh.valid) bvh_build(s);
    const uint32_t nb = s.nb;
    const uint32_t ld = s.ld;
    const size_t nb2 = (size_t)nb * ld;
    if (nb == 0) throw Error(IRTT_ERR_STATE, "empty scene");
    op_reset(s);
    s.has_transport = false;
    if (sparse) {
        s.d_acc.release();
        s.d_touch.release();
        s.d_T.release();
        s.acc_marked = false;
        s.t_fill_gen = ~0ull;
    } else {
        s.sp.reset();
    }

    hipStream_t stm = s.stream;
    const uint64_t total = per64 * s.nt;
    // the hash table's clear (sparse layout) runs while the host derives the
    // photon weights
    PhotonSink sink;
    if (sparse) sparse_table_begin(s, total, sink);

    // photon weights, thermal_tracing_unified.glsl:287,309-310, and their
    // fixed-point form q_t = rint(w_t / Q), Q = max w * 2^-s; with the bin
    // diagonals, derived again only when the scene's assembly or the photon
    // counts changed since the last build (repeated builds of one scene)
    const bool same_weights = s.wq_gen == s.assembly_gen && s.wq_per == per &&
                              s.wq_total == total && s.h_q.size() == s.nt &&
                              s.h_diag.size() == nb;
    if (!same_weights) {
        s.wq_gen = ~0ull;
        std::vector<float> w(s.nt);
        float wmax = 0.0f;
        for (uint32_t t = 0; t < s.nt; t++) {
            float wt = s.tri_area[t] / (float)per;
            wt = wt * (kSigma * s.eps[t]);
            w[t] = wt;
            wmax = std::max(wmax, wt);
        }
        int sh = 62 - (int)ceil_log2_u64(total ? total : 1);
        if (sh > 40) sh = 40;
        s.wq_quantum = wmax > 0.0f ? std::ldexp((double)wmax, -sh) : 1.0;
        s.h_q.resize(s.nt);
        for (uint32_t t = 0; t < s.nt; t++) {
            double r = wmax > 0.0f ? std::ldexp((double)w[t] / (double)wmax, sh) : 0.0;
            s.h_q[t] = (uint64_t)std::rint(r > 0.0 ? r : 0.0);
        }
        s.h_diag.resize(nb);
        for (uint32_t i = 0; i < nb; i++) s.h_diag[i] = kSigmaD * s.bin_eps_area[i];
    }
    const double Q = s.wq_quantum;
    const std::vector<double> &diag = s.h_diag;
    s.quantum = Q;
    s.tparams = p;

    // weights and diagonals are uploaded when they changed or their device
    // buffers are new (the host copies stay in the scene, unchanged until the
    // stream has been synchronised)
    const uint64_t *prev_q = s.d_q.p;
    s.d_q.alloc(s.nt);
    if (s.d_q.p != prev_q || !same_weights)
        IRTT_HIP(hipMemcpyAsync(s.d_q.p, s.h_q.data(), s.nt * sizeof(uint64_t),
                                hipMemcpyHostToDevice, stm));
    const double *prev_diag = s.d_diag.p;
    s.d_diag.alloc(nb);
    s.d_colacc.alloc(nb);
    s.d_stats.alloc(16);
    if (s.d_diag.p != prev_diag || !same_weights)
        IRTT_HIP(hipMemcpyAsync(s.d_diag.p, diag.data(), nb * sizeof(double), hipMemcpyHostToDevice,
                                stm));
    s.wq_gen = s.assembly_gen;
    s.wq_per = per;
    s.wq_total = total;

    // Dense layout.  A repeated build clears only the accumulator chunks the
    // previous one marked; fresh (or reallocated) accumulators are cleared
    // whole.  T (and its stored diagonal) outside the marked chunks keeps
    // what k_fill_T wrote for the scene's assembly: a repeated build of the
    // same assembly sets back only the previous build's marked chunks of T,
    // any other build writes the whole of T.
    bool t_kept = false;
    if (!sparse) {
        const void *prev_acc = s.d_acc.p, *prev_touch = s.d_touch.p, *prev_T = s.d_T.p,
                   *prev_tdiag = s.d_tdiag.p;
        s.d_acc.alloc(nb2);
        s.d_touch.alloc(nb2 >> kTouchShift);
        s.d_T.alloc(nb2);
        s.d_tdiag.alloc(nb);
        const size_t n4 = s.d_touch.bytes() / 4;
        const bool marked = s.acc_marked && s.d_acc.p == prev_acc && s.d_touch.p == prev_touch;
        t_kept = marked && s.d_T.p == prev_T && s.d_tdiag.p == prev_tdiag &&
                 s.t_fill_gen == s.assembly_gen;
        if (marked) {
            k_clear_marked<<<std::min<uint32_t>(cdiv(n4, kClearWords), 2048), 256, 0, stm>>>(
                s.d_acc.p, reinterpret_cast<uint32_t *>(s.d_touch.p), n4,
                t_kept ? s.d_T.p : nullptr, s.d_diag.p, s.d_tdiag.p, ld);
            IRTT_HIP(hipGetLastError());
        } else {
            IRTT_HIP(hipMemsetAsync(s.d_acc.p, 0, s.d_acc.bytes(), stm));
            IRTT_HIP(hipMemsetAsync(s.d_touch.p, 0, s.d_touch.bytes(), stm));
        }
        // neither is trusted again until this build has completed
        s.acc_marked = false;
        s.t_fill_gen = ~0ull;
        sink.acc = s.d_acc.p;
        sink.touch = s.d_touch.p;
    }
    IRTT_HIP(hipMemsetAsync(s.d_colacc.p, 0, s.d_colacc.bytes(), stm));
    IRTT_HIP(hipMemsetAsync(s.d_stats.p, 0, s.d_stats.bytes(), stm));
    // absorbing bin of every photon (4 B each) so the adjoint trace replays
    // each photon once instead of twice; not kept above 2^31 photons or
    // when HBM is short (the adjoint then finds absorbers with a pass of its own)
    PathStore &ps = s.paths;
    ps.valid = false;
    if (!(p.flags & (IRTT_TRANSPORT_KEEP_ABSORBERS | IRTT_TRANSPORT_KEEP_PATHS)) ||
        total > (1ull << 31)) {
        s.d_absorber.release();
    } else {
        try {
            s.d_absorber.alloc(total);
            IRTT_HIP(hipMemsetAsync(s.d_absorber.p, 0xff, s.d_absorber.bytes(), stm));
        } catch (const Error &e) {
            if (e.code != IRTT_ERR_NOMEM) throw;
            s.d_absorber.release();
        }
    }

    // the paths (with the absorbers): record buffers sized for the accepted
    // segments of the last build of the same photon count (else one per
    // photon) and half as many edge records, plus each wave's partial chunks,
    // within 30 % of the free device memory; records that do not fit are
    // replayed by the adjoint trace
    PathRecs pr{};
    const bool keep_paths = (p.flags & IRTT_TRANSPORT_KEEP_PATHS) && s.d_absorber.p;
    if (keep_paths) {
        int dev = 0, ncu = 256;
        IRTT_HIP(hipGetDevice(&dev));
        IRTT_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const uint64_t waves = path_waves_total(ncu, total);
        const uint64_t seg_est =
            (ps.photons == total && s.fwd_segments > 0) ? s.fwd_segments + s.fwd_segments / 32
                                                        : total;
        double budget = std::max(1.0 * (1u << 30), 0.3 * ((double)dev_available() + (double)ps.bytes()));
        uint64_t sch = 0, ech = 0;
        for (;;) {
            sch = (seg_est + kChunk - 1) / kChunk + waves + 1;
            ech = (seg_est / 2 + kChunk - 1) / kChunk + waves + 1;
            const double need = (double)kChunk * (sch * sizeof(SegRec) + ech * sizeof(EdgeRec));
            if (need > budget) {
                const double f = budget / need;
                sch = std::max<uint64_t>(2 * waves + 1, (uint64_t)(sch * f));
                ech = std::max<uint64_t>(2 * waves + 1, (uint64_t)(ech * f));
            }
            try {
                ps.segs.alloc(sch * kChunk * 3);
                ps.edges.alloc(ech * kChunk * 4);
                ps.fills.alloc(sch + ech);
                ps.ctl.alloc(3);
                break;
            } catch (const Error &e) {
                if (e.code != IRTT_ERR_NOMEM) throw;
                ps.release();
                if (budget <= 1.0 * (1u << 30)) break;
                budget = std::max(1.0 * (1u << 30), 0.5 * budget);
            }
        }
        if (ps.ctl.p) {
            static const unsigned long long ctl0[3] = {0ull, 0ull, ~0ull};
            IRTT_HIP(hipMemcpyAsync(ps.ctl.p, ctl0, sizeof(ctl0), hipMemcpyHostToDevice, stm));
            ps.seg_chunks = sch;
            ps.edge_chunks = ech;
            pr.segs = RecBuf{ps.segs.p, (uint32_t)sch, ps.fills.p, ps.ctl.p};
            pr.edges = RecBuf{ps.edges.p, (uint32_t)ech, ps.fills.p + sch, ps.ctl.p + 1};
            pr.first_dropped = ps.ctl.p + 2;
        }
    } else {
        ps.release();
    }
    launch_photons(s, p, per, max_depth, total, sink, s.d_absorber.p, s.d_stats.p, true,
                   pr.first_dropped ? &pr : nullptr);

    // finalize timed from ev1 to ev2; both times are read after the last sync
    if (sparse) {
        sparse_table_finish(s, sink);
    } else {
        if (!t_kept) {
            k_fill_T<<<dim3(cdiv(ld, 1024), cdiv(nb, kFillRows)), 256, 0, stm>>>(
                s.d_diag.p, nb, ld, s.d_T.p, s.d_tdiag.p);
            IRTT_HIP(hipGetLastError());
        }
        dim3 fg(cdiv(ld, kFinCols), cdiv(nb, kFinRows));
        k_finalize<<<fg, 256, 0, stm>>>(s.d_acc.p, s.d_touch.p, s.d_diag.p, nb, ld, Q, s.d_T.p,
                                        s.d_colacc.p, s.d_tdiag.p);
        IRTT_HIP(hipGetLastError());
    }
    IRTT_HIP(hipEventRecord(s.ev2, stm));
    // energy conservation in FP64 on the device (after the column sums)
    s.d_energy.alloc(8);
    k_energy<<<1, kEnergyThreads, 0, stm>>>(reinterpret_cast<const unsigned long long *>(s.d_q.p), s.nt, per,
                                 s.d_colacc.p, s.d_diag.p, nb, s.d_stats.p, Q, s.d_energy.p);
    IRTT_HIP(hipGetLastError());
    // read-backs into the scene's page-locked block: column sums, counters,
    // stored diagonal, energy figures
    auto *colacc = static_cast<unsigned long long *>(
        s.h_pin.get((size_t)nb * sizeof(unsigned long long) + 16 * sizeof(unsigned long long) +
                    8 * sizeof(double) + 4 * sizeof(unsigned long long) +
                    (size_t)nb * sizeof(float)));
    unsigned long long *hstats = colacc + nb;
    double *henergy = reinterpret_cast<double *>(hstats + 16);
    unsigned long long *hctl = reinterpret_cast<unsigned long long *>(henergy + 8);
    float *dcol = reinterpret_cast<float *>(hctl + 4);
    if (pr.first_dropped)
        IRTT_HIP(hipMemcpyAsync(hctl, ps.ctl.p, 3 * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipMemcpyAsync(henergy, s.d_energy.p, 6 * sizeof(double), hipMemcpyDeviceToHost,
                            stm));
    IRTT_HIP(hipMemcpyAsync(colacc, s.d_colacc.p, nb * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipMemcpyAsync(hstats, s.d_stats.p, 16 * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipMemcpyAsync(dcol, s.d_tdiag.p, nb * sizeof(float), hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipStreamSynchronize(stm));
    IRTT_HIP(hipEventElapsedTime(&st.trace_ms, s.ev0, s.ev1));
    IRTT_HIP(hipEventElapsedTime(&st.finalize_ms, s.ev1, s.ev2));

    s.colsum.assign(nb, 0.0);
    for (uint32_t j = 0; j < nb; j++) s.colsum[j] = diag[j] - (double)colacc[j] * Q;
    // diagonal of the stored operator, for the Jacobi preconditioner
    s.Tdiag.assign(dcol, dcol + nb);
    st.photons = total;
    st.rays = hstats[S_RAYS];
    st.absorbed = hstats[S_ABS];
    st.miss = hstats[S_MISS];
    st.too_close = hstats[S_CLOSE];
    st.backface = hstats[S_BACK];
    s.fwd_rays = st.rays;
    s.fwd_segments = st.rays - st.miss - st.too_close - st.backface;
    st.degenerate = hstats[S_DEG];
    st.forced_absorb = hstats[S_FORCED];
    st.node_visits = hstats[S_COUNT];
    st.tri_tests = hstats[S_COUNT + 1];
    st.quantum = Q;
    st.emitted_w = henergy[0];
    st.absorbed_w = henergy[1];
    st.escaped_w = henergy[2];
    st.dropped_w = henergy[3];
    st.energy_rel = henergy[4];
    st.energy_exact = henergy[5] != 0.0 ? 1u : 0u;
    st.bvh_ms = s.bvh.build_ms;
    st.layout = p.layout;
    st.nnz = sparse ? s.sp.nnz : (uint64_t)nb * nb;
    s.layout = p.layout;
    s.op_kind = sparse ? IRTT_OP_SPARSE : IRTT_OP_DENSE;
    if (pr.first_dropped) {
        ps.sc = std::min<uint64_t>(hctl[0], ps.seg_chunks);
        ps.ec = std::min<uint64_t>(hctl[1], ps.edge_chunks);
        ps.first_dropped = hctl[2];
        ps.photons = total;
        ps.valid = true;
    }
    s.has_transport = true;
    s.acc_marked = !sparse;
    s.t_fill_gen = sparse ? ~0ull : s.assembly_gen;
}

}  // namespace irtt
