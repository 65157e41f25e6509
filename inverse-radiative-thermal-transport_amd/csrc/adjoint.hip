// adjoint.hip -- adjoint photon trace: d(objective)/d(vertex positions).
//
// Replaces ThermalTransport::computeAdjoint (thermal_transport.cpp:740-832)
// and the raygen program in TRACING_MODE 2
// (assets/shader/raytracing_thermal/thermal_tracing_unified.glsl:292-599).
// Every photon is replayed from its seed once its absorbing bin is known (the
// reference's pass 0; here the forward build's per-photon record, or a
// forward pass that only records absorbers), weighted by
// sum_h lambda_h[bin] * x_h^4[source bin] (sumOverHistory, :227-244, indexed
// history-major so every history length is handled, not only one step):
//   (i)   -/+ dw/dv of the three source vertices at emission / absorption,
//   (ii)  w * d ln|G| / dv for the six vertices of every path segment, with
//         G = -(N_i.d)(N_j.d) / (|N_i||N_j|(r+eps)^4), d = p_j - p_i,
//   (iii) the approximate edge term: a hit within edge_ratio of an edge casts
//         one occlusion ray from just outside the edge; a front-facing,
//         non-adjacent occluded surface receives +, the path's origin vertex +,
//         the edge vertex - of the edge-irradiance derivative.
//
// Stages instead of one kernel, so that no kernel carries another's state:
//   k_adj_paths     replays the paths (the forward kernel's lane scheduling and
//                   traversal) and writes one 48-byte segment record per
//                   accepted segment and one 64-byte edge record per hit near
//                   an edge (the hit as found), in wave-owned chunks -- or,
//                   when the build kept its paths (IRTT_TRANSPORT_KEEP_PATHS),
//                   the build's photon kernel has written them already;
//   k_edge_prep     turns each edge record into its occlusion ray and
//                   edge-irradiance derivative;
//   k_adj_segments  runs the segment terms (ii) and the hit-area extension
//                   over the records radix-sorted by start triangle and by
//                   hit triangle: a window of 64 records sums the terms of
//                   equal triangles in consecutive lanes (segmented scan) and
//                   adds each run once;
//   k_adj_occl      traces the edge records' occlusion rays and scatters (iii);
//   k_adj_area      sums (i) per source triangle from the seeds and absorbers.
// Gradients are summed exactly as 128-bit fixed-point integers (a 64-bit
// atomic add and, on carry or for large magnitudes, a second one on the high
// word), one copy per XCD, so the result does not depend on the order of the
// adds: the trace is bit-reproducible, alone or beside other work.  When the
// records of all photons do not fit, the photons are replayed in ranges.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "device_math.hpp"
#include "irtt_internal.hpp"
#include "path_records.hpp"
#include "traverse.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace irtt {
namespace {

__device__ __forceinline__ V3 add3(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 scl3(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ bool finite3(V3 a)
{
    return isfinite(a.x) && isfinite(a.y) && isfinite(a.z);
}
__device__ __forceinline__ V3 load_v(const float *xyz, uint32_t v)
{
    return V3{xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2]};
}

// d|N|/2 (area) w.r.t. the three vertices
__device__ __forceinline__ void area_grad(V3 A, V3 B, V3 C, V3 g[3])
{
    V3 e1 = B - A, e2 = C - A;
    V3 n = normalize3(cross3(e1, e2));
    V3 db = scl3(cross3(e2, n), 0.5f);
    V3 dc = scl3(cross3(n, e1), 0.5f);
    g[0] = scl3(add3(db, dc), -1.0f);
    g[1] = db;
    g[2] = dc;
}

// G and d ln|G| / d{a_i,b_i,c_i,a_j,b_j,c_j}: returns G and hands each
// vertex's gradient to emit(v, g) as soon as it is formed, so the six are never
// live at once
template <typename Emit>
__device__ __forceinline__ float geom_grad(V3 Ai, V3 Bi, V3 Ci, V3 Aj, V3 Bj, V3 Cj, float bi1,
                                           float bi2, float bj1, float bj2, float eps, Emit emit)
{
    float wi0 = 1.0f - bi1 - bi2, wj0 = 1.0f - bj1 - bj2;
    V3 pi = bary_point(Ai, Bi, Ci, wi0, bi1, bi2);
    V3 pj = bary_point(Aj, Bj, Cj, wj0, bj1, bj2);
    V3 d = pj - pi;
    float r = sqrt_rn(dot3(d, d));
    V3 dh = scl3(d, 1.0f / r);
    V3 e1i = Bi - Ai, e2i = Ci - Ai, e1j = Bj - Aj, e2j = Cj - Aj;
    V3 Ni = cross3(e1i, e2i), Nj = cross3(e1j, e2j);
    float li = sqrt_rn(dot3(Ni, Ni)), lj = sqrt_rn(dot3(Nj, Nj));
    V3 ni = scl3(Ni, 1.0f / li), nj = scl3(Nj, 1.0f / lj);
    float ci = dot3(Ni, d), cj = dot3(Nj, d);
    float re = r + eps;
    float re2 = re * re;
    float G = -(ci * cj) / (li * lj * (re2 * re2));
    if (!(fabsf(G) > eps * eps)) return G;
    const float ici = 1.0f / ci, icj = 1.0f / cj, ili = 1.0f / li, ilj = 1.0f / lj,
                ire = 4.0f / re;
    // one side at a time: s = 0 the segment's start triangle (i), 1 its end (j)
#pragma unroll
    for (int side = 0; side < 2; side++) {
        const V3 e1 = side ? e1j : e1i, e2 = side ? e2j : e2i, n = side ? nj : ni;
        const V3 e2xd = cross3(e2, d), dxe1 = cross3(d, e1);
        const V3 ab = cross3(e2, n), ac = cross3(n, e1);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int v = 3 * side + c;
            const float kd = side ? (c == 0 ? wj0 : (c == 1 ? bj1 : bj2))
                                  : -(c == 0 ? wi0 : (c == 1 ? bi1 : bi2));
            const V3 dN = c == 0 ? scl3(add3(e2xd, dxe1), -1.0f) : (c == 1 ? e2xd : dxe1);
            const V3 dL = c == 0 ? scl3(add3(ab, ac), -1.0f) : (c == 1 ? ab : ac);
            V3 g_ci = scl3(Ni, kd);
            V3 g_cj = scl3(Nj, kd);
            V3 g_li{0.0f, 0.0f, 0.0f}, g_lj{0.0f, 0.0f, 0.0f};
            if (side == 0) {
                g_ci = add3(g_ci, dN);
                g_li = dL;
            } else {
                g_cj = add3(g_cj, dN);
                g_lj = dL;
            }
            V3 g_r = scl3(dh, kd);
            emit(v, add3(add3(scl3(g_ci, ici), scl3(g_cj, icj)),
                         scl3(add3(add3(scl3(g_li, ili), scl3(g_lj, ilj)), scl3(g_r, ire)), -1.0f)));
        }
    }
    return G;
}

// The collimated beam's transfer from emitter triangle i (photons leave along
// its unit normal n_i) to receiver triangle j: B = -(N_j . n_i) / |N_j|, the
// cosine at the receiver, and d ln|B| / d{six vertices} through n_i on i's
// corners and through N_j, |N_j| on j's (a parallel beam has no distance or
// emitter-cosine factor).  Same emit() and gate as geom_grad.
template <class Emit>
__device__ __forceinline__ float beam_grad(V3 Ai, V3 Bi, V3 Ci, V3 Aj, V3 Bj, V3 Cj, float eps,
                                           Emit emit)
{
    V3 e1i = Bi - Ai, e2i = Ci - Ai, e1j = Bj - Aj, e2j = Cj - Aj;
    V3 Ni = cross3(e1i, e2i), Nj = cross3(e1j, e2j);
    float li = sqrt_rn(dot3(Ni, Ni)), lj = sqrt_rn(dot3(Nj, Nj));
    V3 ni = scl3(Ni, 1.0f / li), nj = scl3(Nj, 1.0f / lj);
    float cij = dot3(Ni, Nj), cn = dot3(Nj, ni);
    float B = -cn / lj;
    if (!(fabsf(B) > eps * eps)) return B;
#pragma unroll
    for (int side = 0; side < 2; side++) {
        // side 0: d(N_i . N_j) with N_j held; side 1: d(N_j . n_i) with n_i held
        const V3 e1 = side ? e1j : e1i, e2 = side ? e2j : e2i, n = side ? nj : ni;
        const V3 av = side ? ni : Nj;
        const float ic = 1.0f / (side ? cn : cij), il = 1.0f / (side ? lj : li);
        const V3 e2xa = cross3(e2, av), axe1 = cross3(av, e1);
        const V3 ab = cross3(e2, n), ac = cross3(n, e1);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const V3 dN = c == 0 ? scl3(add3(e2xa, axe1), -1.0f) : (c == 1 ? e2xa : axe1);
            const V3 dL = c == 0 ? scl3(add3(ab, ac), -1.0f) : (c == 1 ? ab : ac);
            emit(3 * side + c, scl3(dN, ic) - scl3(dL, il));
        }
    }
    return B;
}

// ---- exact gradient sums ---------------------------------------------------
// One gradient component is a signed 128-bit integer in units of q (a power
// of two).  v is rounded to the nearest multiple of q once; the sums of the
// rounded values are exact, so any order of the adds gives the same bits.
// v in units of q, rounded once, as the 128-bit pair (lo, hi), value
// hi * 2^64 + lo
__device__ __forceinline__ void fx_quant(double v, double inv_q, unsigned long long &lo,
                                         long long &hi)
{
    const double y = rint(v * inv_q);
    if (fabs(y) < 9.223372036854775808e18) {
        const long long s = (long long)y;
        lo = (unsigned long long)s;
        hi = s < 0 ? -1 : 0;
    } else if (isfinite(y)) {
        // |y| >= 2^63: y = hd 2^64 + ld with 0 <= ld < 2^64, both exact
        // (y is a multiple of 2^11 here)
        const double hd = floor(y * 5.421010862427522e-20);
        const double ld = y - hd * 1.8446744073709552e19;
        lo = (unsigned long long)ld;
        hi = (long long)hd;
    } else {
        lo = 0;
        hi = 0;
    }
}

// A gradient component is three 64-bit words that take the 128-bit addend
// y = lo + 2^64 hi in parts, each with an atomic add whose result is not
// waited for.  When y fits in 64 signed bits (all but rare outliers): word 0
// takes its low 32 bits (0 .. 2^32 - 1) and word 1 the signed rest y >> 32;
// otherwise word 0 takes the low 32 bits of lo, word 1 its high 32 bits and
// word 2 hi.  Every part added to words 0 and 1 is below 2^32 in magnitude,
// so they hold 2^31 adds without overflow; the host forms word0 +
// 2^32 word1 + 2^64 word2 (words 1 and 2 signed) modulo 2^128, the exact
// sum in any order.
constexpr uint32_t kCompWords = 3;
__device__ __forceinline__ unsigned long long *fx_comp(unsigned long long *g, uint32_t vertex,
                                                       uint32_t axis)
{
    return g + kCompWords * (3 * (size_t)vertex + axis);
}

// adds the 128-bit (lo, hi) to the component at c
__device__ __forceinline__ void fx_add_q(unsigned long long *c, unsigned long long lo, long long hi)
{
    const unsigned long long l0 = lo & 0xffffffffull;
    const long long sl = (long long)lo;
    if (hi == (sl >> 63)) {
        const long long a = sl >> 32;
        if (l0) (void)atomicAdd(c, l0);
        if (a) (void)atomicAdd(c + 1, (unsigned long long)a);
    } else {
        if (l0) (void)atomicAdd(c, l0);
        if (lo >> 32) (void)atomicAdd(c + 1, lo >> 32);
        (void)atomicAdd(c + 2, (unsigned long long)hi);
    }
}

__device__ __forceinline__ void fx_add(unsigned long long *c, double v, double inv_q)
{
    unsigned long long lo;
    long long hi;
    fx_quant(v, inv_q, lo, hi);
    fx_add_q(c, lo, hi);
}

// The eight XCDs each sum into their own copy of the gradient (3 n_vertices
// components apiece), so that atomics to one component from different XCDs
// do not queue on one address; the host adds the copies.  The copy is the
// XCD's id (hardware register XCC_ID, bits 3:0).
constexpr uint32_t kGradCopies = 8;
__device__ __forceinline__ uint32_t xcc_id()
{
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | 20) & (kGradCopies - 1);
}

__device__ __forceinline__ void fx_add3(unsigned long long *g, uint32_t v, V3 d, double f,
                                        double inv_q)
{
    fx_add(fx_comp(g, v, 0), f * (double)d.x, inv_q);
    fx_add(fx_comp(g, v, 1), f * (double)d.y, inv_q);
    fx_add(fx_comp(g, v, 2), f * (double)d.z, inv_q);
}

// ---- shared arguments ---------------------------------------------------------
struct AdjCommon {
    const float *xyz;
    const uint32_t *idx;
    const float *rho_d, *rho_s;
    const uint8_t *flags;
    const float *w, *area;
    const float4 *nodes, *tris;
    uint32_t n_leaves;
    uint32_t nt, nb, bin_mode;
    uint32_t per_tri, rays_per_thread, batch_seed, max_depth;
    float tiny_dist, eps, ratio;
    uint32_t aflags, n_hist;
    const double *lambda;      // n_hist x nb
    const double *x4;          // n_hist x nb
    const uint32_t *absorber;  // per photon (~0: none)
    uint64_t photons;
    unsigned long long *grad;  // kGradCopies x 3 n_vertices components, kCompWords words each
    uint64_t grad_stride;      // words per copy (9 n_vertices)
    double inv_q;
    unsigned long long *stats;
};

enum { A_ABS = 0, A_SEG, A_CONT, A_EDGE, A_OCCL, A_COUNT };

// photon kernel grid of the path replay: waves per CU, lanes striding over
// the photons (the forward kernel's setting)
#ifndef IRTT_PHOTON_WAVES_PER_CU_ADJ
#define IRTT_PHOTON_WAVES_PER_CU_ADJ 128
#endif

__device__ __forceinline__ double hist_factor(const AdjCommon &a, uint32_t bin_l, uint32_t bin_x)
{
    double f = 0.0;
    for (uint32_t h = 0; h < a.n_hist; h++)
        f += a.lambda[(size_t)h * a.nb + bin_l] * a.x4[(size_t)h * a.nb + bin_x];
    return f;
}

struct PathArgs {
    AdjCommon c;
    uint64_t g_begin, g_end;       // photons replayed
    const uint32_t *live;          // or, if set: the photons listed in live[0, *n_live)
    const uint32_t *n_live;
    RecBuf segs, edges;
    unsigned long long *first_dropped;  // smallest photon with a record that did not fit
};

// waves per SIMD the LDS stack admits (160 KiB per CU, four SIMDs), at most
// IRTT_ADJ_WAVES (the replay kernel needs 94 VGPRs: 5 waves without spills)
#ifndef IRTT_ADJ_WAVES
#define IRTT_ADJ_WAVES 5
#endif
constexpr int adjoint_waves(int stack)
{
    return (160 * 1024 / (stack * 4 * kTraceBlock)) / 4 < IRTT_ADJ_WAVES
               ? (160 * 1024 / (stack * 4 * kTraceBlock)) / 4
               : IRTT_ADJ_WAVES;
}
// the occlusion-ray kernel holds far fewer registers: up to 6 waves per SIMD
constexpr int occl_waves(int stack)
{
    return (160 * 1024 / (stack * 4 * kTraceBlock)) / 4 < 6
               ? (160 * 1024 / (stack * 4 * kTraceBlock)) / 4
               : 6;
}

// Replays the photon paths (one segment per lane per iteration, lanes of a
// wave on consecutive photons, a lane whose photon ends takes its next
// photon) and writes one segment record per accepted segment and one edge
// record per hit near an edge.  Photons without an absorbing bin are not
// replayed (unless IRTT_ADJ_REFERENCE_DROPPED replays them with bin 0, as the
// shader does).  A record that does not fit is dropped and its photon noted.
template <int STACK>
__global__ __launch_bounds__(kTraceBlock, adjoint_waves(STACK)) void k_adj_paths(PathArgs pa)
{
    const AdjCommon &a = pa.c;
    __shared__ uint32_t stk_n[STACK * kTraceBlock];
    __shared__ float stk_t[stack_t_entries(STACK) * kTraceBlock];
    uint32_t *sn = stk_n + threadIdx.x;
    float *stt = stk_t + threadIdx.x;
    const bool f_edge = a.aflags & IRTT_ADJ_EDGE, f_proj = a.aflags & IRTT_ADJ_EDGE_PROJECTED;
    const bool dropped_too = a.aflags & IRTT_ADJ_REFERENCE_DROPPED;
    uint32_t nvis = 0, ntri = 0;
    uint32_t s_chunk = kNoChunk, s_fill = 0, e_chunk = kNoChunk, e_fill = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kTraceBlock;
    // g: this lane's next photon (or list position) to look at; gp: the
    // photon being replayed
    const uint64_t g_stop = pa.live ? (uint64_t)*pa.n_live : pa.g_end;
    uint64_t g = (pa.live ? 0 : pa.g_begin) + (uint64_t)blockIdx.x * kTraceBlock + threadIdx.x,
             gp = 0;
    uint32_t t = 0, src = 0, abs_bin = 0, seed = 0, depth = 0, prev_tri = 0;
    int prev_max = 0;
    float pb1 = 0.0f, pb2 = 0.0f;
    V3 o{0.0f, 0.0f, 0.0f}, d{0.0f, 0.0f, 1.0f};
    bool live = false;
    for (;;) {
        if (!live) {
            // this lane's next photon with a path to replay
            for (; g < g_stop; g += stride) {
                const uint64_t ph = pa.live ? (uint64_t)pa.live[g] : g;
                const uint32_t r = a.absorber[ph];
                if (r == 0xffffffffu && !dropped_too) continue;
                t = (uint32_t)(ph / a.per_tri);
                // the emission ray (getEmissionRay, thermal_tracing_unified.glsl:125-161)
                const uint32_t i0 = a.idx[3 * t], i1 = a.idx[3 * t + 1], i2 = a.idx[3 * t + 2];
                const V3 p0 = load_v(a.xyz, i0), p1 = load_v(a.xyz, i1), p2 = load_v(a.xyz, i2);
                {
                    // thermal_tracing_unified.glsl:291
                    const uint32_t k = (uint32_t)(ph - (uint64_t)t * a.per_tri);
                    const uint32_t xx = k / a.rays_per_thread, nn = k - xx * a.rays_per_thread;
                    seed = tea16(a.nt * xx + t, a.batch_seed + nn);
                }
                const float u0 = rnd(seed), u1 = rnd(seed);
                (void)rnd(seed);
                const float su = sqrt_rn(u0);
                const float b0 = 1.0f - su, b1 = su * (1.0f - u1), b2 = su * u1;
                const V3 pos = bary_point(p0, p1, p2, b0, b1, b2);
                float len;
                const V3 nrm = normalize3(cross3(p1 - p0, p2 - p0), len);
                o = offset_ray(pos, nrm);
                d = nrm;
                if (a.flags[t] & IRTT_MAT_DIFFUSE_EMISSION) {
                    const float r0 = rnd(seed), r1 = rnd(seed);
                    (void)rnd(seed);
                    d = tangent_to_world(cosine_hemisphere(r0, r1), nrm);
                }
                if (!(len > 0.0f)) continue;  // degenerate emitter: no terms
                const int sc = max_corner(b0, b1, b2);
                src = a.bin_mode ? t : (sc == 0 ? i0 : (sc == 1 ? i1 : i2));
                abs_bin = r == 0xffffffffu ? 0u : r;
                prev_tri = t;
                pb1 = b1;
                pb2 = b2;
                prev_max = sc;
                depth = 0;
                gp = ph;
                g += stride;
                live = true;
                break;
            }
        }
        // every lane stays in the loop until the wave is done: the record
        // slots are allocated by the whole wave together
        if (__ballot(live) == 0) break;
        Query q{};
        q.cur = kQueryDone;
        if (live) {
            ++depth;
            if (!query_begin(q, a.tris, a.n_leaves, o, d, 1e-12f, 1e12f, ntri)) q.cur = kQueryDone;
        }
        query_run_ref(q, a.nodes, a.tris, 1e-12f, sn, nvis, ntri);
        // the reference's reject rules: miss, too close, (near) backface
        bool seg = false;
        Hit h{-1, 0.0f, 0.0f, 0.0f};
        uint32_t hp = 0, j0 = 0, j1 = 0, j2 = 0;
        V3 q0{0.0f, 0.0f, 0.0f}, q1 = q0, q2 = q0, hn = q0;
        if (live) {
            h = query_end(q, 1e12f);
            if (h.prim >= 0 && h.t >= a.tiny_dist) {
                hp = (uint32_t)h.prim;
                j0 = a.idx[3 * hp];
                j1 = a.idx[3 * hp + 1];
                j2 = a.idx[3 * hp + 2];
                q0 = load_v(a.xyz, j0);
                q1 = load_v(a.xyz, j1);
                q2 = load_v(a.xyz, j2);
                hn = normalize3(cross3(q1 - q0, q2 - q0));
                seg = dot3(hn, d) <= -1e-3f;
            }
            if (!seg) live = false;
        }
        {
            bool ok;
            const uint64_t slot = chunk_slot(seg, s_chunk, s_fill, pa.segs, ok);
            if (seg) {
                if (ok) {
                    SegRec *dst = static_cast<SegRec *>(pa.segs.p) + slot;
                    dst->r0 = make_uint4(prev_tri, hp, src, abs_bin);
                    dst->r1 = make_float4(pb1, pb2, h.b1, h.b2);
                    dst->r2 = make_uint4((uint32_t)gp, t, depth == 1 ? 1u : 0u, 0u);
                } else {
                    atomicMin(pa.first_dropped, (unsigned long long)gp);
                }
            }
        }
        // the approximate edge term, :454-567: a hit near an edge becomes an
        // edge record, its occlusion ray derived later (k_edge_prep) by all
        // lanes of a wave at once instead of by the few lanes here that hit
        // near an edge
        const float hb0 = 1.0f - h.b1 - h.b2;
        bool edge = false;
        if (seg && f_edge) {
            const int opp = min_corner(hb0, h.b1, h.b2);
            edge = (opp == 0 ? hb0 : (opp == 1 ? h.b1 : h.b2)) <= a.ratio;
        }
        if (f_edge) {
            bool ok;
            const uint64_t slot = chunk_slot(edge, e_chunk, e_fill, pa.edges, ok);
            if (edge) {
                if (ok) {
                    EdgeRec *dst = static_cast<EdgeRec *>(pa.edges.p) + slot;
                    dst->r0 = make_float4(d.x, d.y, d.z, h.t);
                    dst->r1 = make_float4(h.b1, h.b2, __uint_as_float(hp), __uint_as_float(t));
                    dst->r2 = make_float4(__uint_as_float(prev_tri), __uint_as_float((uint32_t)prev_max),
                                          __uint_as_float(src), __uint_as_float(abs_bin));
                    dst->r3 = make_uint4(0u, 0u, 0u, (uint32_t)gp);
                } else {
                    atomicMin(pa.first_dropped, (unsigned long long)gp);
                }
            }
        }
        if (seg) {
            prev_tri = hp;
            pb1 = h.b1;
            pb2 = h.b2;
            prev_max = max_corner(hb0, h.b1, h.b2);
            // getNextRay, :163-197
            const V3 hpos = bary_point(q0, q1, q2, hb0, h.b1, h.b2);
            o = offset_ray(hpos, hn);
            const float u = rnd(seed);
            const float rd = a.rho_d[hp], rs = a.rho_s[hp];
            bool absorbed = false;
            if (u >= rd + rs) {
                absorbed = true;
            } else if (u >= rd) {
                const V3 in = normalize3(d);
                const float dn = 2.0f * dot3(hn, in);
                d = V3{in.x - dn * hn.x, in.y - dn * hn.y, in.z - dn * hn.z};
            } else {
                const float r0 = rnd(seed), r1 = rnd(seed);
                (void)rnd(seed);
                d = tangent_to_world(cosine_hemisphere(r0, r1), hn);
            }
            if (absorbed || depth == a.max_depth) live = false;
        }
    }
    chunk_close(s_chunk, s_fill, pa.segs);
    chunk_close(e_chunk, e_fill, pa.edges);
}

// The photons with an absorbing bin, in ascending order, so that the
// replay's lanes take only photons with a path instead of skipping the others
// one by one: per 256-photon block its count, then (after an exclusive scan
// of the counts) its photons at its offset.
__global__ __launch_bounds__(256) void k_live_count(const uint32_t *__restrict__ absorber,
                                                    uint64_t photons, uint32_t *__restrict__ counts)
{
    const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool has = g < photons && absorber[g] != 0xffffffffu;
    const int n = __syncthreads_count(has);
    if (threadIdx.x == 0) counts[blockIdx.x] = (uint32_t)n;
}

__global__ __launch_bounds__(256) void k_live_fill(const uint32_t *__restrict__ absorber,
                                                   uint64_t photons,
                                                   const uint32_t *__restrict__ offsets,
                                                   uint32_t n_blocks, const uint32_t *__restrict__ counts,
                                                   uint32_t *__restrict__ live, uint32_t *__restrict__ n_live)
{
    __shared__ uint32_t wsum[4];
    const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool has = g < photons && absorber[g] != 0xffffffffu;
    const unsigned long long m = __ballot(has);
    const uint32_t wave = threadIdx.x / 64;
    if (__lane_id() == 0) wsum[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = offsets[blockIdx.x];
    for (uint32_t w = 0; w < wave; w++) base += wsum[w];
    if (has) live[base + (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull))] = (uint32_t)g;
    if (blockIdx.x == n_blocks - 1 && threadIdx.x == 0) *n_live = offsets[blockIdx.x] + counts[blockIdx.x];
}

// Edge records, raw form (in) -> ray form (out, the same slot; in == out
// converts in place; empty slots of partly filled chunks left alone): the
// occlusion ray past the edge and the edge's irradiance derivative
// (thermal_tracing_unified.glsl:454-494), from the recorded hit.  A record
// of a photon without an absorbing bin gets photon index ~0 (skipped by
// k_adj_occl) unless IRTT_ADJ_REFERENCE_DROPPED counts it with bin 0.
__global__ __launch_bounds__(256) void k_edge_prep(AdjCommon a, const EdgeRec *in, EdgeRec *out,
                                                   const uint32_t *__restrict__ fills,
                                                   uint32_t n_chunks, uint32_t absorber_from_photon)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n_chunks * kChunk) return;
    if ((uint32_t)(i % kChunk) >= fills[i / kChunk]) return;
    const bool f_proj = a.aflags & IRTT_ADJ_EDGE_PROJECTED;
    const bool dropped_too = a.aflags & IRTT_ADJ_REFERENCE_DROPPED;
    const EdgeRec raw = in[i];
    const V3 d{raw.r0.x, raw.r0.y, raw.r0.z};
    const float ht = raw.r0.w, b1 = raw.r1.x, b2 = raw.r1.y;
    const uint32_t hp = __float_as_uint(raw.r1.z), t = __float_as_uint(raw.r1.w);
    const uint32_t prev_tri = __float_as_uint(raw.r2.x), prev_max = __float_as_uint(raw.r2.y);
    const uint32_t j0 = a.idx[3 * hp], j1 = a.idx[3 * hp + 1], j2 = a.idx[3 * hp + 2];
    const V3 q0 = load_v(a.xyz, j0), q1 = load_v(a.xyz, j1), q2 = load_v(a.xyz, j2);
    const V3 hn = normalize3(cross3(q1 - q0, q2 - q0));
    const float hb0 = 1.0f - b1 - b2;
    const int opp = min_corner(hb0, b1, b2);
    const int nr = max_corner(hb0, b1, b2);
    int other = 0;
    if (opp == 0 || nr == 0) {
        other = 1;
        if (opp == 1 || nr == 1) other = 2;
    }
    const V3 qn = nr == 0 ? q0 : (nr == 1 ? q1 : q2);
    const V3 qo = other == 0 ? q0 : (other == 1 ? q1 : q2);
    const V3 qp = opp == 0 ? q0 : (opp == 1 ? q1 : q2);
    const uint32_t jn = nr == 0 ? j0 : (nr == 1 ? j1 : j2);
    V3 edir = qn - qo;
    const V3 oppe = qn - qp;
    const float elen = sqrt_rn(dot3(edir, edir));
    edir = scl3(edir, 1.0f / elen);
    V3 n_edg = normalize3(cross3(edir, hn));
    float th = dot3(oppe, n_edg);
    if (th < 0.0f) {
        th = -th;
        n_edg = scl3(n_edg, -1.0f);
    }
    const float larea = a.ratio * th * (1.0f - 0.5f * a.ratio) * elen;
    V3 da = normalize3(cross3(edir, d));
    if (dot3(da, n_edg) < 0.0f) da = scl3(da, -1.0f);
    // the reference scales by |da.n_edg| (:486); the extension divides by it:
    // an edge moved by u sweeps (u.da)/|da.n_edg| of the occluder's plane per
    // unit length, seen along the ray
    const float cpl = fabsf(dot3(da, n_edg));
    V3 dE = f_proj ? scl3(da, elen * a.w[t] / (larea * cpl)) : scl3(da, cpl * elen * a.w[t] / larea);
    if (!finite3(dE)) dE = V3{0.0f, 0.0f, 0.0f};
    V3 to = bary_point(q0, q1, q2, hb0, b1, b2);
    to = add3(to, scl3(n_edg, a.eps + dot3(qn - to, n_edg)));
    to = to - scl3(d, a.eps);
    EdgeRec r;
    r.r0 = make_float4(to.x, to.y, to.z, ht);
    r.r1 = make_float4(d.x, d.y, d.z, __uint_as_float(hp));
    r.r2 = make_float4(dE.x, dE.y, dE.z, __uint_as_float(jn));
    // kept paths: the absorbing bin from the per-photon record
    uint32_t gp = raw.r3.w, ab = absorber_from_photon ? a.absorber[gp] : __float_as_uint(raw.r2.w);
    if (ab == 0xffffffffu) {
        ab = 0u;
        if (!dropped_too) gp = 0xffffffffu;
    }
    r.r3 = make_uint4(a.idx[3 * prev_tri + prev_max], __float_as_uint(raw.r2.z), ab, gp);
    out[i] = r;
}

// sort keys of the segment slots: the start (side 0) or hit (side 1)
// triangle, nt for empty slots and, with skip_unabsorbed, for records of
// photons without an absorbing bin (kept paths hold every photon's).  With
// absorber (kept paths, side 0), each record's absorbing bin is first taken
// from the per-photon record into the segment record.
__global__ void k_seg_keys(SegRec *__restrict__ segs, const uint32_t *__restrict__ fills,
                           uint32_t n_chunks, uint32_t nt, uint32_t side, uint32_t skip_unabsorbed,
                           const uint32_t *__restrict__ absorber, uint32_t *__restrict__ keys,
                           uint32_t *__restrict__ vals)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n_chunks * kChunk) return;
    const uint32_t c = (uint32_t)(i / kChunk), k = (uint32_t)(i % kChunk);
    uint32_t key = nt;
    if (k < fills[c]) {
        uint4 r0 = segs[i].r0;
        if (absorber) {
            r0.w = absorber[segs[i].r2.x];
            segs[i].r0.w = r0.w;
        }
        if (!(skip_unabsorbed && r0.w == 0xffffffffu)) key = side ? r0.y : r0.x;
    }
    keys[i] = key;
    vals[i] = (uint32_t)i;
}

struct SegArgs {
    AdjCommon c;
    const SegRec *segs;
    const uint32_t *order;      // slots sorted by the side's triangle
    const uint32_t *keys;       // their triangles (nt: empty slot)
    uint64_t n;                 // slots
    uint32_t side;              // 0: start triangle's vertices, 1: hit triangle's
    uint64_t g_limit;           // records of photons at or above this are skipped
    uint32_t *big;              // slots of records with a component too large for the
    uint32_t *n_big;            // 64-bit window sums (n entries), and their count
};

// v in units of q, rounded once, when |v / q| < 2^55: the window sums of 64
// such values (two terms each) stay below 2^62 and are exact in 64 bits;
// false for larger or non-finite values, which are added on their own
__device__ __forceinline__ bool fx_small(double v, double inv_q, long long &out)
{
    const double y = rint(v * inv_q);
    if (!(fabs(y) < 36028797018963968.0)) return false;
    out = (long long)y;
    return true;
}

// 128-bit (lo, hi) sum
__device__ __forceinline__ void add128(unsigned long long &lo, long long &hi, unsigned long long lo2,
                                       long long hi2)
{
    const unsigned long long s = lo + lo2;
    hi = hi + hi2 + (s < lo ? 1 : 0);
    lo = s;
}

// The terms of one segment record on one side (side 0: its start triangle's
// vertices, side 1: its hit triangle's vertices and the hit-area extension):
// put(e, v) receives component e = 3 corner + axis of the side's triangle.
template <int SIDE, class Put>
__device__ __forceinline__ void seg_terms(const AdjCommon &a, const SegRec &r, bool f_cont,
                                          bool f_hit, Put put, uint32_t &n_cont)
{
    const uint32_t pt = r.r0.x, hp = r.r0.y, t = r.r2.y;
    // a photon without an absorbing bin counts with bin 0 (its records are
    // taken only under IRTT_ADJ_REFERENCE_DROPPED)
    const double fpath = hist_factor(a, r.r0.w == 0xffffffffu ? 0u : r.r0.w, r.r0.z);
    const uint32_t i0 = a.idx[3 * pt], i1 = a.idx[3 * pt + 1], i2 = a.idx[3 * pt + 2];
    const uint32_t j0 = a.idx[3 * hp], j1 = a.idx[3 * hp + 1], j2 = a.idx[3 * hp + 2];
    const V3 q0 = load_v(a.xyz, j0), q1 = load_v(a.xyz, j1), q2 = load_v(a.xyz, j2);
    const float wt = a.w[t];
    if (f_cont) {
        constexpr int lo_v = SIDE == 0 ? 0 : 3;
        auto emit = [&](int v, V3 gv) {
            if (v < lo_v || v >= lo_v + 3) return;
            const V3 dv = scl3(gv, wt);
            if (!finite3(dv)) return;
            const int c = v - lo_v;
            put(3 * c, fpath * (double)dv.x);
            put(3 * c + 1, fpath * (double)dv.y);
            put(3 * c + 2, fpath * (double)dv.z);
        };
        // extension: a collimated emitter's first segment carries the beam's
        // transfer
        const bool beam = (a.aflags & IRTT_ADJ_COLLIMATED) && r.r2.z != 0 &&
                          !(a.flags[t] & IRTT_MAT_DIFFUSE_EMISSION);
        const V3 p0 = load_v(a.xyz, i0), p1 = load_v(a.xyz, i1), p2 = load_v(a.xyz, i2);
        const float G = beam ? beam_grad(p0, p1, p2, q0, q1, q2, a.eps, emit)
                             : geom_grad(p0, p1, p2, q0, q1, q2, r.r1.x, r.r1.y, r.r1.z, r.r1.w,
                                         a.eps, emit);
        if (SIDE == 0 && fabsf(G) > a.eps * a.eps) n_cont++;
    }
    if (SIDE == 1 && f_hit) {
        // extension: the hit triangle's area Jacobian (the hit point is a
        // fixed barycentric point of it)
        V3 ga[3];
        area_grad(q0, q1, q2, ga);
        const float sh = wt / a.area[hp];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const V3 dh = scl3(ga[c], sh);
            if (!finite3(dh)) continue;
            put(3 * c, fpath * (double)dh.x);
            put(3 * c + 1, fpath * (double)dh.y);
            put(3 * c + 2, fpath * (double)dh.z);
        }
    }
}

// The segment terms with a sort-reduce scatter: the continuous geometric
// term (thermal_tracing_unified.glsl:391-448) on one side of each segment --
// side 0 its start triangle's vertices, side 1 its hit triangle's vertices
// and the hit-area extension -- over the records sorted by that triangle.  A
// window of 64 records sums the terms of equal triangles in consecutive
// lanes exactly (64-bit integers in units of q, segmented scan) and adds each
// run once: about one add per triangle and window instead of one per record.
// A record with a component of 2^55 q or more (rare: grazing geometry) is
// left out of the window and listed for k_adj_seg_big.
template <int SIDE>
__global__ __launch_bounds__(256) void k_adj_segments(SegArgs sa)
{
    const AdjCommon &a = sa.c;
    unsigned long long *const gx = a.grad + xcc_id() * a.grad_stride;
    const bool f_cont = a.aflags & IRTT_ADJ_CONTINUOUS;
    const bool f_hit = SIDE == 1 && (a.aflags & IRTT_ADJ_HIT_AREA);
    const uint32_t lane = __lane_id();
    uint32_t n_seg = 0, n_cont = 0;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint32_t n_waves = gridDim.x * blockDim.x / 64;
    const uint64_t n_windows = sa.n / 64;
    for (uint64_t wi = wave; wi < n_windows; wi += n_waves) {
        const uint64_t pos = wi * 64 + lane;
        bool have = sa.keys[pos] != a.nt;
        const uint64_t slot = sa.order[pos];
        if (__ballot(have) == 0) continue;
        SegRec r{};
        if (have) {
            r = sa.segs[slot];
            have = (uint64_t)r.r2.x < sa.g_limit;
        }
        const uint32_t key = have ? (SIDE == 0 ? r.r0.x : r.r0.y) : 0xffffffffu;
        long long acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (have) {
            if (SIDE == 0) n_seg++;
            bool big = false;
            seg_terms<SIDE>(a, r, f_cont, f_hit,
                      [&](int e, double v) {
                          long long y;
                          if (fx_small(v, a.inv_q, y)) acc[e] += y;
                          else big = true;
                      },
                      n_cont);
            if (big) {
#pragma unroll
                for (int e = 0; e < 9; e++) acc[e] = 0;
                sa.big[atomicAdd(sa.n_big, 1u)] = (uint32_t)slot;
            }
        }
        // segmented inclusive scan over the window: runs of equal keys (every
        // lane takes part in each shuffle: a lane outside the active set
        // reads as zero); the 64-bit sums are exact
        const uint32_t key_prev = __shfl_up(key, 1, 64);
        bool head = lane == 0 || key_prev != key;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const bool head_up = __shfl_up(head ? 1 : 0, off, 64) != 0;
            const bool take = lane >= (uint32_t)off && !head;
#pragma unroll
            for (int e = 0; e < 9; e++) {
                const long long y2 = __shfl_up(acc[e], off, 64);
                if (take) acc[e] += y2;
            }
            if (lane >= (uint32_t)off) head = head || head_up;
        }
        // the run's last lane adds the run's sums
        const uint32_t key_next = __shfl_down(key, 1, 64);
        const bool last = lane == 63 || key_next != key;
        if (last && key != 0xffffffffu) {
            const uint32_t tri = key;
#pragma unroll
            for (int e = 0; e < 9; e++)
                fx_add_q(fx_comp(gx, a.idx[3 * tri + e / 3], e % 3), (unsigned long long)acc[e],
                         acc[e] < 0 ? -1ll : 0ll);
        }
    }
    if (SIDE == 0) {
        uint32_t cnt[2] = {n_seg, n_cont};
        const int slot_of[2] = {A_SEG, A_CONT};
#pragma unroll
        for (int k = 0; k < 2; k++) {
            uint32_t v = cnt[k];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (lane == 0 && v) atomicAdd(a.stats + slot_of[k], (unsigned long long)v);
        }
    }
}

// The listed records of k_adj_segments, each component added on its own
// (128-bit exact sums).
template <int SIDE>
__global__ __launch_bounds__(256) void k_adj_seg_big(SegArgs sa)
{
    const AdjCommon &a = sa.c;
    unsigned long long *const gx = a.grad + xcc_id() * a.grad_stride;
    const bool f_cont = a.aflags & IRTT_ADJ_CONTINUOUS;
    const bool f_hit = SIDE == 1 && (a.aflags & IRTT_ADJ_HIT_AREA);
    const uint32_t n = *sa.n_big;
    uint32_t unused = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const SegRec r = sa.segs[sa.big[i]];
        const uint32_t key = SIDE == 0 ? r.r0.x : r.r0.y;
        seg_terms<SIDE>(a, r, f_cont, f_hit,
                  [&](int e, double v) {
                      fx_add(fx_comp(gx, a.idx[3 * key + e / 3], e % 3), v, a.inv_q);
                  },
                  unused);
    }
}

struct OcclArgs {
    AdjCommon c;
    const EdgeRec *edges;
    const uint32_t *fills;
    uint32_t n_chunks;     // chunks written
    uint64_t g_limit;      // records of photons at or above this are skipped
};

// The terms of one answered occlusion ray (the edge term of
// thermal_tracing_unified.glsl:495-567): the edge record er and the closest
// hit h beyond the edge.
__device__ __forceinline__ void occl_terms(const AdjCommon &a, unsigned long long *gx,
                                           const EdgeRec &er, const Hit &h, bool f_gain,
                                           uint32_t &n_occl)
{
    const V3 od{er.r1.x, er.r1.y, er.r1.z};
    const V3 dE{er.r2.x, er.r2.y, er.r2.z};
    const float oht = er.r0.w;
    const uint32_t ohp = __float_as_uint(er.r1.w), edge_v = __float_as_uint(er.r2.w);
    const uint32_t origin_v = er.r3.x, src = er.r3.y, abs_bin = er.r3.z;
    const bool valid = h.prim >= 0 && h.t >= a.tiny_dist;
    bool nbr = false;
    uint32_t shadow_v = 0;
    if (valid) {
        const uint32_t op = (uint32_t)h.prim;
        const uint32_t c0 = a.idx[3 * op], c1 = a.idx[3 * op + 1], c2 = a.idx[3 * op + 2];
        const uint32_t oj0 = a.idx[3 * ohp], oj1 = a.idx[3 * ohp + 1], oj2 = a.idx[3 * ohp + 2];
        nbr = c0 == oj0 || c0 == oj1 || c0 == oj2 || c1 == oj0 || c1 == oj1 || c1 == oj2 ||
              c2 == oj0 || c2 == oj1 || c2 == oj2;
        const V3 s0 = load_v(a.xyz, c0), s1 = load_v(a.xyz, c1), s2 = load_v(a.xyz, c2);
        const V3 tn = normalize3(cross3(s1 - s0, s2 - s0));
        const float ob0 = 1.0f - h.b1 - h.b2;
        const int sm = max_corner(ob0, h.b1, h.b2);
        shadow_v = sm == 0 ? c0 : (sm == 1 ? c1 : c2);
        if (!nbr && dot3(tn, od) < -1e-3f) {
            n_occl++;
            const uint32_t shadow_bin = a.bin_mode ? op : shadow_v;
            const float sa = 1.0f - a.rho_d[op] - a.rho_s[op];
            const float d1 = oht, d2 = h.t, tot = d1 + d2;
            const double fs = hist_factor(a, shadow_bin, src);
            fx_add3(gx, shadow_v, scl3(dE, sa * (d1 / tot)), fs, a.inv_q);
            fx_add3(gx, origin_v, scl3(dE, sa * (d2 / tot)), fs, a.inv_q);
            fx_add3(gx, edge_v, scl3(dE, -sa), fs, a.inv_q);
        }
    }
    // Extension (not in the reference): the power the moving edge takes
    // from the shadow side lands on the hit triangle and follows this
    // photon's path, so it is credited to the path's absorbing bin with
    // the opposite sign, also when nothing lies behind the edge.
    if (f_gain && (h.prim < 0 || (valid && !nbr))) {
        const double fpath = hist_factor(a, abs_bin, src);
        const float f_or = valid ? h.t / (oht + h.t) : 1.0f;
        fx_add3(gx, edge_v, dE, fpath, a.inv_q);
        fx_add3(gx, origin_v, scl3(dE, f_or), -fpath, a.inv_q);
        if (valid) fx_add3(gx, shadow_v, scl3(dE, 1.0f - f_or), -fpath, a.inv_q);
    }
}

// Answered lanes that make a wave of k_adj_occl take new records.  Measured
// on the city block (10.8 M occlusion rays beside the segment sort keys):
// 64 (each window of 64 rays walked to its last answer) 3.55 ms, 48 3.4,
// 32 3.39, 16 3.4, 8 3.5 (adjoint trace 22.13 / 21.96 / 21.92 / 22.07 /
// 22.39 ms).
#ifndef IRTT_OCCL_REFILL
#define IRTT_OCCL_REFILL 32
#endif

// The edge records' occlusion rays.  Each lane walks the record slots
// lane, lane + stride, ... (empty slots of partly filled chunks and records
// of photons at or above g_limit skipped) and holds one ray at a time as a
// resumable closest-hit query; the wave walks the tree until IRTT_OCCL_REFILL
// lanes hold an answer, whose terms are then added and whose next records
// taken while the other lanes keep their place in the tree.  The gradient
// sums are exact integers, so the order of the rays does not matter.
template <int STACK>
__global__ __launch_bounds__(kTraceBlock, occl_waves(STACK)) void k_adj_occl(OcclArgs oa)
{
    const AdjCommon &a = oa.c;
    unsigned long long *const gx = a.grad + xcc_id() * a.grad_stride;
    __shared__ uint32_t stk_n[STACK * kTraceBlock];
    __shared__ float stk_t[stack_t_entries(STACK) * kTraceBlock];
    uint32_t *sn = stk_n + threadIdx.x;
    float *stt = stk_t + threadIdx.x;
    const bool f_gain = a.aflags & IRTT_ADJ_EDGE_GAIN;
    uint32_t n_edge = 0, n_occl = 0, nvis = 0, ntri = 0;
    const uint64_t n_slots = (uint64_t)oa.n_chunks * kChunk;
    const uint64_t stride = (uint64_t)gridDim.x * kTraceBlock;
    uint64_t r = (uint64_t)blockIdx.x * kTraceBlock + threadIdx.x;
    EdgeRec er{};
    Query q{};
    q.cur = kQueryDone;
    bool open = false;  // this lane holds a record whose ray has not been answered and used
    for (;;) {
        if (open && q.cur == kQueryDone) {
            occl_terms(a, gx, er, query_end(q, 1e12f), f_gain, n_occl);
            open = false;
        }
        if (!open) {
            while (r < n_slots) {
                const uint64_t slot = r;
                r += stride;
                if ((uint32_t)(slot % kChunk) >= oa.fills[slot / kChunk]) continue;
                er = oa.edges[slot];
                if ((uint64_t)er.r3.w >= oa.g_limit) continue;
                n_edge++;
                const V3 to{er.r0.x, er.r0.y, er.r0.z}, od{er.r1.x, er.r1.y, er.r1.z};
                if (!query_begin(q, a.tris, a.n_leaves, to, od, 1e-12f, 1e12f, ntri))
                    q.cur = kQueryDone;
                open = true;
                break;
            }
        }
        if (__ballot(open) == 0) break;
        query_run_until_ref(q, open, a.nodes, a.tris, 1e-12f, sn, nvis, ntri, IRTT_OCCL_REFILL);
    }
    uint32_t cnt[2] = {n_edge, n_occl};
    const int slot_of[2] = {A_EDGE, A_OCCL};
#pragma unroll
    for (int k = 0; k < 2; k++) {
        uint32_t v = cnt[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (threadIdx.x == 0 && v) atomicAdd(a.stats + slot_of[k], (unsigned long long)v);
    }
}

// The source-area terms (thermal_tracing_unified.glsl:327-351 and :588-592),
// one block per source triangle: every replayed photon adds
// -lambda[src] x4[src] dw/dv at emission and, if absorbed,
// +lambda[abs] x4[src] dw/dv; dw/dv = w dA/dv / A is the triangle's own, so
// the factors are summed over its photons first (in a fixed order) and
// multiplied once.  Also counts the absorbed photons.
__global__ __launch_bounds__(256) void k_adj_area(AdjCommon a, uint32_t do_area)
{
    const uint32_t t = blockIdx.x;
    unsigned long long *const gx = a.grad + xcc_id() * a.grad_stride;
    __shared__ double sm[256];
    __shared__ unsigned int sabs;
    if (threadIdx.x == 0) sabs = 0;
    __syncthreads();
    const uint32_t i0 = a.idx[3 * t], i1 = a.idx[3 * t + 1], i2 = a.idx[3 * t + 2];
    const V3 p0 = load_v(a.xyz, i0), p1 = load_v(a.xyz, i1), p2 = load_v(a.xyz, i2);
    float len;
    (void)normalize3(cross3(p1 - p0, p2 - p0), len);
    const bool live = len > 0.0f;
    const bool dropped_too = a.aflags & IRTT_ADJ_REFERENCE_DROPPED;
    double c = 0.0;
    uint32_t nabs = 0;
    const uint64_t g0 = (uint64_t)t * a.per_tri;
    for (uint32_t k = threadIdx.x; k < a.per_tri; k += blockDim.x) {
        const uint32_t r = a.absorber[g0 + k];
        if (r != 0xffffffffu) nabs++;
        if (!live || !do_area) continue;
        const uint32_t xx = k / a.rays_per_thread, nn = k - xx * a.rays_per_thread;
        uint32_t seed = tea16(a.nt * xx + t, a.batch_seed + nn);
        const float u0 = rnd(seed), u1 = rnd(seed);
        const float su = sqrt_rn(u0);
        const float b0 = 1.0f - su, b1 = su * (1.0f - u1), b2 = su * u1;
        const int sc = max_corner(b0, b1, b2);
        const uint32_t src = a.bin_mode ? t : (sc == 0 ? i0 : (sc == 1 ? i1 : i2));
        c -= hist_factor(a, src, src);
        if (r != 0xffffffffu) c += hist_factor(a, r, src);
        (void)dropped_too;
    }
    sm[threadIdx.x] = c;
    atomicAdd(&sabs, nabs);
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (sabs) atomicAdd(a.stats + A_ABS, (unsigned long long)sabs);
        if (live && do_area && sm[0] != 0.0) {
            V3 ga[3];
            area_grad(p0, p1, p2, ga);
            const float sa = a.w[t] / a.area[t];
            const uint32_t iv[3] = {i0, i1, i2};
            for (int cc = 0; cc < 3; cc++) {
                const V3 dw = scl3(ga[cc], sa);
                if (finite3(dw)) fx_add3(gx, iv[cc], dw, sm[0], a.inv_q);
            }
        }
    }
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

// Device bytes the adjoint trace's records take for a replay of `segments`
// accepted segments (segment records with their sort keys, edge records for
// up to half of them, per-wave partial chunks), for serving's admission.
uint64_t adjoint_record_bytes(uint64_t photons, uint64_t segments)
{
    const uint64_t waves = std::min<uint64_t>((photons + 1023) / 1024, 256ull * 128);
    return (segments + (waves + 1) * kChunk) * (sizeof(SegRec) + 16) +
           (segments / 2 + (waves + 1) * kChunk) * sizeof(EdgeRec) + 4 * photons;
}

void adjoint_trace(Scene &s, const irtt_adjoint_params &ap, uint32_t n_hist, const double *lambda,
                   const double *x, double *grad, irtt_adjoint_stats &st)
{
    if (!s.has_transport)
        throw Error(IRTT_ERR_STATE, "build the transport first: the adjoint replays its photons");
    if (n_hist == 0) throw Error(IRTT_ERR_ARG, "n_hist must be positive");
    if (!s.bvh.valid) bvh_build(s);
    const irtt_transport_params &tp = s.tparams;
    const uint32_t per = tp.threads_per_tri * tp.rays_per_thread;
    const uint32_t max_depth =
        (tp.max_depth == 0 || tp.max_depth > kMaxDepthDefault) ? kMaxDepthDefault : tp.max_depth;
    const uint32_t nb = s.nb;
    const uint64_t photons = (uint64_t)per * s.nt;
    if (photons > 0xffffffffull)
        throw Error(IRTT_ERR_LIMIT, "the adjoint trace replays at most 2^32 photons");
    std::vector<float> w(s.nt);
    float wmax = 0.0f;
    for (uint32_t t = 0; t < s.nt; t++) {
        float wt = s.tri_area[t] / (float)per;
        w[t] = wt * (kSigma * s.eps[t]);
        wmax = std::max(wmax, w[t]);
    }
    // fixed-point unit q: 2^-48 of the largest photon weight times the
    // largest history factor, as a power of two (exact scaling); every sum
    // is exact in units of q, a single contribution is rounded to q once
    double fmax = 0.0;
    for (uint32_t h = 0; h < n_hist; h++) {
        double lm = 0.0, xm = 0.0;
        for (uint32_t i = 0; i < nb; i++) {
            lm = std::max(lm, std::fabs(lambda[(size_t)h * nb + i]));
            xm = std::max(xm, std::fabs(x[(size_t)h * nb + i]));
        }
        fmax += lm * xm;
    }
    const double scale = fmax * (double)wmax;
    double inv_q = 1.0;
    if (scale > 0.0 && std::isfinite(scale)) {
        int e;
        std::frexp(scale, &e);  // scale < 2^e
        inv_q = std::ldexp(1.0, 48 - e);
    }

    hipStream_t stm = s.stream;
    AdjWork &ws = s.adj;
    DBuf<float> &d_w = ws.w, &d_area = ws.area;
    DBuf<double> &d_lam = ws.lam, &d_x4 = ws.x4;
    DBuf<unsigned long long> &d_grad = ws.grad, &d_st = ws.st;
    d_w.alloc(s.nt);
    d_area.alloc(s.nt);
    d_lam.alloc((size_t)n_hist * nb);
    d_x4.alloc((size_t)n_hist * nb);
    d_grad.alloc(kGradCopies * kCompWords * 3 * (size_t)s.nv);
    d_st.alloc(16);
    IRTT_HIP(hipMemcpyAsync(d_w.p, w.data(), s.nt * 4, hipMemcpyHostToDevice, stm));
    IRTT_HIP(hipMemcpyAsync(d_area.p, s.tri_area.data(), s.nt * 4, hipMemcpyHostToDevice, stm));
    IRTT_HIP(hipMemcpyAsync(d_lam.p, lambda, d_lam.bytes(), hipMemcpyHostToDevice, stm));
    IRTT_HIP(hipMemcpyAsync(d_x4.p, x, d_x4.bytes(), hipMemcpyHostToDevice, stm));
    IRTT_HIP(hipMemsetAsync(d_grad.p, 0, d_grad.bytes(), stm));
    IRTT_HIP(hipMemsetAsync(d_st.p, 0, d_st.bytes(), stm));

    IRTT_HIP(hipEventRecord(s.ev0, stm));
    // absorbing bins: the forward build's record, or a forward pass of our own
    DBuf<uint32_t> &d_abs_own = ws.abs_own;
    const uint32_t *absorber = nullptr;
    uint64_t rays = s.fwd_rays, segments = s.fwd_segments;
    if (s.d_absorber.n == photons) {
        absorber = s.d_absorber.p;
        d_abs_own.release();
    } else {
        d_abs_own.alloc(photons);
        absorber_pass(s, d_abs_own.p, rays, segments);
        absorber = d_abs_own.p;
    }

    AdjCommon c{};
    c.xyz = s.d_xyz.p;
    c.idx = s.d_idx.p;
    c.rho_d = s.d_rho_d.p;
    c.rho_s = s.d_rho_s.p;
    c.flags = s.d_flags.p;
    c.w = d_w.p;
    c.area = d_area.p;
    c.nodes = s.bvh.nodes.p;
    c.tris = s.bvh.tris.p;
    c.n_leaves = s.bvh.n_leaves;
    c.nt = s.nt;
    c.nb = nb;
    c.bin_mode = s.bin_mode;
    c.per_tri = per;
    c.rays_per_thread = tp.rays_per_thread;
    c.batch_seed = tp.batch_seed;
    c.max_depth = max_depth;
    c.tiny_dist = tp.tiny_dist;
    c.eps = ap.tiny_eps > 0 ? ap.tiny_eps : 1e-2f;
    c.ratio = ap.edge_ratio > 0 ? ap.edge_ratio : 0.05f;
    c.aflags = ap.flags ? ap.flags : (IRTT_ADJ_AREA | IRTT_ADJ_CONTINUOUS | IRTT_ADJ_EDGE);
    c.n_hist = n_hist;
    c.lambda = d_lam.p;
    c.x4 = d_x4.p;
    c.absorber = absorber;
    c.photons = photons;
    c.grad = d_grad.p;
    c.grad_stride = kCompWords * 3 * (uint64_t)s.nv;
    c.inv_q = inv_q;
    c.stats = d_st.p;

    int dev = 0, ncu = 256;
    IRTT_HIP(hipGetDevice(&dev));
    IRTT_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int S = s.bvh.max_depth + 1 <= 24 ? 24 : stack_size_for(s.bvh.max_depth);

    // A second stream runs the stages that need no other stage's output
    // beside the ones that do: the source-area terms beside the path replay,
    // each range's occlusion rays beside its sorted segment terms.  Every
    // gradient sum is exact, so the overlap does not change a bit.
    struct Side {
        hipStream_t st = nullptr;
        hipEvent_t fork = nullptr, area_done = nullptr, occl_done = nullptr;
        ~Side()
        {
            for (hipEvent_t e : {fork, area_done, occl_done})
                if (e) (void)hipEventDestroy(e);
            if (st) stream_give(st);
        }
    } side2;
    side2.st = stream_take();
    for (hipEvent_t *e : {&side2.fork, &side2.area_done, &side2.occl_done})
        IRTT_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    hipStream_t stm2 = side2.st;
    bool occl_pending = false;

    // source-area terms and the absorbed count
    IRTT_HIP(hipEventRecord(side2.fork, stm));
    IRTT_HIP(hipStreamWaitEvent(stm2, side2.fork, 0));
    k_adj_area<<<s.nt, 256, 0, stm2>>>(c, (c.aflags & IRTT_ADJ_AREA) ? 1u : 0u);
    IRTT_HIP(hipGetLastError());
    IRTT_HIP(hipEventRecord(side2.area_done, stm2));

    const bool want_edges = c.aflags & IRTT_ADJ_EDGE;
    const bool dropped_too = c.aflags & IRTT_ADJ_REFERENCE_DROPPED;
    uint32_t key_bits = 1;
    while (key_bits < 32 && (1ull << key_bits) <= s.nt) key_bits++;
    DBuf<uint32_t> &d_fills = ws.fills, &d_keys = ws.keys, &d_keys2 = ws.keys2,
                   &d_order = ws.order, &d_order2 = ws.order2;
    DBuf<unsigned long long> &d_ctl = ws.ctl;  // chunks taken (segments, edges), first dropped
    DBuf<uint8_t> &d_sort_tmp = ws.sort_tmp;
    size_t sort_tmp_b = 0;
    ws.n_big.alloc(1);
    static_assert(sizeof(SegRec) == 3 * sizeof(uint4) && sizeof(EdgeRec) == 4 * sizeof(uint4),
                  "record layouts");

    // The terms of one set of records: sc segment chunks and ec edge chunks
    // (raw edge records in edges_in, their ray form into edges_out), records
    // of photons at or above g_limit skipped.  The occlusion rays run on the
    // side stream beside the sorted segment terms.
    auto terms = [&](SegRec *segs_p, const uint32_t *seg_fills, uint32_t sc,
                     const EdgeRec *edges_in, EdgeRec *edges_out, const uint32_t *edge_fills,
                     uint32_t ec, uint64_t g_limit, bool kept_records) {
        if (ec > 0) {
            OcclArgs oa{};
            oa.c = c;
            oa.edges = edges_out;
            oa.fills = edge_fills;
            oa.n_chunks = ec;
            oa.g_limit = g_limit;
            const uint32_t og = (uint32_t)std::max<uint64_t>(
                1, std::min<uint64_t>((uint64_t)ec * (kChunk / 64), (uint64_t)ncu * 64));
            // after the records are complete, beside the segment terms
            // enqueued next on the scene's stream (measured: beside the start
            // side's terms instead, both slow to 4.1 and 4.9 ms and the trace
            // takes 0.3 ms longer)
            IRTT_HIP(hipEventRecord(side2.fork, stm));
            IRTT_HIP(hipStreamWaitEvent(stm2, side2.fork, 0));
            k_edge_prep<<<cdiv((uint64_t)ec * kChunk, 256), 256, 0, stm2>>>(
                c, edges_in, edges_out, edge_fills, ec, kept_records ? 1u : 0u);
            IRTT_HIP(hipGetLastError());
            if (S == 24) k_adj_occl<24><<<og, kTraceBlock, 0, stm2>>>(oa);
            else if (S == 32) k_adj_occl<32><<<og, kTraceBlock, 0, stm2>>>(oa);
            else if (S == 64) k_adj_occl<64><<<og, kTraceBlock, 0, stm2>>>(oa);
            else k_adj_occl<128><<<og, kTraceBlock, 0, stm2>>>(oa);
            IRTT_HIP(hipGetLastError());
            IRTT_HIP(hipEventRecord(side2.occl_done, stm2));
            occl_pending = true;
        }
        if (sc == 0) return;
        // each side over the records sorted by its triangle
        const uint64_t n = (uint64_t)sc * kChunk;
        for (uint32_t side = 0; side < 2; side++) {
            k_seg_keys<<<cdiv(n, 256), 256, 0, stm>>>(
                segs_p, seg_fills, sc, s.nt, side, kept_records && !dropped_too ? 1u : 0u,
                kept_records && side == 0 ? absorber : nullptr, d_keys.p, d_order.p);
            IRTT_HIP(hipGetLastError());
            rocprim::double_buffer<uint32_t> kd(d_keys.p, d_keys2.p), vd(d_order.p, d_order2.p);
            size_t tb = sort_tmp_b;
            IRTT_HIP(rocprim::radix_sort_pairs(d_sort_tmp.p, tb, kd, vd, (size_t)n, 0, key_bits,
                                               stm));
            SegArgs sa{};
            sa.c = c;
            sa.segs = segs_p;
            sa.keys = kd.current();
            sa.order = vd.current();
            sa.n = n;
            sa.side = side;
            sa.g_limit = g_limit;
            // the sort's other key buffer lists the records too large for
            // the window sums
            sa.big = kd.current() == d_keys.p ? d_keys2.p : d_keys.p;
            sa.n_big = ws.n_big.p;
            IRTT_HIP(hipMemsetAsync(ws.n_big.p, 0, sizeof(uint32_t), stm));
            const uint32_t sg = (uint32_t)std::max<uint64_t>(
                1, std::min<uint64_t>(cdiv(n / 64, 4), (uint64_t)ncu * 16));
            if (side == 0) {
                k_adj_segments<0><<<sg, 256, 0, stm>>>(sa);
                k_adj_seg_big<0><<<ncu, 256, 0, stm>>>(sa);
            } else {
                k_adj_segments<1><<<sg, 256, 0, stm>>>(sa);
                k_adj_seg_big<1><<<ncu, 256, 0, stm>>>(sa);
            }
            IRTT_HIP(hipGetLastError());
        }
    };
    // sort keys, orders and the sort's work space for n_slots segment slots
    auto alloc_sort = [&](uint64_t n_slots) {
        d_keys.alloc(n_slots);
        d_keys2.alloc(n_slots);
        d_order.alloc(n_slots);
        d_order2.alloc(n_slots);
        sort_tmp_b = 0;
        rocprim::double_buffer<uint32_t> kb(nullptr, nullptr), vb(nullptr, nullptr);
        IRTT_HIP(rocprim::radix_sort_pairs(nullptr, sort_tmp_b, kb, vb, (size_t)n_slots, 0,
                                           key_bits));
        d_sort_tmp.alloc(std::max<size_t>(sort_tmp_b, 1));
    };

    // Paths the build kept (IRTT_TRANSPORT_KEEP_PATHS): their records stand
    // for the replay of every photon below the first one with a record that
    // did not fit; the photons from there on are replayed.
    uint64_t g_first = 0;
    PathStore &ps = s.paths;
    const bool kept = ps.valid && ps.photons == photons && absorber == s.d_absorber.p &&
                      c.ratio == kKeptEdgeRatio;
    st.kept_paths = 0;
    if (kept) {
        const uint32_t sc = (uint32_t)ps.sc, ec = want_edges ? (uint32_t)ps.ec : 0u;
        alloc_sort(std::max<uint64_t>((uint64_t)sc * kChunk, 1));
        if (ec > 0) ws.edges.alloc((size_t)ec * kChunk * 4);
        g_first = std::min<uint64_t>(ps.first_dropped, photons);
        const EdgeRec *kept_edges = reinterpret_cast<const EdgeRec *>(ps.edges.p);
        terms(reinterpret_cast<SegRec *>(ps.segs.p), ps.fills.p, sc, kept_edges,
              reinterpret_cast<EdgeRec *>(ws.edges.p), ps.fills.p + ps.seg_chunks, ec, g_first,
              true);
        st.kept_paths = g_first < photons ? 2u : 1u;
        if (g_first < photons) {
            // the replay below reallocates the buffers these stages use
            IRTT_HIP(hipStreamSynchronize(stm));
            IRTT_HIP(hipStreamSynchronize(stm2));
            occl_pending = false;
        }
    }

    // Path replay into segment and edge records, then their terms.  The
    // records of all photons fit when the device has room (the forward
    // build's accepted segments bound them; edge records are taken as at most
    // half of those); otherwise, and whenever a buffer overflows, the photons
    // are replayed in ranges, each range's terms computed up to its first
    // photon with a dropped record.
    if (g_first < photons) {
        // replay grid: at most IRTT_PHOTON_WAVES_PER_CU_ADJ waves per CU, and
        // at least 16 photons per lane (each wave's last chunk is partly empty)
        const uint64_t left = photons - g_first;
        const uint64_t waves = std::max<uint64_t>(
            1, std::min<uint64_t>((uint64_t)ncu * IRTT_PHOTON_WAVES_PER_CU_ADJ,
                                  (left + 16 * kTraceBlock - 1) / (16 * kTraceBlock)));
        const uint64_t seg_bytes = sizeof(SegRec) + 4 * sizeof(uint32_t);  // + sort keys and order
        // the record buffers this scene already holds count as available;
        // when the device cannot give them (memory held elsewhere, or not in
        // pieces this large), the budget halves and the photons run in more
        // ranges
        double budget =
            std::max(2.0 * (1u << 30), 0.3 * ((double)dev_available() + (double)ws.record_bytes()));
        const uint64_t seg_est =
            (uint64_t)((double)segments * ((double)left / (double)photons)) + 1;
        uint64_t seg_chunks = 0, edge_chunks = 0, n_ranges = 1;
        for (;;) {
            seg_chunks = (seg_est + kChunk - 1) / kChunk + waves + 1;
            edge_chunks = want_edges ? (seg_est / 2 + kChunk - 1) / kChunk + waves + 1 : 1;
            const double need =
                (double)kChunk * (seg_chunks * seg_bytes + edge_chunks * sizeof(EdgeRec));
            n_ranges = 1;
            if (need > budget) {
                const double f = budget / need;
                seg_chunks = std::max<uint64_t>(2 * waves + 1, (uint64_t)(seg_chunks * f));
                edge_chunks = want_edges
                                  ? std::max<uint64_t>(2 * waves + 1, (uint64_t)(edge_chunks * f))
                                  : 1;
                n_ranges = (uint64_t)std::ceil(1.0 / f);
            }
            try {
                ws.segs.alloc(seg_chunks * kChunk * 3);
                ws.edges.alloc(edge_chunks * kChunk * 4);
                d_fills.alloc(seg_chunks + edge_chunks);
                alloc_sort(seg_chunks * kChunk);
                break;
            } catch (const Error &e) {
                if (e.code != IRTT_ERR_NOMEM || budget <= 2.0 * (1u << 30)) throw;
                for (auto *b : {&d_fills, &d_keys, &d_keys2, &d_order, &d_order2}) b->release();
                ws.segs.release();
                ws.edges.release();
                d_sort_tmp.release();
                budget = std::max(2.0 * (1u << 30), 0.5 * budget);
            }
        }
        d_ctl.alloc(3);
        SegRec *const segs_p = reinterpret_cast<SegRec *>(ws.segs.p);
        EdgeRec *const edges_p = reinterpret_cast<EdgeRec *>(ws.edges.p);
        const uint64_t range = (left + n_ranges - 1) / n_ranges;
        static const unsigned long long ctl0[3] = {0ull, 0ull, ~0ull};
        // one range from the first photon, photons without an absorbing bin
        // not replayed: the replay's first launch takes its photons from the
        // list of those with one
        const uint32_t *live_p = nullptr;
        if (n_ranges == 1 && g_first == 0 && !dropped_too) {
            const uint32_t nblk = cdiv(photons, 256);
            ws.live.alloc(photons);
            ws.live_counts.alloc(2 * (size_t)nblk + 1);
            uint32_t *counts = ws.live_counts.p, *offsets = counts + nblk, *n_live = offsets + nblk;
            size_t tb = 0;
            IRTT_HIP(rocprim::exclusive_scan(nullptr, tb, counts, offsets, 0u, (size_t)nblk,
                                             rocprim::plus<uint32_t>(), stm));
            ws.scan_tmp.alloc(std::max<size_t>(tb, 1));
            k_live_count<<<nblk, 256, 0, stm>>>(absorber, photons, counts);
            IRTT_HIP(hipGetLastError());
            IRTT_HIP(rocprim::exclusive_scan(ws.scan_tmp.p, tb, counts, offsets, 0u, (size_t)nblk,
                                             rocprim::plus<uint32_t>(), stm));
            k_live_fill<<<nblk, 256, 0, stm>>>(absorber, photons, offsets, nblk, counts,
                                              ws.live.p, n_live);
            IRTT_HIP(hipGetLastError());
            live_p = ws.live.p;
        }
        for (uint64_t gb = g_first; gb < photons; gb += range) {
            const uint64_t ge = std::min<uint64_t>(photons, gb + range);
            uint64_t g_start = gb;
            for (;;) {
                PathArgs pa{};
                pa.c = c;
                pa.g_begin = g_start;
                pa.g_end = ge;
                pa.segs = RecBuf{segs_p, (uint32_t)seg_chunks, d_fills.p, d_ctl.p};
                pa.edges =
                    RecBuf{edges_p, (uint32_t)edge_chunks, d_fills.p + seg_chunks, d_ctl.p + 1};
                pa.first_dropped = d_ctl.p + 2;
                if (live_p && g_start == 0) {
                    pa.live = live_p;
                    pa.n_live = ws.live_counts.p + 2 * (size_t)cdiv(photons, 256);
                }
                // the previous range's occlusion rays end before this range's
                // replay rewrites the records
                if (occl_pending) IRTT_HIP(hipStreamWaitEvent(stm, side2.occl_done, 0));
                IRTT_HIP(hipMemcpyAsync(d_ctl.p, ctl0, sizeof(ctl0), hipMemcpyHostToDevice, stm));
                const uint32_t grid = (uint32_t)std::max<uint64_t>(
                    1, std::min<uint64_t>(cdiv(ge - g_start, kTraceBlock), waves));
                if (S == 24) k_adj_paths<24><<<grid, kTraceBlock, 0, stm>>>(pa);
                else if (S == 32) k_adj_paths<32><<<grid, kTraceBlock, 0, stm>>>(pa);
                else if (S == 64) k_adj_paths<64><<<grid, kTraceBlock, 0, stm>>>(pa);
                else k_adj_paths<128><<<grid, kTraceBlock, 0, stm>>>(pa);
                IRTT_HIP(hipGetLastError());
                unsigned long long ctl[3];
                IRTT_HIP(hipMemcpyAsync(ctl, d_ctl.p, sizeof(ctl), hipMemcpyDeviceToHost, stm));
                IRTT_HIP(hipStreamSynchronize(stm));
                const uint32_t sc = (uint32_t)std::min<uint64_t>(ctl[0], seg_chunks);
                const uint32_t ec =
                    want_edges ? (uint32_t)std::min<uint64_t>(ctl[1], edge_chunks) : 0;
                terms(segs_p, d_fills.p, sc, edges_p, edges_p, d_fills.p + seg_chunks, ec, ctl[2],
                      false);
                if (ctl[2] == ~0ull) break;  // every record of the range fitted
                if (ctl[2] <= g_start)
                    throw Error(IRTT_ERR_LIMIT, "adjoint trace: the records of one photon exceed "
                                                "the record buffers");
                g_start = ctl[2];
            }
        }
    }
    IRTT_HIP(hipStreamWaitEvent(stm, side2.area_done, 0));
    if (occl_pending) IRTT_HIP(hipStreamWaitEvent(stm, side2.occl_done, 0));
    IRTT_HIP(hipEventRecord(s.ev1, stm));
    std::vector<unsigned long long> gx(kGradCopies * kCompWords * 3 * (size_t)s.nv), hs(16);
    IRTT_HIP(hipMemcpyAsync(gx.data(), d_grad.p, d_grad.bytes(), hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipMemcpyAsync(hs.data(), d_st.p, 16 * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipStreamSynchronize(stm));
    IRTT_HIP(hipEventElapsedTime(&st.trace_ms, s.ev0, s.ev1));
    const double q = 1.0 / inv_q;
    const size_t words = kCompWords * 3 * (size_t)s.nv;
    for (size_t k = 0; k < 3 * (size_t)s.nv; k++) {
        unsigned __int128 sum = 0;  // exact: the copies' words added modulo 2^128
        for (uint32_t c = 0; c < kGradCopies; c++) {
            const unsigned long long *w = gx.data() + c * words + kCompWords * k;
            sum += (unsigned __int128)w[0] +
                   (unsigned __int128)((__int128)(long long)w[1] * ((__int128)1 << 32)) +
                   ((unsigned __int128)w[2] << 64);
        }
        grad[k] = (double)(__int128)sum * q;
    }
    st.photons = photons;
    st.absorbed = hs[A_ABS];
    st.segments = hs[A_SEG];
    st.continuous = hs[A_CONT];
    st.edge_hits = hs[A_EDGE];
    st.occluded = hs[A_OCCL];
    st.rays = rays;
}

}  // namespace irtt
