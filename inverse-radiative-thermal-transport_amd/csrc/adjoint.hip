// adjoint.hip -- adjoint photon trace: d(objective)/d(vertex positions).
//
// Replaces ThermalTransport::computeAdjoint (thermal_transport.cpp:740-832)
// and the raygen program in TRACING_MODE 2
// (assets/shader/raytracing_thermal/thermal_tracing_unified.glsl:292-599).
// Each photon is replayed from its seed: pass 0 finds the absorbing bin (or
// the forward build's per-photon record supplies it and pass 0 is skipped),
// pass 1 accumulates, weighted by sum_h lambda_h[bin] * x_h^4[source bin]
// (sumOverHistory, :227-244, indexed history-major so every history length
// is handled, not only one step):
//   (i)   -/+ dw/dv of the three source vertices at emission / absorption,
//   (ii)  w * d ln|G| / dv for the six vertices of every path segment, with
//         G = -(N_i.d)(N_j.d) / (|N_i||N_j|(r+eps)^4), d = p_j - p_i,
//   (iii) the approximate edge term: a hit within edge_ratio of an edge casts
//         one occlusion ray from just outside the edge; a front-facing,
//         non-adjacent occluded surface receives +, the path's origin vertex +,
//         the edge vertex - of the edge-irradiance derivative.
// Contributions are FP32 (as the shader computes them), scaled and summed in
// FP64 with hardware atomics on a 3 x n_vertices buffer; the source triangle's
// three vertices are summed in each lane's registers over its photons of that
// triangle first.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "device_math.hpp"
#include "irtt_internal.hpp"
#include "traverse.hpp"

namespace irtt {
namespace {

__device__ __forceinline__ V3 add3(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 scl3(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ bool finite3(V3 a)
{
    return isfinite(a.x) && isfinite(a.y) && isfinite(a.z);
}
__device__ __forceinline__ int min_corner(float b0, float b1, float b2)
{
    int idx = 0;
    float v = b0;
    if (b1 < v) {
        idx = 1;
        v = b1;
    }
    if (b2 < v) idx = 2;
    return idx;
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
// live at once (the adjoint kernel's register peak)
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

__device__ __forceinline__ void acc3(double *g, uint32_t v, V3 d, double f)
{
    unsafeAtomicAdd(g + 3 * (size_t)v, f * (double)d.x);
    unsafeAtomicAdd(g + 3 * (size_t)v + 1, f * (double)d.y);
    unsafeAtomicAdd(g + 3 * (size_t)v + 2, f * (double)d.z);
}

struct AdjArgs {
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
    const double *lambda;  // n_hist x nb
    const double *x4;      // n_hist x nb
    uint64_t photons;
    uint32_t run;              // photons per lane per chunk, 1 .. kAdjRun
    const uint32_t *absorber;  // per photon from the forward build, or nullptr
    double *grad;
    unsigned long long *stats;
};

enum { A_ABS = 0, A_SEG, A_CONT, A_EDGE, A_OCCL, A_RAYS, A_COUNT };

__device__ __forceinline__ double hist_factor(const AdjArgs &a, uint32_t bin_l, uint32_t bin_x)
{
    double f = 0.0;
    for (uint32_t h = 0; h < a.n_hist; h++)
        f += a.lambda[(size_t)h * a.nb + bin_l] * a.x4[(size_t)h * a.nb + bin_x];
    return f;
}

// Photons per lane per chunk (a.run, at most kAdjRun): a wave takes chunks of
// 64 * a.run consecutive photons, lane l the run [l * a.run, (l + 1) * a.run)
// of each, so a lane's photons nearly always share their source triangle and
// the source vertices' sums stay in the lane's registers between photons.
// The host shortens the runs when there are too few photons to give every
// wave of the grid a chunk.
// (city block, 2^26 photons: 16 -> 37.5-38.3 ms, 32 -> 35.0-36.0, 64 -> 32.7-33.4)
#ifndef IRTT_ADJ_RUN
#define IRTT_ADJ_RUN 64
#endif
constexpr uint32_t kAdjRun = IRTT_ADJ_RUN;

// what a lane does next: start a photon, trace a path segment, trace an
// occlusion ray, end the current pass's path, nothing (no photons left)
enum { M_NONE = 0, M_PATH, M_OCCL, M_END, M_DONE };

// waves per SIMD: IRTT_ADJ_WAVES, or what the LDS stack admits (160 KiB per
// CU, four SIMDs).  Measured on the city block (2^26 photons, all terms):
// 3 waves (168 VGPRs, 48 B spilled per lane) 36.6 ms, 4 waves (128 VGPRs,
// 240 B spilled) 35.0-35.3 ms.  The area terms are summed when a photon ends
// and the counters are kept per wave, so that little stays live across the
// path; with the area state carried along, 4 waves spilled 368 B and took
// 56.5 ms.
#ifndef IRTT_ADJ_WAVES
#define IRTT_ADJ_WAVES 4
#endif
constexpr int adjoint_waves(int stack)
{
    return (160 * 1024 / (stack * 4 * kTraceBlock)) / 4 < IRTT_ADJ_WAVES
               ? (160 * 1024 / (stack * 4 * kTraceBlock)) / 4
               : IRTT_ADJ_WAVES;
}

// One loop iteration traces one ray per lane: a segment of the photon path
// being replayed, or the occlusion ray of the previous segment's edge term.
// A lane whose photon ends starts its next photon in the next iteration, so
// lanes do not wait for the longest path of the wave, and occlusion rays are
// traced together with other lanes' path segments instead of by the few
// lanes that need one.
template <int STACK>
__global__ __launch_bounds__(kTraceBlock, adjoint_waves(STACK)) void k_adjoint(AdjArgs a)
{
    __shared__ uint32_t stk_n[STACK * kTraceBlock];
    __shared__ float stk_t[stack_t_entries(STACK) * kTraceBlock];
    uint32_t *sn = stk_n + threadIdx.x;
    float *stt = stk_t + threadIdx.x;
    const uint32_t lane = threadIdx.x;
    // per-wave event counts (A_ABS per lane: several photons may start in
    // one iteration); a lane marks an iteration's events in ev (bit k = slot k)
    uint32_t wcnt[A_COUNT] = {0, 0, 0, 0, 0, 0};
    uint32_t n_abs = 0;
    uint32_t nvis = 0, ntri = 0;
    const bool f_area = a.aflags & IRTT_ADJ_AREA, f_cont = a.aflags & IRTT_ADJ_CONTINUOUS;
    const bool f_edge = a.aflags & IRTT_ADJ_EDGE, f_gain = a.aflags & IRTT_ADJ_EDGE_GAIN;
    const bool f_hit = a.aflags & IRTT_ADJ_HIT_AREA, f_proj = a.aflags & IRTT_ADJ_EDGE_PROJECTED;

    // this lane's run of photons: the next is run_g, `left` remain, in chunk `chunk`
    uint32_t chunk = blockIdx.x;
    const uint32_t run = a.run;
    uint64_t run_g = (uint64_t)chunk * kTraceBlock * run + (uint64_t)lane * run;
    uint32_t left = run_g < a.photons ? (uint32_t)min<uint64_t>(run, a.photons - run_g) : 0u;
    // the source vertices' gradient of this lane's photons of triangle acc_t
    uint32_t acc_t = 0xffffffffu;
    double sacc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};

    // the photon being replayed
    uint64_t g = 0;
    uint32_t t = 0, src = 0, abs_bin = 0;
    bool absorbed_any = false;
    int pass = 0;
    double fpath = 0.0;
    // area terms: emission side (pass 1 started on a non-degenerate emitter),
    // absorption side (absorbed in pass 1); summed when the photon ends
    bool area_on = false, area_abs = false;
    // its path
    uint32_t seed = 0, depth = 0, prev0 = 0, prev1 = 0, prev2 = 0;
    int prev_max = 0;
    float pb1 = 0.0f, pb2 = 0.0f;
    V3 o{0.0f, 0.0f, 0.0f}, d{0.0f, 0.0f, 1.0f};
    bool path_on = false;  // a next path segment follows the pending occlusion ray
    // the pending occlusion ray's edge term
    V3 od{0.0f, 0.0f, 1.0f}, dE{0.0f, 0.0f, 0.0f};
    float oht = 0.0f;
    uint32_t ohp = 0, edge_v = 0, origin_v = 0;  // ohp: the segment's hit triangle

    int mode = M_NONE;
    Query q;
    bool tracing = false;
    for (;;) {
        // ---- lanes without a ray: end the current pass or photon, start the next ----
        // ph: 0 = nothing pending, 1 = the current pass's path has ended, 2 = start a photon
        int ph = mode == M_NONE ? 2 : (mode == M_END ? 1 : 0);
        while (ph != 0) {
            if (ph == 1) {
                // the current pass's path has ended
                if (pass == 0) {
                    pass = 1;
                } else {
                    // the photon's area terms on its source vertices,
                    // thermal_tracing_unified.glsl:327-351 and :588-592
                    const double c_area =
                        area_on ? -hist_factor(a, src, src) + (area_abs ? fpath : 0.0) : 0.0;
                    if (c_area != 0.0) {
                        const uint32_t v0 = a.idx[3 * t], v1 = a.idx[3 * t + 1], v2 = a.idx[3 * t + 2];
                        V3 ga[3];
                        area_grad(V3{a.xyz[3 * v0], a.xyz[3 * v0 + 1], a.xyz[3 * v0 + 2]},
                                  V3{a.xyz[3 * v1], a.xyz[3 * v1 + 1], a.xyz[3 * v1 + 2]},
                                  V3{a.xyz[3 * v2], a.xyz[3 * v2 + 1], a.xyz[3 * v2 + 2]}, ga);
                        const float sa = a.w[t] / a.area[t];
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            const V3 dw = scl3(ga[c], sa);
                            if (!finite3(dw)) continue;
                            sacc[3 * c] += c_area * (double)dw.x;
                            sacc[3 * c + 1] += c_area * (double)dw.y;
                            sacc[3 * c + 2] += c_area * (double)dw.z;
                        }
                    }
                    ph = 2;
                    continue;
                }
            } else {
                // next photon of this lane
                while (left == 0) {
                    chunk += gridDim.x;
                    const uint64_t base = (uint64_t)chunk * kTraceBlock * run;
                    if (base >= a.photons) break;
                    run_g = base + (uint64_t)lane * run;
                    left = run_g < a.photons ? (uint32_t)min<uint64_t>(run, a.photons - run_g)
                                             : 0u;
                }
                if (left == 0) {
                    mode = M_DONE;
                    break;
                }
                g = run_g++;
                left--;
                t = (uint32_t)(g / a.per_tri);
                if (t != acc_t) {
                    // another source triangle: flush the sums of the last one
                    if (acc_t != 0xffffffffu) {
#pragma unroll
                        for (int e = 0; e < 9; e++)
                            if (sacc[e] != 0.0)
                                unsafeAtomicAdd(a.grad + 3 * (size_t)a.idx[3 * acc_t + e / 3] + e % 3,
                                                sacc[e]);
                    }
#pragma unroll
                    for (int e = 0; e < 9; e++) sacc[e] = 0.0;
                    acc_t = t;
                }
                area_on = area_abs = false;
                absorbed_any = false;
                abs_bin = 0;
                pass = 0;
                // pass 0 finds the absorbing bin; the forward build's record saves it
                if (a.absorber) {
                    const uint32_t r = a.absorber[g];
                    absorbed_any = r != 0xffffffffu;
                    abs_bin = absorbed_any ? r : 0u;
                    if (absorbed_any) n_abs++;
                    pass = 1;
                }
            }
            // start pass `pass` of photon g: its emission ray
            // (getEmissionRay, thermal_tracing_unified.glsl:125-161)
            const uint32_t i0 = a.idx[3 * t], i1 = a.idx[3 * t + 1], i2 = a.idx[3 * t + 2];
            const V3 p0{a.xyz[3 * i0], a.xyz[3 * i0 + 1], a.xyz[3 * i0 + 2]};
            const V3 p1{a.xyz[3 * i1], a.xyz[3 * i1 + 1], a.xyz[3 * i1 + 2]};
            const V3 p2{a.xyz[3 * i2], a.xyz[3 * i2 + 1], a.xyz[3 * i2 + 2]};
            {
                // thermal_tracing_unified.glsl:291 (photon g = run_g - 1)
                const uint32_t k = (uint32_t)(run_g - 1 - (uint64_t)t * a.per_tri);
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
            if (!(len > 0.0f)) {
                // degenerate emitter: no contribution
                area_on = false;
                pass = 1;
                ph = 1;
                continue;
            }
            const int sc = max_corner(b0, b1, b2);
            src = a.bin_mode ? t : (sc == 0 ? i0 : (sc == 1 ? i1 : i2));
            if (pass == 1 && f_area) area_on = true;  // emission-side area term
            if (pass == 1 && !absorbed_any && !(a.aflags & IRTT_ADJ_REFERENCE_DROPPED)) {
                ph = 1;
                continue;
            }
            fpath = pass ? hist_factor(a, abs_bin, src) : 0.0;
            prev0 = i0;
            prev1 = i1;
            prev2 = i2;
            pb1 = b1;
            pb2 = b2;
            prev_max = sc;
            depth = 1;
            mode = M_PATH;
            tracing = query_begin(q, a.tris, a.n_leaves, o, d, 1e-12f, 1e12f, ntri);
            ph = 0;
        }
        if (__ballot(mode != M_DONE) == 0) break;

        // ---- every lane with a ray traces it ----
        query_run(q, tracing && mode != M_DONE, a.nodes, a.tris, 1e-12f, sn, stt, nvis, ntri);
        uint32_t ev = 0;
        if (mode != M_DONE) {
        const Hit h = query_end(q, 1e12f);
        tracing = false;
        int next = 1;  // 0: the pass's path ends, 1: its next segment, 2: occlusion ray begun

        if (mode == M_OCCL) {
            // the edge term of the previous segment, thermal_tracing_unified.glsl:454-567
            const bool valid = h.prim >= 0 && h.t >= a.tiny_dist;
            bool nbr = false;
            uint32_t shadow_v = 0;
            if (valid) {
                const uint32_t op = (uint32_t)h.prim;
                const uint32_t c0 = a.idx[3 * op], c1 = a.idx[3 * op + 1], c2 = a.idx[3 * op + 2];
                const uint32_t oj0 = a.idx[3 * ohp], oj1 = a.idx[3 * ohp + 1],
                               oj2 = a.idx[3 * ohp + 2];
                nbr = c0 == oj0 || c0 == oj1 || c0 == oj2 || c1 == oj0 || c1 == oj1 || c1 == oj2 ||
                      c2 == oj0 || c2 == oj1 || c2 == oj2;
                const V3 s0{a.xyz[3 * c0], a.xyz[3 * c0 + 1], a.xyz[3 * c0 + 2]};
                const V3 s1{a.xyz[3 * c1], a.xyz[3 * c1 + 1], a.xyz[3 * c1 + 2]};
                const V3 s2{a.xyz[3 * c2], a.xyz[3 * c2 + 1], a.xyz[3 * c2 + 2]};
                const V3 tn = normalize3(cross3(s1 - s0, s2 - s0));
                const float ob0 = 1.0f - h.b1 - h.b2;
                const int sm = max_corner(ob0, h.b1, h.b2);
                shadow_v = sm == 0 ? c0 : (sm == 1 ? c1 : c2);
                if (!nbr && dot3(tn, od) < -1e-3f) {
                    ev |= 1u << A_OCCL;
                    const uint32_t shadow_bin = a.bin_mode ? op : shadow_v;
                    const float sa = 1.0f - a.rho_d[op] - a.rho_s[op];
                    const float d1 = oht, d2 = h.t, tot = d1 + d2;
                    const double fs = hist_factor(a, shadow_bin, src);
                    acc3(a.grad, shadow_v, scl3(dE, sa * (d1 / tot)), fs);
                    acc3(a.grad, origin_v, scl3(dE, sa * (d2 / tot)), fs);
                    acc3(a.grad, edge_v, scl3(dE, -sa), fs);
                }
            }
            // Extension (not in the reference): the power the moving edge
            // takes from the shadow side lands on the hit triangle and follows
            // this photon's path, so it is credited to the path's absorbing
            // bin with the opposite sign, also when nothing lies behind the
            // edge (the power escaped before).
            if (f_gain && (h.prim < 0 || (valid && !nbr))) {
                const float f_or = valid ? h.t / (oht + h.t) : 1.0f;
                acc3(a.grad, edge_v, dE, fpath);
                acc3(a.grad, origin_v, scl3(dE, f_or), -fpath);
                if (valid) acc3(a.grad, shadow_v, scl3(dE, 1.0f - f_or), -fpath);
            }
            next = path_on ? 1 : 0;
        } else {
            // a path segment
            if (pass == 0) ev |= 1u << A_RAYS;
            bool stop = h.prim < 0 || h.t < a.tiny_dist;
            uint32_t hp = 0, j0 = 0, j1 = 0, j2 = 0;
            V3 q0{0.0f, 0.0f, 0.0f}, q1 = q0, q2 = q0, hn = q0;
            if (!stop) {
                hp = (uint32_t)h.prim;
                j0 = a.idx[3 * hp];
                j1 = a.idx[3 * hp + 1];
                j2 = a.idx[3 * hp + 2];
                q0 = V3{a.xyz[3 * j0], a.xyz[3 * j0 + 1], a.xyz[3 * j0 + 2]};
                q1 = V3{a.xyz[3 * j1], a.xyz[3 * j1 + 1], a.xyz[3 * j1 + 2]};
                q2 = V3{a.xyz[3 * j2], a.xyz[3 * j2 + 1], a.xyz[3 * j2 + 2]};
                hn = normalize3(cross3(q1 - q0, q2 - q0));
                stop = dot3(hn, d) > -1e-3f;
            }
            bool occl = false, ended = false;
            if (stop) {
                ended = true;
            } else {
                const float hb0 = 1.0f - h.b1 - h.b2;
                if (pass == 1) {
                    ev |= 1u << A_SEG;
                    if (f_cont) {
                        // continuous geometric term, :391-448
                        const V3 r0{a.xyz[3 * prev0], a.xyz[3 * prev0 + 1], a.xyz[3 * prev0 + 2]};
                        const V3 r1{a.xyz[3 * prev1], a.xyz[3 * prev1 + 1], a.xyz[3 * prev1 + 2]};
                        const V3 r2{a.xyz[3 * prev2], a.xyz[3 * prev2 + 1], a.xyz[3 * prev2 + 2]};
                        const uint32_t vid[6] = {prev0, prev1, prev2, j0, j1, j2};
                        const bool first = depth == 1;
                        const float wt = a.w[t];
                        const float G = geom_grad(
                            r0, r1, r2, q0, q1, q2, pb1, pb2, h.b1, h.b2, a.eps,
                            [&](int v, V3 g) {
                                const V3 dv = scl3(g, wt);
                                if (!finite3(dv)) return;
                                if (v < 3 && first) {
                                    // the first segment starts on the source triangle
                                    sacc[3 * v] += fpath * (double)dv.x;
                                    sacc[3 * v + 1] += fpath * (double)dv.y;
                                    sacc[3 * v + 2] += fpath * (double)dv.z;
                                } else {
                                    acc3(a.grad, vid[v], dv, fpath);
                                }
                            });
                        if (fabsf(G) > a.eps * a.eps) ev |= 1u << A_CONT;
                    }
                    if (f_hit) {
                        // extension: the hit triangle's area Jacobian (the hit
                        // point is a fixed barycentric point of it)
                        V3 ga[3];
                        area_grad(q0, q1, q2, ga);
                        const float sh = a.w[t] / a.area[hp];
                        const uint32_t jv[3] = {j0, j1, j2};
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            const V3 dh = scl3(ga[c], sh);
                            if (finite3(dh)) acc3(a.grad, jv[c], dh, fpath);
                        }
                    }
                    const float hb[3] = {hb0, h.b1, h.b2};
                    const int opp = min_corner(hb0, h.b1, h.b2);
                    if (f_edge && hb[opp] <= a.ratio) {
                        // approximate edge term, :454-567: its occlusion ray
                        // is traced in the next iteration
                        ev |= 1u << A_EDGE;
                        const int nr = max_corner(hb0, h.b1, h.b2);
                        int other = 0;
                        if (opp == 0 || nr == 0) {
                            other = 1;
                            if (opp == 1 || nr == 1) other = 2;
                        }
                        const V3 qv[3] = {q0, q1, q2};
                        const uint32_t jv[3] = {j0, j1, j2};
                        edge_v = jv[nr];
                        V3 edir = qv[nr] - qv[other];
                        const V3 oppe = qv[nr] - qv[opp];
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
                        // the reference scales by |da.n_edg| (:486); the
                        // extension divides by it: an edge moved by u sweeps
                        // (u.da)/|da.n_edg| of the occluder's plane per unit
                        // length, seen along the ray
                        const float cpl = fabsf(dot3(da, n_edg));
                        dE = f_proj ? scl3(da, elen * a.w[t] / (larea * cpl))
                                    : scl3(da, cpl * elen * a.w[t] / larea);
                        if (!finite3(dE)) dE = V3{0.0f, 0.0f, 0.0f};
                        V3 to = bary_point(q0, q1, q2, hb0, h.b1, h.b2);
                        to = add3(to, scl3(n_edg, a.eps + dot3(qv[nr] - to, n_edg)));
                        to = to - scl3(d, a.eps);
                        origin_v = prev_max == 0 ? prev0 : (prev_max == 1 ? prev1 : prev2);
                        ohp = hp;
                        oht = h.t;
                        od = d;
                        occl = true;
                        tracing = query_begin(q, a.tris, a.n_leaves, to, d, 1e-12f, 1e12f, ntri);
                    }
                    prev0 = j0;
                    prev1 = j1;
                    prev2 = j2;
                    pb1 = h.b1;
                    pb2 = h.b2;
                    prev_max = max_corner(hb0, h.b1, h.b2);
                }
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
                if (absorbed || depth == a.max_depth) {
                    const int hc = max_corner(hb0, h.b1, h.b2);
                    const uint32_t ab = a.bin_mode ? hp : (hc == 0 ? j0 : (hc == 1 ? j1 : j2));
                    if (pass == 0) {
                        abs_bin = ab;
                        absorbed_any = true;
                        n_abs++;
                    } else if (f_area) {
                        area_abs = true;  // absorption-side area term
                    }
                    ended = true;
                }
            }
            path_on = !ended;
            next = occl ? 2 : (ended ? 0 : 1);
        }
        if (next == 2) {
            mode = M_OCCL;
        } else if (next == 1 && depth < a.max_depth) {
            ++depth;
            mode = M_PATH;
            tracing = query_begin(q, a.tris, a.n_leaves, o, d, 1e-12f, 1e12f, ntri);
        } else {
            mode = M_END;
        }
        }
        // the wave's events of this iteration (every lane is here)
#pragma unroll
        for (int k = A_SEG; k < A_COUNT; k++)
            wcnt[k] += (uint32_t)__builtin_popcountll(__ballot((ev >> k) & 1u));
    }
    // this lane's last source triangle
    if (acc_t != 0xffffffffu) {
#pragma unroll
        for (int e = 0; e < 9; e++)
            if (sacc[e] != 0.0)
                unsafeAtomicAdd(a.grad + 3 * (size_t)a.idx[3 * acc_t + e / 3] + e % 3, sacc[e]);
    }
    for (int off = 32; off > 0; off >>= 1) n_abs += __shfl_down(n_abs, off, 64);
    wcnt[A_ABS] = n_abs;
    for (int k = 0; k < A_COUNT; k++)
        if (threadIdx.x == 0 && wcnt[k]) atomicAdd(a.stats + k, (unsigned long long)wcnt[k]);
}

}  // namespace

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
    std::vector<float> w(s.nt);
    for (uint32_t t = 0; t < s.nt; t++) {
        float wt = s.tri_area[t] / (float)per;
        w[t] = wt * (kSigma * s.eps[t]);
    }
    const double *x4 = x;  // emission factors per bin, n_hist x nb
    hipStream_t stm = s.stream;
    DBuf<float> d_w, d_area;
    DBuf<double> d_lam, d_x4, d_grad;
    DBuf<unsigned long long> d_st;
    d_w.alloc(s.nt);
    d_area.alloc(s.nt);
    d_lam.alloc((size_t)n_hist * nb);
    d_x4.alloc((size_t)n_hist * nb);
    d_grad.alloc(3 * (size_t)s.nv);
    d_st.alloc(8);
    IRTT_HIP(hipMemcpyAsync(d_w.p, w.data(), s.nt * 4, hipMemcpyHostToDevice, stm));
    IRTT_HIP(hipMemcpyAsync(d_area.p, s.tri_area.data(), s.nt * 4, hipMemcpyHostToDevice, stm));
    IRTT_HIP(hipMemcpyAsync(d_lam.p, lambda, d_lam.bytes(), hipMemcpyHostToDevice, stm));
    IRTT_HIP(hipMemcpyAsync(d_x4.p, x4, d_x4.bytes(), hipMemcpyHostToDevice, stm));
    IRTT_HIP(hipMemsetAsync(d_grad.p, 0, d_grad.bytes(), stm));
    IRTT_HIP(hipMemsetAsync(d_st.p, 0, d_st.bytes(), stm));
    AdjArgs a{};
    a.xyz = s.d_xyz.p;
    a.idx = s.d_idx.p;
    a.rho_d = s.d_rho_d.p;
    a.rho_s = s.d_rho_s.p;
    a.flags = s.d_flags.p;
    a.w = d_w.p;
    a.area = d_area.p;
    a.nodes = s.bvh.nodes.p;
    a.tris = s.bvh.tris.p;
    a.n_leaves = s.bvh.n_leaves;
    a.nt = s.nt;
    a.nb = nb;
    a.bin_mode = s.bin_mode;
    a.per_tri = per;
    a.rays_per_thread = tp.rays_per_thread;
    a.batch_seed = tp.batch_seed;
    a.max_depth = max_depth;
    a.tiny_dist = tp.tiny_dist;
    a.eps = ap.tiny_eps > 0 ? ap.tiny_eps : 1e-2f;
    a.ratio = ap.edge_ratio > 0 ? ap.edge_ratio : 0.05f;
    a.aflags = ap.flags ? ap.flags : (IRTT_ADJ_AREA | IRTT_ADJ_CONTINUOUS | IRTT_ADJ_EDGE);
    a.n_hist = n_hist;
    a.lambda = d_lam.p;
    a.x4 = d_x4.p;
    a.photons = (uint64_t)per * s.nt;
    a.absorber = s.d_absorber.n == a.photons ? s.d_absorber.p : nullptr;
    a.grad = d_grad.p;
    a.stats = d_st.p;
    int dev = 0, ncu = 256;
    IRTT_HIP(hipGetDevice(&dev));
    IRTT_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // the forward kernel's stack choice: 24 entries (more waves per CU) for
    // trees up to 23 levels
    const int S = s.bvh.max_depth + 1 <= 24 ? 24 : stack_size_for(s.bvh.max_depth);
    // runs as long as kAdjRun while the chunks still cover ncu * 64 waves
    const uint64_t waves = (uint64_t)ncu * 64;
    a.run = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(kAdjRun, a.photons / (waves * kTraceBlock)));
    uint64_t want = (a.photons + (uint64_t)kTraceBlock * a.run - 1) / ((uint64_t)kTraceBlock * a.run);
    uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, waves));
    IRTT_HIP(hipEventRecord(s.ev0, stm));
    if (S == 24) k_adjoint<24><<<grid, kTraceBlock, 0, stm>>>(a);
    else if (S == 32) k_adjoint<32><<<grid, kTraceBlock, 0, stm>>>(a);
    else if (S == 64) k_adjoint<64><<<grid, kTraceBlock, 0, stm>>>(a);
    else k_adjoint<128><<<grid, kTraceBlock, 0, stm>>>(a);
    IRTT_HIP(hipGetLastError());
    IRTT_HIP(hipEventRecord(s.ev1, stm));
    std::vector<unsigned long long> hs(8);
    IRTT_HIP(hipMemcpyAsync(grad, d_grad.p, d_grad.bytes(), hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipMemcpyAsync(hs.data(), d_st.p, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                            stm));
    IRTT_HIP(hipStreamSynchronize(stm));
    IRTT_HIP(hipEventElapsedTime(&st.trace_ms, s.ev0, s.ev1));
    st.photons = a.photons;
    st.absorbed = hs[A_ABS];
    st.segments = hs[A_SEG];
    st.continuous = hs[A_CONT];
    st.edge_hits = hs[A_EDGE];
    st.occluded = hs[A_OCCL];
    st.rays = a.absorber ? s.fwd_rays : hs[A_RAYS];
}

}  // namespace irtt
