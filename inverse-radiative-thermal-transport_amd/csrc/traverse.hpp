// traverse.hpp -- closest-hit LBVH traversal shared by the forward and the
// adjoint photon kernels and the ray query entry point.
//
// Replaces traceRayEXT + transport_ray.rchit of the reference
// (assets/shader/raytracing_thermal/thermal_tracing_unified.glsl:362,
// transport_ray.rchit:12-16).  Nodes: 64 B (two child boxes, two child refs);
// leaves: 48 B (three vertices, triangle id).  Per-lane stack in LDS.
#pragma once
#include "device_math.hpp"
#include "irtt_internal.hpp"

// 1: stack entries keep the box entry distance and are skipped on pop when
// they start beyond the best hit (8 B per entry); 0 (default): node refs only
// (4 B per entry).  Measured on the city block (2^26 photons): 0 runs the
// photon kernel in 19.0 ms against 28.1 ms for 1 (20 against 10 waves per CU,
// 22.1 against 21.0 node visits per ray).
#ifndef IRTT_STACK_TNEAR
#define IRTT_STACK_TNEAR 0
#endif

// Lanes that must wait at a leaf before the wave tests their triangles
// (1 .. 64; 64: when every lane of the wave is at a leaf or answered).
// Measured on the city block (2^26 photons, 21-level tree), photon kernel:
// 64: 14.03 ms, 32: 13.57, 16: 13.25, 8: 13.44.
#ifndef IRTT_LEAF_LANES
#define IRTT_LEAF_LANES 16
#endif

namespace irtt {

constexpr int kTraceBlock = 64;  // one wave per block: the LDS stack is per lane
// LDS floats per lane for the entry distances
constexpr int stack_t_entries(int stack) { return IRTT_STACK_TNEAR ? stack : 1; }

struct Hit {
    int32_t prim;
    float t, b1, b2;
};

// Conservative slab test, one fused multiply-add per slab end:
// t = fma(l, inv, -oi_l), where per ray and axis oi_l is o * inv moved by
// s = 4u |o * inv| (u = 2^-24) so that the slab end the ray meets first comes
// out no later, and the one it meets last no earlier, than (l - o) * inv
// exactly (oi_lo for the box's low bound, oi_hi for its high bound; the shift
// covers the rounding of o * inv and of the shift itself).  What remains is
// the relative error of inv and of the fma (2u), covered by widening the far
// distance by 1 + 2^-21.  Pruning against the best hit keeps a 1e-4 relative
// margin: box entry and triangle distances round differently, and exact ties
// must still be visited.
__device__ __forceinline__ bool box_hit(float lx, float ly, float lz, float hx, float hy, float hz,
                                        V3 inv, V3 oi_lo, V3 oi_hi, float tmin, float tbest,
                                        float &tnear)
{
    float tx0 = fmaf(lx, inv.x, -oi_lo.x), tx1 = fmaf(hx, inv.x, -oi_hi.x);
    float ty0 = fmaf(ly, inv.y, -oi_lo.y), ty1 = fmaf(hy, inv.y, -oi_hi.y);
    float tz0 = fmaf(lz, inv.z, -oi_lo.z), tz1 = fmaf(hz, inv.z, -oi_hi.z);
    float n = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    float f = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)) * 1.00000048f;
    tnear = n;
    return (n <= f) && (n <= tbest * 1.0001f);
}

// o * inv shifted by 4u |o * inv| towards the near (lo: sign of inv) or far
// (hi) side of a slab, for box_hit
__device__ __forceinline__ void slab_offsets(float o, float inv, float &lo, float &hi)
{
    const float oi = o * inv;
    const float sh = fabsf(oi) * 2.38418579e-7f;
    lo = inv >= 0.0f ? oi + sh : oi - sh;
    hi = inv >= 0.0f ? oi - sh : oi + sh;
}

__device__ __forceinline__ void leaf_test(const float4 *__restrict__ tris, uint32_t leaf,
                                          const RayPre &rp, V3 o, float tmin, Hit &h)
{
    const float4 *tr =
        reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(tris) + leaf * 48u);
    float4 a = tr[0], b = tr[1], c = tr[2];
    int32_t id = (int32_t)__float_as_uint(a.w);
    float t, b1, b2;
    if (tri_hit(rp, o, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, V3{c.x, c.y, c.z}, tmin, h.t, t, b1,
                b2)) {
        if (t < h.t || h.prim < 0 || (t == h.t && id < h.prim)) {
            h.prim = id;
            h.t = t;
            h.b1 = b1;
            h.b2 = b2;
        }
    }
}

// A closest-hit query (smallest t, ties to the smallest triangle id) as
// resumable state, so that a kernel can run the traversal of one lane in
// phases and interleave other work: query_begin, then query_nodes and
// query_leaf alternately until one returns false, then query_end.  Near child
// first; a dead end pops the next stack entry at once (with entry distances,
// entries that start beyond the best hit are skipped).  stk_n / stk_t: this
// lane's LDS stack, entry k at [k * kTraceBlock]; the stack must exceed the
// tree depth (checked on the host).
struct Query {
    V3 o, inv;           // origin, 1 / direction
    V3 oi_lo, oi_hi;     // o * inv shifted for the low and high slab ends (slab_offsets)
    RayPre rp;
    Hit h;
    uint32_t cur;
    int sp;
};

// false when the query is already answered (tree of zero or one leaf)
__device__ __forceinline__ bool query_begin(Query &q, const float4 *__restrict__ tris,
                                            uint32_t n_leaves, V3 o, V3 d, float tmin, float tmax,
                                            uint32_t &ntri)
{
    q.o = o;
    q.h = Hit{-1, tmax, 0.0f, 0.0f};
    q.cur = 0;
    q.sp = 0;
    if (n_leaves == 0) return false;
    q.rp = ray_pre(d);
    if (n_leaves == 1) {
        ntri++;
        leaf_test(tris, 0, q.rp, o, tmin, q.h);
        return false;
    }
    q.inv = V3{1.0f / (d.x != 0.0f ? d.x : 1e-30f), 1.0f / (d.y != 0.0f ? d.y : 1e-30f),
               1.0f / (d.z != 0.0f ? d.z : 1e-30f)};
    slab_offsets(o.x, q.inv.x, q.oi_lo.x, q.oi_hi.x);
    slab_offsets(o.y, q.inv.y, q.oi_lo.y, q.oi_hi.y);
    slab_offsets(o.z, q.inv.z, q.oi_lo.z, q.oi_hi.z);
    return true;
}

// next stack entry into q.cur; false when the stack is exhausted
__device__ __forceinline__ bool query_pop(Query &q, const uint32_t *stk_n, const float *stk_t)
{
    while (q.sp > 0) {
        q.sp--;
        if (!IRTT_STACK_TNEAR || stk_t[q.sp * kTraceBlock] <= q.h.t * 1.0001f) {
            q.cur = stk_n[q.sp * kTraceBlock];
            return true;
        }
    }
    return false;
}

// one interior node: descend, or pop at a dead end; false when the stack is
// exhausted (answered)
__device__ __forceinline__ bool query_node(Query &q, const float4 *__restrict__ nodes, float tmin,
                                           uint32_t *stk_n, float *stk_t, uint32_t &nvis)
{
    nvis++;
    // 32-bit byte offset from the array base (fewer than 2^26 nodes, checked on the host)
    const float4 *nd =
        reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(nodes) + (q.cur << 6));
    float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
    float t0, t1;
    bool h0 = box_hit(n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, q.inv, q.oi_lo, q.oi_hi, tmin, q.h.t, t0);
    bool h1 = box_hit(n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, q.inv, q.oi_lo, q.oi_hi, tmin, q.h.t, t1);
    uint32_t c0 = __float_as_uint(n3.x), c1 = __float_as_uint(n3.y);
    if (h0 && h1) {
        uint32_t nr, fr;
        float ft;
        if (t1 < t0) {
            nr = c1; fr = c0; ft = t0;
        } else {
            nr = c0; fr = c1; ft = t1;
        }
        stk_n[q.sp * kTraceBlock] = fr;
        if (IRTT_STACK_TNEAR) stk_t[q.sp * kTraceBlock] = ft;
        (void)ft;
        q.sp++;
        q.cur = nr;
    } else if (h0) {
        q.cur = c0;
    } else if (h1) {
        q.cur = c1;
    } else {
        return query_pop(q, stk_n, stk_t);
    }
    return true;
}

// interior nodes until the lane stands at a leaf (true) or has exhausted its
// stack (false: answered)
__device__ __forceinline__ bool query_nodes(Query &q, const float4 *__restrict__ nodes, float tmin,
                                            uint32_t *stk_n, float *stk_t, uint32_t &nvis)
{
    while (!(q.cur & kLeafBit))
        if (!query_node(q, nodes, tmin, stk_n, stk_t, nvis)) return false;
    return true;
}

// the triangle of the leaf under the lane, then the next stack entry; false
// when answered
__device__ __forceinline__ bool query_leaf(Query &q, const float4 *__restrict__ tris, float tmin,
                                           const uint32_t *stk_n, const float *stk_t,
                                           uint32_t &ntri)
{
    ntri++;
    leaf_test(tris, q.cur & ~kLeafBit, q.rp, q.o, tmin, q.h);
    return query_pop(q, stk_n, stk_t);
}

// the answer: prim -1 and t = tmax on a miss
__device__ __forceinline__ Hit query_end(const Query &q, float tmax)
{
    Hit h = q.h;
    if (h.prim < 0) h.t = tmax;
    return h;
}

// Runs the queries the wave's lanes have begun (tracing: this lane's query is
// open) until every one is answered.  The lanes of a wave walk interior nodes
// together; a lane that reaches a leaf waits there until IRTT_LEAF_LANES
// lanes wait (or no lane is left among the interior nodes), and then the
// waiting lanes test their triangles together.  Each ray's own order of
// visits is that of an uninterrupted walk.
__device__ __forceinline__ void query_run(Query &q, bool tracing, const float4 *__restrict__ nodes,
                                          const float4 *__restrict__ tris, float tmin,
                                          uint32_t *stk_n, float *stk_t, uint32_t &nvis,
                                          uint32_t &ntri)
{
    for (;;) {
        for (;;) {
            bool inner = tracing && !(q.cur & kLeafBit);
            if (inner) {
                tracing = query_node(q, nodes, tmin, stk_n, stk_t, nvis);
                inner = tracing && !(q.cur & kLeafBit);
            }
            if (__ballot(inner) == 0) break;
            if (IRTT_LEAF_LANES < kTraceBlock &&
                __builtin_popcountll(__ballot(tracing && !inner)) >= IRTT_LEAF_LANES)
                break;
        }
        if (tracing && (q.cur & kLeafBit)) tracing = query_leaf(q, tris, tmin, stk_n, stk_t, ntri);
        if (__ballot(tracing) == 0) break;
    }
}

// The same walk with the lane's state held in q.cur alone: an interior node
// ((int)cur >= 0), a leaf (leaf bit set), or kQueryDone once the stack is
// exhausted.  The loop tests are then plain compares whose wave masks feed
// the ballots directly, with no per-lane flags carried from one iteration to
// the next (measured on the city block: photon kernel 11.98 -> 11.72 ms).
// Node refs only (no entry distances on the stack).
constexpr uint32_t kQueryDone = 0xffffffffu;  // not a leaf: leaf indices stay below 2^26

__device__ __forceinline__ uint32_t query_pop_ref(Query &q, const uint32_t *stk_n)
{
    if (q.sp > 0) {
        q.sp--;
        return stk_n[q.sp * kTraceBlock];
    }
    return kQueryDone;
}

// one interior node: the next node or leaf, or the next stack entry (or
// kQueryDone) at a dead end
__device__ __forceinline__ void query_node_ref(Query &q, const float4 *__restrict__ nodes,
                                               float tmin, uint32_t *stk_n, uint32_t &nvis)
{
    nvis++;
    const float4 *nd =
        reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(nodes) + (q.cur << 6));
    float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
    float t0, t1;
    bool h0 = box_hit(n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, q.inv, q.oi_lo, q.oi_hi, tmin, q.h.t, t0);
    bool h1 = box_hit(n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, q.inv, q.oi_lo, q.oi_hi, tmin, q.h.t, t1);
    uint32_t c0 = __float_as_uint(n3.x), c1 = __float_as_uint(n3.y);
    if (h0 && h1) {
        const bool swap = t1 < t0;
        stk_n[q.sp * kTraceBlock] = swap ? c0 : c1;
        q.sp++;
        q.cur = swap ? c1 : c0;
    } else if (h0) {
        q.cur = c0;
    } else if (h1) {
        q.cur = c1;
    } else {
        q.cur = query_pop_ref(q, stk_n);
    }
}

// Runs the wave's queries (a lane not tracing holds cur = kQueryDone) until
// every one is answered; leaf tests wait for IRTT_LEAF_LANES lanes as in
// query_run.  Each ray's own order of visits is that of an uninterrupted walk.
__device__ __forceinline__ void query_run_ref(Query &q, const float4 *__restrict__ nodes,
                                              const float4 *__restrict__ tris, float tmin,
                                              uint32_t *stk_n, uint32_t &nvis, uint32_t &ntri)
{
    static_assert(!IRTT_STACK_TNEAR, "query_run_ref keeps node refs only");
    for (;;) {
        for (;;) {
            if ((int)q.cur >= 0) query_node_ref(q, nodes, tmin, stk_n, nvis);
            if (__ballot((int)q.cur >= 0) == 0) break;
            if (IRTT_LEAF_LANES < kTraceBlock &&
                __builtin_popcountll(__ballot((int)q.cur < 0 && q.cur != kQueryDone)) >=
                    IRTT_LEAF_LANES)
                break;
        }
        if ((int)q.cur < 0 && q.cur != kQueryDone) {
            ntri++;
            leaf_test(tris, q.cur & ~kLeafBit, q.rp, q.o, tmin, q.h);
            q.cur = query_pop_ref(q, stk_n);
        }
        if (__ballot(q.cur != kQueryDone) == 0) break;
    }
}

// query_run_ref that returns once `want` of the lanes with `open` set (a ray
// given to this call or an earlier one, its answer not yet taken) hold their
// answer (cur = kQueryDone), or no lane is tracing: the wave then starts new
// rays on those lanes while the others keep their place in the tree.
__device__ __forceinline__ void query_run_until_ref(Query &q, bool open,
                                                    const float4 *__restrict__ nodes,
                                                    const float4 *__restrict__ tris, float tmin,
                                                    uint32_t *stk_n, uint32_t &nvis,
                                                    uint32_t &ntri, uint32_t want)
{
    for (;;) {
        for (;;) {
            if ((int)q.cur >= 0) query_node_ref(q, nodes, tmin, stk_n, nvis);
            if (__ballot((int)q.cur >= 0) == 0) break;
            if (IRTT_LEAF_LANES < kTraceBlock &&
                __builtin_popcountll(__ballot((int)q.cur < 0 && q.cur != kQueryDone)) >=
                    IRTT_LEAF_LANES)
                break;
        }
        if ((int)q.cur < 0 && q.cur != kQueryDone) {
            ntri++;
            leaf_test(tris, q.cur & ~kLeafBit, q.rp, q.o, tmin, q.h);
            q.cur = query_pop_ref(q, stk_n);
        }
        if (__ballot(q.cur != kQueryDone) == 0) break;
        if ((uint32_t)__builtin_popcountll(__ballot(open && q.cur == kQueryDone)) >= want) break;
    }
}

// the whole query in one call
template <int STACK>
__device__ __forceinline__ Hit trace_closest(const float4 *__restrict__ nodes,
                                             const float4 *__restrict__ tris, uint32_t n_leaves,
                                             V3 o, V3 d, float tmin, float tmax, uint32_t *stk_n,
                                             float *stk_t, uint32_t &nvis, uint32_t &ntri)
{
    Query q;
    const bool tracing = query_begin(q, tris, n_leaves, o, d, tmin, tmax, ntri);
    query_run(q, tracing, nodes, tris, tmin, stk_n, stk_t, nvis, ntri);
    return query_end(q, tmax);
}

// stack entries needed for a tree of the given depth
int stack_size_for(uint32_t depth);

}  // namespace irtt
