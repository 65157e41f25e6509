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

namespace irtt {

constexpr int kTraceBlock = 64;  // one wave per block: the LDS stack is per lane
// LDS floats per lane for the entry distances
constexpr int stack_t_entries(int stack) { return IRTT_STACK_TNEAR ? stack : 1; }

struct Hit {
    int32_t prim;
    float t, b1, b2;
};

// conservative slab test; the far distance is widened by 1 + 2^-21 (>= 1 + 2*gamma_3).
// Pruning against the best hit keeps a 1e-4 relative margin: box entry and
// triangle distances round differently, and exact ties must still be visited.
__device__ __forceinline__ bool box_hit(float lx, float ly, float lz, float hx, float hy, float hz,
                                        V3 o, V3 inv, float tmin, float tbest, float &tnear)
{
    float tx0 = (lx - o.x) * inv.x, tx1 = (hx - o.x) * inv.x;
    float ty0 = (ly - o.y) * inv.y, ty1 = (hy - o.y) * inv.y;
    float tz0 = (lz - o.z) * inv.z, tz1 = (hz - o.z) * inv.z;
    float n = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    float f = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)) * 1.00000048f;
    tnear = n;
    return (n <= f) && (n <= tbest * 1.0001f);
}

__device__ __forceinline__ void leaf_test(const float4 *__restrict__ tris, uint32_t leaf,
                                          const RayPre &rp, V3 o, float tmin, Hit &h)
{
    float4 a = tris[3 * leaf], b = tris[3 * leaf + 1], c = tris[3 * leaf + 2];
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

// Closest hit (smallest t, ties to the smallest triangle id), near child
// first; a popped entry is skipped when its box starts beyond the best hit.
// stk_n / stk_t: this lane's LDS stack, entry k at [k * kTraceBlock]; STACK
// must exceed the tree depth (checked on the host).
template <int STACK>
__device__ __forceinline__ Hit trace_closest(const float4 *__restrict__ nodes,
                                             const float4 *__restrict__ tris, uint32_t n_leaves,
                                             V3 o, V3 d, float tmin, float tmax, uint32_t *stk_n,
                                             float *stk_t, uint32_t &nvis, uint32_t &ntri)
{
    Hit h{-1, tmax, 0.0f, 0.0f};
    if (n_leaves == 0) return h;
    RayPre rp = ray_pre(d);
    if (n_leaves == 1) {
        ntri++;
        leaf_test(tris, 0, rp, o, tmin, h);
        if (h.prim < 0) h.t = tmax;
        return h;
    }
    V3 inv{1.0f / (d.x != 0.0f ? d.x : 1e-30f), 1.0f / (d.y != 0.0f ? d.y : 1e-30f),
           1.0f / (d.z != 0.0f ? d.z : 1e-30f)};
    int sp = 0;
    uint32_t cur = 0;
    for (;;) {
        if (!(cur & kLeafBit)) {
            nvis++;
            const float4 *nd = nodes + 4 * (size_t)cur;
            float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
            float t0, t1;
            bool h0 = box_hit(n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, o, inv, tmin, h.t, t0);
            bool h1 = box_hit(n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, o, inv, tmin, h.t, t1);
            uint32_t c0 = __float_as_uint(n3.x), c1 = __float_as_uint(n3.y);
            if (h0 && h1) {
                uint32_t nr, fr;
                float ft;
                if (t1 < t0) {
                    nr = c1; fr = c0; ft = t0;
                } else {
                    nr = c0; fr = c1; ft = t1;
                }
                stk_n[sp * kTraceBlock] = fr;
                if (IRTT_STACK_TNEAR) stk_t[sp * kTraceBlock] = ft;
                (void)ft;
                sp++;
                cur = nr;
                continue;
            } else if (h0) {
                cur = c0;
                continue;
            } else if (h1) {
                cur = c1;
                continue;
            }
        } else {
            ntri++;
            leaf_test(tris, cur & ~kLeafBit, rp, o, tmin, h);
        }
        bool found = false;
        while (sp > 0) {
            sp--;
            if (!IRTT_STACK_TNEAR || stk_t[sp * kTraceBlock] <= h.t * 1.0001f) {
                cur = stk_n[sp * kTraceBlock];
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    if (h.prim < 0) h.t = tmax;
    return h;
}

// stack entries needed for a tree of the given depth
int stack_size_for(uint32_t depth);

}  // namespace irtt
