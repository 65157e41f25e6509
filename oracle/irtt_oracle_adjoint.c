/*
 * irtt_oracle_adjoint.c -- CPU oracle of the adjoint photon trace
 * (test infrastructure only; see irtt_oracle.h).
 *
 * Restates assets/shader/raytracing_thermal/thermal_tracing_unified.glsl:292-599
 * (TRACING_MODE 2): every photon is traced twice with the same seed; the first
 * pass finds the absorbing bin, the second accumulates
 *   (i)   source-area terms  -/+ dw/dv at emission / absorption (:327-351,588-592),
 *   (ii)  continuous geometric terms w * d ln|G| / dv for the six vertices of
 *         each path segment (:391-448),
 *   (iii) the approximate edge term: hits within gradEdgeRatio of an edge
 *         cast one occlusion ray from just outside the edge (:454-567),
 * each weighted by lambda[absorbing bin] * x^4[source bin] (sumOverHistory,
 * :227-244; one history step).  G = -(N_i.d)(N_j.d) / (|N_i||N_j|(r+eps)^4)
 * with unnormalised face normals N and d = p_j - p_i, the quantity whose
 * gradient the reference's generated symGradGeomTerm3D.h evaluates; its
 * logarithmic derivative is written out here term by term.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "irtt_oracle.h"

typedef struct { float x, y, z; } f3;

static inline f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
static inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 scl(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross(f3 a, f3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float len3(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline f3 nrm(f3 a)
{
    float l = len3(a);
    return mk(a.x / l, a.y / l, a.z / l);
}
static inline int finite3(f3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
static inline f3 bary(f3 p0, f3 p1, f3 p2, float b0, float b1, float b2)
{
    return mk(p0.x * b0 + p1.x * b1 + p2.x * b2, p0.y * b0 + p1.y * b1 + p2.y * b2,
              p0.z * b0 + p1.z * b1 + p2.z * b2);
}
static inline int argmax3(float b0, float b1, float b2)
{
    int idx = 0;
    float v = b0;
    if (b1 >= v) { idx = 1; v = b1; }
    if (b2 >= v) idx = 2;
    return idx;
}
static inline int argmin3(float b0, float b1, float b2)
{
    int idx = 0;
    float v = b0;
    if (b1 < v) { idx = 1; v = b1; }
    if (b2 < v) idx = 2;
    return idx;
}

/* gradient of the triangle area 1/2 |(b-a) x (c-a)| w.r.t. a, b, c
 * (the quantity of symGradArea3D.h) */
void or_area_gradient(const float *a, const float *b, const float *c, float *ga, float *gb,
                      float *gc)
{
    f3 A = ld3(a), B = ld3(b), C = ld3(c);
    f3 e1 = sub(B, A), e2 = sub(C, A);
    f3 n = nrm(cross(e1, e2));
    f3 db = scl(cross(e2, n), 0.5f);
    f3 dc = scl(cross(n, e1), 0.5f);
    f3 da = scl(add(db, dc), -1.0f);
    ga[0] = da.x; ga[1] = da.y; ga[2] = da.z;
    gb[0] = db.x; gb[1] = db.y; gb[2] = db.z;
    gc[0] = dc.x; gc[1] = dc.y; gc[2] = dc.z;
}

/* G and d ln|G| / d{a_i,b_i,c_i,a_j,b_j,c_j} (18 floats in out).  beta_* are
 * the barycentric weights of corners 1 and 2.  Returns G. */
float or_geom_term_grad(const float *ai, const float *bi, const float *ci, const float *aj,
                        const float *bj, const float *cj, float bi1, float bi2, float bj1,
                        float bj2, float eps, float *out)
{
    f3 Ai = ld3(ai), Bi = ld3(bi), Ci = ld3(ci), Aj = ld3(aj), Bj = ld3(bj), Cj = ld3(cj);
    float wi0 = 1.0f - bi1 - bi2, wj0 = 1.0f - bj1 - bj2;
    f3 pi = bary(Ai, Bi, Ci, wi0, bi1, bi2);
    f3 pj = bary(Aj, Bj, Cj, wj0, bj1, bj2);
    f3 d = sub(pj, pi);
    float r = len3(d);
    f3 dh = scl(d, 1.0f / r);
    f3 e1i = sub(Bi, Ai), e2i = sub(Ci, Ai), e1j = sub(Bj, Aj), e2j = sub(Cj, Aj);
    f3 Ni = cross(e1i, e2i), Nj = cross(e1j, e2j);
    float li = len3(Ni), lj = len3(Nj);
    f3 ni = scl(Ni, 1.0f / li), nj = scl(Nj, 1.0f / lj);
    float ci_ = dot(Ni, d), cj_ = dot(Nj, d);
    float re = r + eps;
    float re2 = re * re;
    float G = -(ci_ * cj_) / (li * lj * (re2 * re2));
    /* coefficients of d w.r.t. the six vertices: d = pj - pi */
    const float kd[6] = {-wi0, -bi1, -bi2, wj0, bj1, bj2};
    /* d(N_i.d)/dv: normal part on i's vertices, d part on all */
    f3 e2xd_i = cross(e2i, d), dxe1_i = cross(d, e1i);
    f3 e2xd_j = cross(e2j, d), dxe1_j = cross(d, e1j);
    f3 dNi[3] = {scl(add(e2xd_i, dxe1_i), -1.0f), e2xd_i, dxe1_i};
    f3 dNj[3] = {scl(add(e2xd_j, dxe1_j), -1.0f), e2xd_j, dxe1_j};
    /* d|N|/dv */
    f3 ab_i = cross(e2i, ni), ac_i = cross(ni, e1i);
    f3 ab_j = cross(e2j, nj), ac_j = cross(nj, e1j);
    f3 dLi[3] = {scl(add(ab_i, ac_i), -1.0f), ab_i, ac_i};
    f3 dLj[3] = {scl(add(ab_j, ac_j), -1.0f), ab_j, ac_j};
    for (int v = 0; v < 6; v++) {
        f3 g_ci = scl(Ni, kd[v]);
        f3 g_cj = scl(Nj, kd[v]);
        f3 g_li = mk(0, 0, 0), g_lj = mk(0, 0, 0);
        if (v < 3) {
            g_ci = add(g_ci, dNi[v]);
            g_li = dLi[v];
        } else {
            g_cj = add(g_cj, dNj[v - 3]);
            g_lj = dLj[v - 3];
        }
        f3 g_r = scl(dh, kd[v]);
        f3 g = add(add(scl(g_ci, 1.0f / ci_), scl(g_cj, 1.0f / cj_)),
                   scl(add(add(scl(g_li, 1.0f / li), scl(g_lj, 1.0f / lj)),
                           scl(g_r, 4.0f / re)),
                       -1.0f));
        out[3 * v] = g.x; out[3 * v + 1] = g.y; out[3 * v + 2] = g.z;
    }
    return G;
}

/* The collimated beam's transfer to a receiver triangle j from an emitter
 * triangle i whose photons leave along its unit normal n_i:
 * B = -(N_j . n_i) / |N_j| (the cosine at the receiver), and d ln|B| / d{six
 * vertices} (18 floats in out): on i's corners through n_i = N_i / |N_i|, on
 * j's through N_j and |N_j|; the hit point's motion along a parallel beam
 * changes nothing.  Returns B. */
float or_beam_term_grad(const float *ai, const float *bi, const float *ci, const float *aj,
                        const float *bj, const float *cj, float *out)
{
    f3 Ai = ld3(ai), Bi = ld3(bi), Ci = ld3(ci), Aj = ld3(aj), Bj = ld3(bj), Cj = ld3(cj);
    f3 e1i = sub(Bi, Ai), e2i = sub(Ci, Ai), e1j = sub(Bj, Aj), e2j = sub(Cj, Aj);
    f3 Ni = cross(e1i, e2i), Nj = cross(e1j, e2j);
    float li = len3(Ni), lj = len3(Nj);
    f3 ni = scl(Ni, 1.0f / li), nj = scl(Nj, 1.0f / lj);
    float cij = dot(Ni, Nj);     /* N_i . N_j */
    float cn = dot(Nj, ni);      /* N_j . n_i */
    float B = -cn / lj;
    /* d(N_i . N_j)/dv_i with N_j fixed, d(N_j . n_i)/dv_j with n_i fixed */
    f3 e2xa_i = cross(e2i, Nj), axe1_i = cross(Nj, e1i);
    f3 e2xa_j = cross(e2j, ni), axe1_j = cross(ni, e1j);
    f3 dNi[3] = {scl(add(e2xa_i, axe1_i), -1.0f), e2xa_i, axe1_i};
    f3 dNj[3] = {scl(add(e2xa_j, axe1_j), -1.0f), e2xa_j, axe1_j};
    f3 ab_i = cross(e2i, ni), ac_i = cross(ni, e1i);
    f3 ab_j = cross(e2j, nj), ac_j = cross(nj, e1j);
    f3 dLi[3] = {scl(add(ab_i, ac_i), -1.0f), ab_i, ac_i};
    f3 dLj[3] = {scl(add(ab_j, ac_j), -1.0f), ab_j, ac_j};
    for (int v = 0; v < 6; v++) {
        f3 g = v < 3 ? sub(scl(dNi[v], 1.0f / cij), scl(dLi[v], 1.0f / li))
                     : sub(scl(dNj[v - 3], 1.0f / cn), scl(dLj[v - 3], 1.0f / lj));
        out[3 * v] = g.x; out[3 * v + 1] = g.y; out[3 * v + 2] = g.z;
    }
    return B;
}

typedef struct {
    int32_t prim;
    float t, b1, b2;
} hit_t;

static inline hit_t closest(const or_bvh *bvh, f3 o, f3 d)
{
    float ray[6] = {o.x, o.y, o.z, d.x, d.y, d.z};
    hit_t h;
    or_trace_bvh(bvh, 1, ray, 1e-12f, 1e12f, &h.prim, &h.t, &h.b1, &h.b2, NULL);
    return h;
}

static inline void acc3(double *g, uint32_t v, f3 d, double f)
{
    double x = f * (double)d.x, y = f * (double)d.y, z = f * (double)d.z;
#ifdef _OPENMP
#pragma omp atomic
#endif
    g[3 * v] += x;
#ifdef _OPENMP
#pragma omp atomic
#endif
    g[3 * v + 1] += y;
#ifdef _OPENMP
#pragma omp atomic
#endif
    g[3 * v + 2] += z;
}

int or_adjoint(const or_scene *sc, const or_params *p, const or_adjoint_params *ap,
               const double *lambda, const double *x4, double *grad, int n_threads,
               or_adjoint_stats *st_out)
{
    float *area = (float *)malloc(sizeof(float) * sc->nt);
    float *w = (float *)malloc(sizeof(float) * sc->nt);
    uint64_t *q = (uint64_t *)malloc(sizeof(uint64_t) * sc->nt);
    or_triangle_areas(sc, area);
    or_photon_weights(sc, area, p, w, q);
    or_bvh bvh;
    or_bvh_build(sc, &bvh);
    const uint32_t per = p->rays_per_thread * p->threads_per_tri;
    const uint64_t nph = (uint64_t)sc->nt * per;
    const float eps = ap->tiny_eps, ratio = ap->edge_ratio;
    uint64_t s_abs = 0, s_seg = 0, s_cont = 0, s_edge = 0, s_occl = 0, s_rays = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : s_abs, s_seg, s_cont, s_edge, s_occl, s_rays)
#endif
    for (uint64_t g = 0; g < nph; g++) {
        uint32_t t = (uint32_t)(g / per);
        uint32_t k = (uint32_t)(g % per);
        uint32_t xx = k / p->rays_per_thread, nn = k % p->rays_per_thread;
        const uint32_t seed0 = or_tea16(sc->nt * xx + t, p->batch_seed + nn);
        const uint32_t *ti = sc->idx + 3 * t;
        f3 p0 = ld3(sc->xyz + 3 * ti[0]), p1 = ld3(sc->xyz + 3 * ti[1]), p2 = ld3(sc->xyz + 3 * ti[2]);
        uint32_t abs_bin = 0;
        int absorbed_any = 0;
        for (int pass = 0; pass <= 1; pass++) {
            uint32_t seed = seed0;
            float u0 = or_rnd(&seed), u1 = or_rnd(&seed);
            (void)or_rnd(&seed);
            float b[3], o_[3], n_[3], d_[3], pos_[3];
            or_sample_triangle(u0, u1, b);
            f3 pos = bary(p0, p1, p2, b[0], b[1], b[2]);
            f3 c = cross(sub(p1, p0), sub(p2, p0));
            float len = len3(c);
            f3 nr = mk(c.x / len, c.y / len, c.z / len);
            pos_[0] = pos.x; pos_[1] = pos.y; pos_[2] = pos.z;
            n_[0] = nr.x; n_[1] = nr.y; n_[2] = nr.z;
            or_offset_ray(pos_, n_, o_);
            if (sc->flags[t] & OR_FLAG_DIFFUSE_EMISSION) {
                float r0 = or_rnd(&seed), r1 = or_rnd(&seed);
                (void)or_rnd(&seed);
                float l[3];
                or_sample_cosine_hemisphere(r0, r1, l);
                or_tangent_to_world(l, n_, d_);
            } else {
                d_[0] = n_[0]; d_[1] = n_[1]; d_[2] = n_[2];
            }
            if (!(len > 0.0f)) break;
            f3 o = ld3(o_), d = ld3(d_);
            const uint32_t src = p->bin_mode ? t : ti[argmax3(b[0], b[1], b[2])];
            const float wt = w[t];
            const double fsrc = pass ? lambda[src] * x4[src] : 0.0;
            /* (i) emission-side area term, thermal_tracing_unified.glsl:327-351 */
            f3 dw[3];
            if (pass == 1 && (ap->flags & OR_ADJ_AREA)) {
                float ga[3], gb[3], gc[3];
                float pa[3] = {p0.x, p0.y, p0.z}, pb[3] = {p1.x, p1.y, p1.z}, pc[3] = {p2.x, p2.y, p2.z};
                or_area_gradient(pa, pb, pc, ga, gb, gc);
                float s = wt / area[t];
                dw[0] = scl(ld3(ga), s); dw[1] = scl(ld3(gb), s); dw[2] = scl(ld3(gc), s);
                for (int cc = 0; cc < 3; cc++)
                    if (finite3(dw[cc])) acc3(grad, ti[cc], dw[cc], -fsrc);
            }
            if (pass == 1 && !absorbed_any && !(ap->flags & OR_ADJ_REFERENCE_DROPPED)) break;
            const double fpath = pass ? lambda[abs_bin] * x4[src] : 0.0;
            uint32_t prev[3] = {ti[0], ti[1], ti[2]};
            float pb1 = b[1], pb2 = b[2];
            int prev_max = argmax3(b[0], b[1], b[2]);
            uint32_t depth = 0;
            for (;;) {
                ++depth;
                if (depth > p->max_depth) break;
                hit_t h = closest(&bvh, o, d);
                if (pass == 0) s_rays++;
                if (h.prim < 0 || h.t < p->tiny_dist) break;
                const uint32_t *hi = sc->idx + 3 * (uint32_t)h.prim;
                f3 q0 = ld3(sc->xyz + 3 * hi[0]), q1 = ld3(sc->xyz + 3 * hi[1]), q2 = ld3(sc->xyz + 3 * hi[2]);
                f3 hn = nrm(cross(sub(q1, q0), sub(q2, q0)));
                if (dot(hn, d) > -1e-3f) break;
                float hb0 = 1.0f - h.b1 - h.b2;
                if (pass == 1) {
                    s_seg++;
                    /* (ii) continuous geometric term, :391-448 */
                    if (ap->flags & OR_ADJ_CONTINUOUS) {
                        float gg[18];
                        const float *pv[6] = {sc->xyz + 3 * prev[0], sc->xyz + 3 * prev[1],
                                              sc->xyz + 3 * prev[2], sc->xyz + 3 * hi[0],
                                              sc->xyz + 3 * hi[1], sc->xyz + 3 * hi[2]};
                        /* extension (OR_ADJ_COLLIMATED): a collimated first
                         * segment carries the beam's transfer */
                        const int beam = (ap->flags & OR_ADJ_COLLIMATED) && depth == 1 &&
                                         !(sc->flags[t] & OR_FLAG_DIFFUSE_EMISSION);
                        float G = beam ? or_beam_term_grad(pv[0], pv[1], pv[2], pv[3], pv[4],
                                                           pv[5], gg)
                                       : or_geom_term_grad(pv[0], pv[1], pv[2], pv[3], pv[4],
                                                           pv[5], pb1, pb2, h.b1, h.b2, eps, gg);
                        if (fabsf(G) > eps * eps) {
                            s_cont++;
                            uint32_t vid[6] = {prev[0], prev[1], prev[2], hi[0], hi[1], hi[2]};
                            for (int v = 0; v < 6; v++) {
                                f3 dv = scl(ld3(gg + 3 * v), wt);
                                if (finite3(dv)) acc3(grad, vid[v], dv, fpath);
                            }
                        }
                    }
                    /* extension (OR_ADJ_HIT_AREA): the hit point is a fixed
                     * barycentric point of the hit triangle, whose area
                     * element scales the path density there, so the path's
                     * weight gains w * dA_j/dv / A_j on the hit triangle's
                     * corners (the reference keeps this factor for the
                     * emitting triangle only, :327-351,588-592) */
                    if (ap->flags & OR_ADJ_HIT_AREA) {
                        float ga[3], gb[3], gc[3];
                        float pa[3] = {q0.x, q0.y, q0.z}, pb[3] = {q1.x, q1.y, q1.z},
                              pc[3] = {q2.x, q2.y, q2.z};
                        or_area_gradient(pa, pb, pc, ga, gb, gc);
                        const float s = wt / area[h.prim];
                        const f3 dh[3] = {scl(ld3(ga), s), scl(ld3(gb), s), scl(ld3(gc), s)};
                        for (int cc = 0; cc < 3; cc++)
                            if (finite3(dh[cc])) acc3(grad, hi[cc], dh[cc], fpath);
                    }
                    /* (iii) edge term, :454-567 */
                    int opp = argmin3(hb0, h.b1, h.b2);
                    float hb[3] = {hb0, h.b1, h.b2};
                    if ((ap->flags & OR_ADJ_EDGE) && hb[opp] <= ratio) {
                        s_edge++;
                        int near_ = argmax3(hb0, h.b1, h.b2);
                        int other = 0;
                        if (opp == 0 || near_ == 0) {
                            other = 1;
                            if (opp == 1 || near_ == 1) other = 2;
                        }
                        f3 qv[3] = {q0, q1, q2};
                        uint32_t edge_v = hi[near_];
                        f3 edir = sub(qv[near_], qv[other]);
                        f3 oppe = sub(qv[near_], qv[opp]);
                        float elen = len3(edir);
                        edir = scl(edir, 1.0f / elen);
                        f3 n_edg = nrm(cross(edir, hn));
                        float th = dot(oppe, n_edg);
                        if (th < 0.0f) { th = -th; n_edg = scl(n_edg, -1.0f); }
                        float larea = ratio * th * (1.0f - 0.5f * ratio) * elen;
                        f3 da = nrm(cross(edir, d));
                        if (dot(da, n_edg) < 0.0f) da = scl(da, -1.0f);
                        /* the reference scales by |da.n_edg| (:486); moving the
                         * edge by u sweeps (u.da)/|da.n_edg| of the occluder's
                         * plane per unit length, seen along the ray, so
                         * OR_ADJ_EDGE_PROJECTED divides by it instead */
                        const float cpl = fabsf(dot(da, n_edg));
                        f3 dE = (ap->flags & OR_ADJ_EDGE_PROJECTED)
                                    ? scl(da, elen * wt / (larea * cpl))
                                    : scl(da, cpl * elen * wt / larea);
                        if (!finite3(dE)) dE = mk(0, 0, 0);
                        f3 to = bary(q0, q1, q2, hb0, h.b1, h.b2);
                        to = add(to, scl(n_edg, eps + dot(sub(qv[near_], to), n_edg)));
                        to = sub(to, scl(d, eps));
                        hit_t oc = closest(&bvh, to, d);
                        const int valid = oc.prim >= 0 && oc.t >= p->tiny_dist;
                        const uint32_t origin_v = prev[prev_max];
                        int nb_ = 0;
                        uint32_t shadow_v = 0;
                        if (valid) {
                            const uint32_t *ci = sc->idx + 3 * (uint32_t)oc.prim;
                            for (int a = 0; a < 3; a++)
                                for (int bb = 0; bb < 3; bb++)
                                    if (ci[a] == hi[bb]) nb_ = 1;
                            f3 c0 = ld3(sc->xyz + 3 * ci[0]), c1 = ld3(sc->xyz + 3 * ci[1]),
                               c2 = ld3(sc->xyz + 3 * ci[2]);
                            f3 tn = nrm(cross(sub(c1, c0), sub(c2, c0)));
                            float ob0 = 1.0f - oc.b1 - oc.b2;
                            int sm = argmax3(ob0, oc.b1, oc.b2);
                            shadow_v = ci[sm];
                            if (!nb_ && dot(tn, d) < -1e-3f) {
                                s_occl++;
                                uint32_t shadow_bin = p->bin_mode ? (uint32_t)oc.prim : shadow_v;
                                float sa = 1.0f - sc->rho_d[oc.prim] - sc->rho_s[oc.prim];
                                float d1 = h.t, d2 = oc.t, tot = d1 + d2;
                                double fs = lambda[shadow_bin] * x4[src];
                                acc3(grad, shadow_v, scl(dE, sa * (d1 / tot)), fs);
                                acc3(grad, origin_v, scl(dE, sa * (d2 / tot)), fs);
                                acc3(grad, edge_v, scl(dE, -sa), fs);
                            }
                        }
                        /* extension (OR_ADJ_EDGE_GAIN): the occluding side gains
                         * what the shadow side (or escape) loses, credited to
                         * this path's absorbing bin with the opposite sign */
                        if ((ap->flags & OR_ADJ_EDGE_GAIN) && (oc.prim < 0 || (valid && !nb_))) {
                            double fg = lambda[abs_bin] * x4[src];
                            float f_or = valid ? oc.t / (h.t + oc.t) : 1.0f;
                            acc3(grad, edge_v, dE, fg);
                            acc3(grad, origin_v, scl(dE, f_or), -fg);
                            if (valid) acc3(grad, shadow_v, scl(dE, 1.0f - f_or), -fg);
                        }
                    }
                    prev[0] = hi[0]; prev[1] = hi[1]; prev[2] = hi[2];
                    pb1 = h.b1; pb2 = h.b2;
                    prev_max = argmax3(hb0, h.b1, h.b2);
                }
                /* getNextRay, :163-197 */
                f3 hp = bary(q0, q1, q2, hb0, h.b1, h.b2);
                float hp_[3] = {hp.x, hp.y, hp.z}, hn_[3] = {hn.x, hn.y, hn.z}, on[3];
                or_offset_ray(hp_, hn_, on);
                o = ld3(on);
                float u = or_rnd(&seed);
                float rd = sc->rho_d[h.prim], rs = sc->rho_s[h.prim];
                int absorbed = 0;
                if (u >= rd + rs) {
                    absorbed = 1;
                } else if (u >= rd) {
                    f3 in = nrm(d);
                    float dn = 2.0f * dot(hn, in);
                    d = mk(in.x - dn * hn.x, in.y - dn * hn.y, in.z - dn * hn.z);
                } else {
                    float r0 = or_rnd(&seed), r1 = or_rnd(&seed);
                    (void)or_rnd(&seed);
                    float l[3], dd[3];
                    or_sample_cosine_hemisphere(r0, r1, l);
                    or_tangent_to_world(l, hn_, dd);
                    d = ld3(dd);
                }
                if (absorbed || depth == p->max_depth) {
                    uint32_t ab = p->bin_mode ? (uint32_t)h.prim : hi[argmax3(hb0, h.b1, h.b2)];
                    if (pass == 0) {
                        abs_bin = ab;
                        absorbed_any = 1;
                        s_abs++;
                    } else if (ap->flags & OR_ADJ_AREA) {
                        /* absorption-side area term, :588-592 */
                        for (int cc = 0; cc < 3; cc++)
                            if (finite3(dw[cc])) acc3(grad, ti[cc], dw[cc], fpath);
                    }
                    break;
                }
            }
        }
    }
    if (st_out) {
        st_out->photons = nph; st_out->absorbed = s_abs; st_out->segments = s_seg;
        st_out->continuous = s_cont; st_out->edge_hits = s_edge; st_out->occluded = s_occl;
        st_out->rays = s_rays;
    }
    or_bvh_free(&bvh);
    free(area); free(w); free(q);
    return 0;
}
