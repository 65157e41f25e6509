/*
 * irtt_oracle.c -- CPU oracle (test infrastructure only; see irtt_oracle.h).
 *
 * Build with -O2 -ffp-contract=off (no -ffast-math): every FP32 expression is
 * evaluated as written, one IEEE operation at a time, so the HIP kernels that
 * follow the same operation order reproduce it bit for bit.
 */
#include "irtt_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* RNG: assets/shader/utils/glsl/rendering_utils.glsl:361-392               */
/* ------------------------------------------------------------------------ */

/* 16-round TEA seed mix, rendering_utils.glsl:361-374 */
uint32_t or_tea16(uint32_t v0, uint32_t v1)
{
    uint32_t s0 = 0;
    for (int n = 0; n < 16; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

/* Numerical Recipes LCG, rendering_utils.glsl:376-380 */
uint32_t or_lcg(uint32_t *state)
{
    *state = 1664525u * *state + 1013904223u;
    return *state;
}

/* mantissa-mask float in [0,1), rendering_utils.glsl:383-392 */
float or_rnd(uint32_t *state)
{
    uint32_t bits = 0x3f800000u | (0x007fffffu & (or_lcg(state) >> 9));
    float f;
    memcpy(&f, &bits, 4);
    return f - 1.0f;
}

void or_rnd_stream(uint32_t seed, uint32_t n, float *out)
{
    for (uint32_t i = 0; i < n; i++) out[i] = or_rnd(&seed);
}

/* sin/cos of 2*pi*u: v = 4u, q = nearest quadrant, f = v - q in [-1/2,1/2],
 * sin(pi/2 f) and cos(pi/2 f) from their Taylor series (degree 9 / 10,
 * truncation < 2e-9 on the interval), then the quadrant rotation. */
void or_sincos_turns(float u, float *s, float *c)
{
    float v = u * 4.0f;
    int q = (int)(v + 0.5f);
    float f = v - (float)q;
    float f2 = f * f;
    float ps = 1.604411809e-04f;
    ps = ps * f2 + -4.681753926e-03f;
    ps = ps * f2 + 7.969262451e-02f;
    ps = ps * f2 + -6.459640861e-01f;
    ps = ps * f2 + 1.570796371e+00f;
    float sf = ps * f;
    float pc = -2.520204180e-05f;
    pc = pc * f2 + 9.192602593e-04f;
    pc = pc * f2 + -2.086348087e-02f;
    pc = pc * f2 + 2.536695004e-01f;
    pc = pc * f2 + -1.233700514e+00f;
    float cf = pc * f2 + 1.0f;
    switch (q & 3) {
    case 0: *s = sf; *c = cf; break;
    case 1: *s = cf; *c = -sf; break;
    case 2: *s = -sf; *c = -cf; break;
    default: *s = -cf; *c = sf; break;
    }
}

/* ------------------------------------------------------------------------ */
/* small vector helpers (GLSL semantics, operation order fixed)              */
/* ------------------------------------------------------------------------ */
static inline void v3_sub(const float a[3], const float b[3], float r[3])
{
    r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2];
}
static inline void v3_cross(const float a[3], const float b[3], float r[3])
{
    float x = a[1] * b[2] - a[2] * b[1];
    float y = a[2] * b[0] - a[0] * b[2];
    float z = a[0] * b[1] - a[1] * b[0];
    r[0] = x; r[1] = y; r[2] = z;
}
static inline float v3_dot(const float a[3], const float b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
/* normalize(v) = v / length(v); returns the length */
static inline float v3_normalize(const float v[3], float r[3])
{
    float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    r[0] = v[0] / len; r[1] = v[1] / len; r[2] = v[2] / len;
    return len;
}
/* face normal normalize(cross(p1-p0, p2-p0)), thermal_tracing_unified.glsl:101-117 */
static inline float face_normal(const float *p0, const float *p1, const float *p2,
                                float n[3])
{
    float e1[3], e2[3], c[3];
    v3_sub(p1, p0, e1);
    v3_sub(p2, p0, e2);
    v3_cross(e1, e2, c);
    return v3_normalize(c, n);
}
/* p0*b0 + p1*b1 + p2*b2, evaluated left to right */
static inline void bary_point(const float *p0, const float *p1, const float *p2,
                              const float b[3], float r[3])
{
    for (int k = 0; k < 3; k++) r[k] = p0[k] * b[0] + p1[k] * b[1] + p2[k] * b[2];
}
/* findMaxIdx, thermal_tracing_unified.glsl:199-211 (ties go to the later corner) */
static inline int max_corner(const float b[3])
{
    int idx = 0;
    float v = b[0];
    if (b[1] >= v) { idx = 1; v = b[1]; }
    if (b[2] >= v) idx = 2;
    return idx;
}

/* ------------------------------------------------------------------------ */
/* samplers and frames                                                       */
/* ------------------------------------------------------------------------ */

/* ray_tracing_utils.glsl:57-60 */
void or_sample_triangle(float u0, float u1, float b[3])
{
    float su = sqrtf(u0);
    b[0] = 1.0f - su;
    b[1] = su * (1.0f - u1);
    b[2] = su * u1;
}

/* Malley's method via the uniform disk, ray_tracing_utils.glsl:178-182,217-220 */
void or_sample_cosine_hemisphere(float u0, float u1, float d[3])
{
    float r = sqrtf(u0);
    float s, c;
    or_sincos_turns(u1, &s, &c);
    float x = r * c, y = r * s;
    float zz = 1.0f - x * x - y * y;
    d[0] = x; d[1] = y;
    d[2] = sqrtf(zz > 0.0f ? zz : 0.0f);
}

/* getPerpendicularVector + tangentSpaceToWorldSpace, rendering_utils.glsl:212-219,261-272 */
void or_tangent_to_world(const float v[3], const float n[3], float out[3])
{
    float a[3] = {fabsf(n[0]), fabsf(n[1]), fabsf(n[2])};
    uint32_t xm = ((a[0] - a[1]) < 0.0f && (a[0] - a[2]) < 0.0f) ? 1u : 0u;
    uint32_t ym = (a[1] - a[2]) < 0.0f ? (1u ^ xm) : 0u;
    uint32_t zm = 1u ^ (xm | ym);
    float m[3] = {(float)xm, (float)ym, (float)zm};
    float p[3], t[3], bc[3], b[3];
    v3_cross(n, m, p);
    v3_normalize(p, t);
    v3_cross(n, t, bc);
    v3_normalize(bc, b);
    for (int k = 0; k < 3; k++) out[k] = t[k] * v[0] + b[k] * v[1] + n[k] * v[2];
}

/* Ray Tracing Gems ch. 6 integer-ULP offset, ray_tracing_utils.glsl:242-259 */
void or_offset_ray(const float p[3], const float n[3], float out[3])
{
    for (int k = 0; k < 3; k++) {
        int32_t of = (int32_t)(256.0f * n[k]);
        int32_t bits;
        memcpy(&bits, &p[k], 4);
        bits += (p[k] < 0.0f) ? -of : of;
        float pi;
        memcpy(&pi, &bits, 4);
        out[k] = fabsf(p[k]) < (1.0f / 32.0f) ? p[k] + (1.0f / 65536.0f) * n[k] : pi;
    }
}

/* ------------------------------------------------------------------------ */
/* watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), two-sided */
/* ------------------------------------------------------------------------ */
int or_tri_hit(const float o[3], const float d[3], const float p0[3],
               const float p1[3], const float p2[3], float tmin, float tmax,
               float *t_out, float *b1_out, float *b2_out)
{
    float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    int kz = 0;
    float am = ax;
    if (ay > am) { kz = 1; am = ay; }
    if (az > am) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) { int tmp = kx; kx = ky; ky = tmp; }
    float Sx = d[kx] / d[kz];
    float Sy = d[ky] / d[kz];
    float Sz = 1.0f / d[kz];
    float A[3], B[3], C[3];
    v3_sub(p0, o, A);
    v3_sub(p1, o, B);
    v3_sub(p2, o, C);
    float Ax = A[kx] - Sx * A[kz], Ay = A[ky] - Sy * A[kz];
    float Bx = B[kx] - Sx * B[kz], By = B[ky] - Sy * B[kz];
    float Cx = C[kx] - Sx * C[kz], Cy = C[ky] - Sy * C[kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
        return 0;
    float det = U + V + W;
    if (det == 0.0f) return 0;
    float Az = Sz * A[kz], Bz = Sz * B[kz], Cz = Sz * C[kz];
    float T = U * Az + V * Bz + W * Cz;
    float t = T / det;
    if (!(t >= tmin && t <= tmax)) return 0;
    *t_out = t;
    *b1_out = V / det;
    *b2_out = W / det;
    return 1;
}

/* conservative slab test, one fused multiply-add per slab end:
 * t = fma(l, inv, -oi_l), where per ray and axis oi_l is o * inv moved by
 * s = 4u |o * inv| (u = 2^-24) so that the slab end the ray meets first comes
 * out no later, and the one it meets last no earlier, than (l - o) * inv
 * exactly; the remaining relative error of inv and of the fma (2u) is covered
 * by widening the far distance by 1 + 2^-21.  Zero direction components are
 * replaced by 1e-30 before inversion.  Boxes are pruned against the best hit
 * only beyond a 1e-4 relative margin: the box entry distance and the triangle
 * distance round differently, and a triangle that ties the best hit exactly
 * must still be visited. */
static inline int box_hit(const float *lo, const float *hi, const float inv[3],
                          const float oi_lo[3], const float oi_hi[3], float tmin, float tbest,
                          float *tnear)
{
    float tx0 = fmaf(lo[0], inv[0], -oi_lo[0]), tx1 = fmaf(hi[0], inv[0], -oi_hi[0]);
    float ty0 = fmaf(lo[1], inv[1], -oi_lo[1]), ty1 = fmaf(hi[1], inv[1], -oi_hi[1]);
    float tz0 = fmaf(lo[2], inv[2], -oi_lo[2]), tz1 = fmaf(hi[2], inv[2], -oi_hi[2]);
    float n = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    float f = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)) * 1.00000048f;
    *tnear = n;
    return (n <= f) && (n <= tbest * 1.0001f);
}

/* 1 / d and o * (1 / d) shifted by 4u |o / d| towards the near (lo: sign of
 * 1 / d) or far (hi) side of each slab */
static inline void ray_inv(const float o[3], const float d[3], float inv[3], float oi_lo[3],
                           float oi_hi[3])
{
    for (int k = 0; k < 3; k++) {
        inv[k] = 1.0f / (d[k] != 0.0f ? d[k] : 1e-30f);
        float oi = o[k] * inv[k];
        float sh = fabsf(oi) * 2.38418579e-7f;
        oi_lo[k] = inv[k] >= 0.0f ? oi + sh : oi - sh;
        oi_hi[k] = inv[k] >= 0.0f ? oi - sh : oi + sh;
    }
}

/* ------------------------------------------------------------------------ */
/* scene quantities                                                          */
/* ------------------------------------------------------------------------ */
void or_triangle_areas(const or_scene *sc, float *area)
{
    for (uint32_t t = 0; t < sc->nt; t++) {
        const float *p0 = sc->xyz + 3 * sc->idx[3 * t];
        const float *p1 = sc->xyz + 3 * sc->idx[3 * t + 1];
        const float *p2 = sc->xyz + 3 * sc->idx[3 * t + 2];
        double e1[3], e2[3];
        for (int k = 0; k < 3; k++) {
            e1[k] = (double)p1[k] - (double)p0[k];
            e2[k] = (double)p2[k] - (double)p0[k];
        }
        double cx = e1[1] * e2[2] - e1[2] * e2[1];
        double cy = e1[2] * e2[0] - e1[0] * e2[2];
        double cz = e1[0] * e2[1] - e1[1] * e2[0];
        double a = sqrt(cx * cx + cy * cy + cz * cz) * 0.5;
        if (a < 10.0 * FLT_EPSILON) a = 10.0 * FLT_EPSILON;
        area[t] = (float)a;
    }
}

/* ------------------------------------------------------------------------ */
/* LBVH                                                                      */
/* ------------------------------------------------------------------------ */
static uint32_t expand_bits10(uint32_t v)
{
    v &= 0x3ffu;
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

static uint32_t quant10(float s, float lo, float ext)
{
    float f = ext > 0.0f ? (s - lo) / ext : 0.0f;
    float g = f * 1024.0f;
    if (!(g > 0.0f)) g = 0.0f;
    if (g > 1023.0f) g = 1023.0f;
    return (uint32_t)g;
}

typedef struct { uint64_t key; } key_rec;
static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

static inline int delta(const uint64_t *k, int n, int i, int j)
{
    if (j < 0 || j > n - 1) return -1;
    return clz64(k[i] ^ k[j]);
}

int or_bvh_build(const or_scene *sc, or_bvh *bvh)
{
    memset(bvh, 0, sizeof(*bvh));
    uint32_t n = 0;
    for (uint32_t t = 0; t < sc->nt; t++)
        if (sc->flags[t] & OR_FLAG_TRACEABLE) n++;
    bvh->n_leaves = n;
    if (n == 0) return 0;
    float *cs = (float *)malloc(sizeof(float) * 3 * n);
    uint32_t *ids = (uint32_t *)malloc(sizeof(uint32_t) * n);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t m = 0;
    for (uint32_t t = 0; t < sc->nt; t++) {
        if (!(sc->flags[t] & OR_FLAG_TRACEABLE)) continue;
        const float *p0 = sc->xyz + 3 * sc->idx[3 * t];
        const float *p1 = sc->xyz + 3 * sc->idx[3 * t + 1];
        const float *p2 = sc->xyz + 3 * sc->idx[3 * t + 2];
        for (int k = 0; k < 3; k++) {
            float s = p0[k] + p1[k] + p2[k];
            cs[3 * m + k] = s;
            lo[k] = fminf(lo[k], s);
            hi[k] = fmaxf(hi[k], s);
        }
        ids[m++] = t;
    }
    float ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * n);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t code = (expand_bits10(quant10(cs[3 * i], lo[0], ext[0])) << 2) |
                        (expand_bits10(quant10(cs[3 * i + 1], lo[1], ext[1])) << 1) |
                        expand_bits10(quant10(cs[3 * i + 2], lo[2], ext[2]));
        keys[i] = ((uint64_t)code << 32) | ids[i];
    }
    qsort(keys, n, sizeof(uint64_t), cmp_u64);
    bvh->morton = (uint32_t *)malloc(sizeof(uint32_t) * n);
    bvh->tri_of_leaf = (uint32_t *)malloc(sizeof(uint32_t) * n);
    bvh->tris = (uint32_t *)malloc(sizeof(uint32_t) * 12 * n);
    for (uint32_t i = 0; i < n; i++) {
        bvh->morton[i] = (uint32_t)(keys[i] >> 32);
        uint32_t t = (uint32_t)keys[i];
        bvh->tri_of_leaf[i] = t;
        uint32_t *r = bvh->tris + 12 * i;
        for (int c = 0; c < 3; c++) {
            memcpy(r + 4 * c, sc->xyz + 3 * sc->idx[3 * t + c], 12);
            r[4 * c + 3] = 0;
        }
        r[3] = t;
    }
    free(cs);
    free(ids);
    uint32_t ni = n > 1 ? n - 1 : 0;
    bvh->nodes = (uint32_t *)calloc(16 * (size_t)(ni ? ni : 1), sizeof(uint32_t));
    uint32_t *parent_i = (uint32_t *)malloc(sizeof(uint32_t) * (ni ? ni : 1));
    /* Karras 2012, Algorithm 1 (hierarchy from sorted unique 64-bit keys) */
    for (int i = 0; i < (int)ni; i++) {
        int dl = delta(keys, n, i, i - 1), dr = delta(keys, n, i, i + 1);
        int d = (dr - dl) > 0 ? 1 : -1;
        int dmin = delta(keys, n, i, i - d);
        int lmax = 2;
        while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d;
        int dnode = delta(keys, n, i, j);
        int s = 0, div = 2, t;
        do {
            t = (l + div - 1) / div;
            if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
            div *= 2;
        } while (t > 1);
        int g = i + s * d + (d < 0 ? d : 0);
        int mn = i < j ? i : j, mx = i < j ? j : i;
        uint32_t c0 = (mn == g) ? (OR_LEAF_BIT | (uint32_t)g) : (uint32_t)g;
        uint32_t c1 = (mx == g + 1) ? (OR_LEAF_BIT | (uint32_t)(g + 1)) : (uint32_t)(g + 1);
        bvh->nodes[16 * i + 12] = c0;
        bvh->nodes[16 * i + 13] = c1;
        if (!(c0 & OR_LEAF_BIT)) parent_i[c0] = (uint32_t)i;
        if (!(c1 & OR_LEAF_BIT)) parent_i[c1] = (uint32_t)i;
    }
    free(keys);
    /* boxes: post-order by explicit stack; depth by the same walk */
    if (ni > 0) {
        float *nb = (float *)malloc(sizeof(float) * 6 * ni);
        uint32_t *stk = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (ni + 1));
        uint32_t *dep = (uint32_t *)malloc(sizeof(uint32_t) * ni);
        uint8_t *state = (uint8_t *)calloc(ni, 1);
        int sp = 0;
        stk[sp++] = 0;
        dep[0] = 0;
        uint32_t maxd = 0;
        while (sp > 0) {
            uint32_t i = stk[sp - 1];
            if (state[i] == 0) {
                state[i] = 1;
                for (int c = 0; c < 2; c++) {
                    uint32_t ch = bvh->nodes[16 * i + 12 + c];
                    if (ch & OR_LEAF_BIT) {
                        if (dep[i] + 1 > maxd) maxd = dep[i] + 1;
                    } else {
                        dep[ch] = dep[i] + 1;
                        stk[sp++] = ch;
                    }
                }
                continue;
            }
            sp--;
            float *box = nb + 6 * i;
            for (int k = 0; k < 3; k++) { box[k] = INFINITY; box[3 + k] = -INFINITY; }
            for (int c = 0; c < 2; c++) {
                uint32_t ch = bvh->nodes[16 * i + 12 + c];
                float cb[6];
                if (ch & OR_LEAF_BIT) {
                    const float *r = (const float *)(bvh->tris + 12 * (ch & ~OR_LEAF_BIT));
                    for (int k = 0; k < 3; k++) {
                        cb[k] = fminf(fminf(r[k], r[4 + k]), r[8 + k]);
                        cb[3 + k] = fmaxf(fmaxf(r[k], r[4 + k]), r[8 + k]);
                    }
                } else {
                    memcpy(cb, nb + 6 * ch, sizeof(cb));
                }
                memcpy(bvh->nodes + 16 * i + 6 * c, cb, sizeof(cb));
                for (int k = 0; k < 3; k++) {
                    box[k] = fminf(box[k], cb[k]);
                    box[3 + k] = fmaxf(box[3 + k], cb[3 + k]);
                }
            }
        }
        bvh->max_depth = maxd;
        free(nb); free(stk); free(dep); free(state);
    }
    free(parent_i);
    return 0;
}

void or_bvh_free(or_bvh *bvh)
{
    free(bvh->morton); free(bvh->tri_of_leaf); free(bvh->nodes); free(bvh->tris);
    memset(bvh, 0, sizeof(*bvh));
}

typedef struct { int32_t prim; float t, b1, b2; } hit_rec;

static inline void leaf_test(const or_bvh *bvh, uint32_t leaf, const float o[3],
                             const float d[3], float tmin, hit_rec *h)
{
    const float *r = (const float *)(bvh->tris + 12 * leaf);
    int32_t id = (int32_t)bvh->tris[12 * leaf + 3];
    float t, b1, b2;
    if (or_tri_hit(o, d, r, r + 4, r + 8, tmin, h->t, &t, &b1, &b2)) {
        if (t < h->t || h->prim < 0 || (t == h->t && id < h->prim)) {
            h->prim = id; h->t = t; h->b1 = b1; h->b2 = b2;
        }
    }
}

/* Closest-hit traversal, near child first (farther child pushed).  A popped
 * node's children are box-tested against the current best hit, with a 1e-4
 * relative margin so entries at exactly the best distance are kept and ties
 * resolve to the smallest id. */
static void trace_one(const or_bvh *bvh, const float o[3], const float d[3],
                      float tmin, float tmax, hit_rec *h, or_trav_stats *st)
{
    h->prim = -1; h->t = tmax; h->b1 = 0.0f; h->b2 = 0.0f;
    if (bvh->n_leaves == 0) return;
    if (bvh->n_leaves == 1) {
        if (st) st->tri_tests++;
        leaf_test(bvh, 0, o, d, tmin, h);
        if (h->prim < 0) h->t = tmax;
        return;
    }
    float inv[3], oi_lo[3], oi_hi[3];
    ray_inv(o, d, inv, oi_lo, oi_hi);
    uint32_t stk[128];
    float stt[128];
    int sp = 0;
    uint32_t cur = 0;
    for (;;) {
        if (!(cur & OR_LEAF_BIT)) {
            if (st) st->node_visits++;
            const uint32_t *nd = bvh->nodes + 16 * cur;
            const float *bx = (const float *)nd;
            float t0, t1;
            int h0 = box_hit(bx, bx + 3, inv, oi_lo, oi_hi, tmin, h->t, &t0);
            int h1 = box_hit(bx + 6, bx + 9, inv, oi_lo, oi_hi, tmin, h->t, &t1);
            if (h0 && h1) {
                uint32_t nr, fr;
                float ft;
                if (t1 < t0) { nr = nd[13]; fr = nd[12]; ft = t0; }
                else { nr = nd[12]; fr = nd[13]; ft = t1; }
                stk[sp] = fr; stt[sp] = ft; sp++;
                cur = nr;
                continue;
            } else if (h0) { cur = nd[12]; continue; }
            else if (h1) { cur = nd[13]; continue; }
        } else {
            if (st) st->tri_tests++;
            leaf_test(bvh, cur & ~OR_LEAF_BIT, o, d, tmin, h);
        }
        /* pop */
        int found = 0;
        while (sp > 0) {
            sp--;
            /* entries are popped without a distance check, as the device
             * traversal does (its stack keeps node refs only) */
            (void)stt;
            cur = stk[sp]; found = 1; break;
        }
        if (!found) break;
    }
    if (h->prim < 0) h->t = tmax;
}

void or_trace_bvh(const or_bvh *bvh, uint32_t nrays, const float *rays,
                  float tmin, float tmax, int32_t *prim, float *t, float *b1,
                  float *b2, or_trav_stats *st)
{
    for (uint32_t r = 0; r < nrays; r++) {
        hit_rec h;
        trace_one(bvh, rays + 6 * r, rays + 6 * r + 3, tmin, tmax, &h, st);
        prim[r] = h.prim; t[r] = h.t; b1[r] = h.b1; b2[r] = h.b2;
    }
}

void or_trace_brute(const or_scene *sc, uint32_t nrays, const float *rays,
                    float tmin, float tmax, int32_t *prim, float *t_out,
                    float *b1_out, float *b2_out, uint32_t *n_ties)
{
    uint32_t ties = 0;
    for (uint32_t r = 0; r < nrays; r++) {
        const float *o = rays + 6 * r, *d = rays + 6 * r + 3;
        int32_t best = -1;
        float bt = tmax, bb1 = 0.0f, bb2 = 0.0f;
        int tie = 0;
        for (uint32_t tr = 0; tr < sc->nt; tr++) {
            if (!(sc->flags[tr] & OR_FLAG_TRACEABLE)) continue;
            float t, b1, b2;
            if (!or_tri_hit(o, d, sc->xyz + 3 * sc->idx[3 * tr],
                            sc->xyz + 3 * sc->idx[3 * tr + 1],
                            sc->xyz + 3 * sc->idx[3 * tr + 2], tmin, tmax, &t, &b1, &b2))
                continue;
            if (best >= 0 && t == bt) tie = 1;
            if (best < 0 || t < bt) {
                if (best >= 0 && t < bt) tie = 0;
                best = (int32_t)tr; bt = t; bb1 = b1; bb2 = b2;
            }
        }
        ties += (uint32_t)tie;
        prim[r] = best; t_out[r] = bt; b1_out[r] = bb1; b2_out[r] = bb2;
    }
    if (n_ties) *n_ties = ties;
}

/* ------------------------------------------------------------------------ */
/* forward photon transport                                                  */
/* ------------------------------------------------------------------------ */
static const float OR_SIGMA = 5.670374419e-8f; /* thermal_common.hpp:39 */

static uint32_t ceil_log2_u64(uint64_t x)
{
    uint32_t l = 0;
    while (l < 64 && ((uint64_t)1 << l) < x) l++;
    return l;
}

double or_photon_weights(const or_scene *sc, const float *area,
                         const or_params *p, float *w, uint64_t *q)
{
    /* thermal_tracing_unified.glsl:287,309-310:
     *   w = A_t / float(rays * threads); w *= sigma * (1 - rho_d - rho_s) */
    uint32_t per = p->rays_per_thread * p->threads_per_tri;
    float wmax = 0.0f;
    for (uint32_t t = 0; t < sc->nt; t++) {
        float eps = 1.0f - sc->rho_d[t] - sc->rho_s[t];
        float wt = area[t] / (float)per;
        wt = wt * (OR_SIGMA * eps);
        w[t] = wt;
        if (wt > wmax) wmax = wt;
    }
    uint64_t total = (uint64_t)per * (uint64_t)sc->nt;
    int s = 62 - (int)ceil_log2_u64(total ? total : 1);
    if (s > 40) s = 40;
    double Q = wmax > 0.0f ? ldexp((double)wmax, -s) : 1.0;
    for (uint32_t t = 0; t < sc->nt; t++) {
        double r = wmax > 0.0f ? ldexp((double)w[t] / (double)wmax, s) : 0.0;
        q[t] = (uint64_t)rint(r > 0.0 ? r : 0.0);
    }
    return Q;
}

int or_transport(const or_scene *sc, const or_params *p, uint32_t tri_begin,
                 uint32_t tri_end, uint32_t tri_stride, uint64_t *acc, uint32_t *rec_abs,
                 uint32_t *rec_src, uint32_t *rec_tri, uint64_t rec_cap, uint64_t *rec_count,
                 int n_threads, or_stats *st_out)
{
    if (tri_end > sc->nt) tri_end = sc->nt;
    if (tri_begin >= tri_end) { if (rec_count) *rec_count = 0; return 0; }
    float *area = (float *)malloc(sizeof(float) * sc->nt);
    float *w = (float *)malloc(sizeof(float) * sc->nt);
    uint64_t *q = (uint64_t *)malloc(sizeof(uint64_t) * sc->nt);
    or_triangle_areas(sc, area);
    or_photon_weights(sc, area, p, w, q);
    or_bvh bvh;
    or_bvh_build(sc, &bvh);
    const uint32_t nb = p->bin_mode ? sc->nt : sc->nv;
    const uint32_t per = p->rays_per_thread * p->threads_per_tri;
    if (tri_stride == 0) tri_stride = 1;
    const uint32_t ntri_run = (tri_end - tri_begin + tri_stride - 1) / tri_stride;
    const uint64_t nph = (uint64_t)ntri_run * per;
    uint64_t nrec = 0;
    or_stats st;
    memset(&st, 0, sizeof(st));
    uint64_t s_rays = 0, s_abs = 0, s_miss = 0, s_close = 0, s_back = 0, s_deg = 0,
             s_forced = 0, s_nodes = 0, s_tris = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : s_rays, s_abs, s_miss, s_close, s_back, s_deg, s_forced, s_nodes, s_tris)
#endif
    for (uint64_t g = 0; g < nph; g++) {
        uint32_t t = tri_begin + (uint32_t)(g / per) * tri_stride;
        uint32_t k = (uint32_t)(g % per);
        uint32_t x = k / p->rays_per_thread, n = k % p->rays_per_thread;
        /* thermal_tracing_unified.glsl:291 */
        uint32_t seed = or_tea16(sc->nt * x + t, p->batch_seed + n);
        const uint32_t *ti = sc->idx + 3 * t;
        const float *p0 = sc->xyz + 3 * ti[0], *p1 = sc->xyz + 3 * ti[1], *p2 = sc->xyz + 3 * ti[2];
        /* emission, thermal_tracing_unified.glsl:125-161 (rand() draws 3) */
        float u0 = or_rnd(&seed), u1 = or_rnd(&seed);
        (void)or_rnd(&seed);
        float b[3], pos[3], nrm[3], o[3], d[3];
        or_sample_triangle(u0, u1, b);
        bary_point(p0, p1, p2, b, pos);
        float len = face_normal(p0, p1, p2, nrm);
        or_offset_ray(pos, nrm, o);
        if (sc->flags[t] & OR_FLAG_DIFFUSE_EMISSION) {
            float r0 = or_rnd(&seed), r1 = or_rnd(&seed);
            (void)or_rnd(&seed);
            float l[3];
            or_sample_cosine_hemisphere(r0, r1, l);
            or_tangent_to_world(l, nrm, d);
        } else {
            d[0] = nrm[0]; d[1] = nrm[1]; d[2] = nrm[2];
        }
        if (!(len > 0.0f)) { s_deg++; continue; }
        uint32_t src = p->bin_mode ? t : ti[max_corner(b)];
        uint32_t depth = 0;
        for (;;) {
            ++depth;
            if (depth > p->max_depth) break;
            hit_rec h;
            or_trav_stats ts = {0, 0};
            trace_one(&bvh, o, d, 1e-12f, 1e12f, &h, &ts);
            s_rays++;
            s_nodes += ts.node_visits;
            s_tris += ts.tri_tests;
            if (h.prim < 0) { s_miss++; break; }
            if (h.t < p->tiny_dist) { s_close++; break; }
            const uint32_t *hi = sc->idx + 3 * (uint32_t)h.prim;
            const float *q0 = sc->xyz + 3 * hi[0], *q1 = sc->xyz + 3 * hi[1], *q2 = sc->xyz + 3 * hi[2];
            float hn[3];
            face_normal(q0, q1, q2, hn);
            if (v3_dot(hn, d) > -1e-3f) { s_back++; break; }
            float hb[3] = {1.0f - h.b1 - h.b2, h.b1, h.b2};
            /* getNextRay, thermal_tracing_unified.glsl:163-197 */
            float hp[3];
            bary_point(q0, q1, q2, hb, hp);
            or_offset_ray(hp, hn, o);
            float u = or_rnd(&seed);
            float rd = sc->rho_d[h.prim], rs = sc->rho_s[h.prim];
            int absorbed = 0;
            if (u >= rd + rs) {
                absorbed = 1;
            } else if (u >= rd) {
                float in[3];
                v3_normalize(d, in);
                float dn = 2.0f * v3_dot(hn, in);
                for (int kk = 0; kk < 3; kk++) d[kk] = in[kk] - dn * hn[kk];
            } else {
                float r0 = or_rnd(&seed), r1 = or_rnd(&seed);
                (void)or_rnd(&seed);
                float l[3];
                or_sample_cosine_hemisphere(r0, r1, l);
                or_tangent_to_world(l, hn, d);
            }
            if (absorbed || depth == p->max_depth) {
                /* thermal_tracing_unified.glsl:578-583, row = absorber, column = source */
                uint32_t ab = p->bin_mode ? (uint32_t)h.prim : hi[max_corner(hb)];
                if (absorbed) s_abs++; else s_forced++;
                if (acc) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                    acc[(uint64_t)ab * nb + src] += q[t];
                }
                if (rec_abs) {
                    uint64_t slot;
#ifdef _OPENMP
#pragma omp atomic capture
#endif
                    slot = nrec++;
                    if (slot < rec_cap) {
                        rec_abs[slot] = ab;
                        rec_src[slot] = src;
                        if (rec_tri) rec_tri[slot] = t;
                    }
                }
                break;
            }
        }
    }
    st.photons = nph; st.rays = s_rays; st.absorbed = s_abs; st.miss = s_miss;
    st.too_close = s_close; st.backface = s_back; st.degenerate = s_deg;
    st.forced_absorb = s_forced; st.node_visits = s_nodes; st.tri_tests = s_tris;
    if (st_out) *st_out = st;
    if (rec_count) *rec_count = nrec < rec_cap ? nrec : rec_cap;
    or_bvh_free(&bvh);
    free(area); free(w); free(q);
    return 0;
}
