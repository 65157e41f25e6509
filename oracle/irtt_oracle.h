/*
 * irtt_oracle.h -- CPU oracle for the radiative-transport hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline), never as a product code path.
 *
 * A plain-C restatement of the reference's photon tracer and the pieces under
 * it.  Each function names the reference file:line it follows (paths are
 * relative to the reference repository root).  FP32 arithmetic is written out
 * operation by operation (compile with -ffp-contract=off) so that the HIP
 * kernels, which follow the same operation order, can be compared with it
 * bit for bit.
 */
#ifndef IRTT_ORACLE_H
#define IRTT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers: assets/shader/utils/glsl/rendering_utils.glsl:361-392 ---- */
uint32_t or_tea16(uint32_t v0, uint32_t v1);
uint32_t or_lcg(uint32_t *state);
float or_rnd(uint32_t *state);
/* fills out[0..n) with successive or_rnd draws starting from seed */
void or_rnd_stream(uint32_t seed, uint32_t n, float *out);

/* sin(2*pi*u), cos(2*pi*u) for u in [0,1) by quadrant reduction + Taylor
 * polynomials (the reference calls GLSL sin/cos,
 * ray_tracing_utils.glsl:178-182; those are not bit-reproducible). */
void or_sincos_turns(float u, float *s, float *c);

/* ---- samplers / frames: ray_tracing_utils.glsl:57-60,178-182,217-220,242-259,
 *      rendering_utils.glsl:212-219,261-272 ---- */
void or_sample_triangle(float u0, float u1, float b[3]);
void or_sample_cosine_hemisphere(float u0, float u1, float d[3]);
void or_tangent_to_world(const float v[3], const float n[3], float out[3]);
void or_offset_ray(const float p[3], const float n[3], float out[3]);

/* ---- ray / triangle (watertight, Woop-Benthin-Wald 2013) ---- */
int or_tri_hit(const float o[3], const float d[3], const float p0[3],
               const float p1[3], const float p2[3], float tmin, float tmax,
               float *t, float *b1, float *b2);

/* ---- scene ---- */
typedef struct {
    uint32_t nv, nt;
    const float *xyz;       /* 3*nv world-space positions */
    const uint32_t *idx;    /* 3*nt vertex indices */
    const float *rho_d;     /* nt diffuse reflectance */
    const float *rho_s;     /* nt specular reflectance */
    const uint8_t *flags;   /* nt: bit0 diffuse emission, bit1 traceable */
} or_scene;

#define OR_FLAG_DIFFUSE_EMISSION 1u
#define OR_FLAG_TRACEABLE 2u

/* triangle areas, FP64 cross product, clamp >= 10*FLT_EPSILON, cast to FP32
 * (thermal_data.cpp:56-81) */
void or_triangle_areas(const or_scene *sc, float *area);

/* ---- LBVH (Karras 2012) over the traceable triangles ---- */
/* node record: 16 x 32-bit words = lo0[3] hi0[3] lo1[3] hi1[3] child0 child1 0 0.
 * child ref: bit31 set = leaf, low bits = leaf slot; leaf slot j holds
 * triangle record j.  triangle record: 12 words = p0[3] id p1[3] 0 p2[3] 0. */
#define OR_LEAF_BIT 0x80000000u
typedef struct {
    uint32_t n_leaves;      /* traceable triangles */
    uint32_t *morton;       /* n_leaves, in leaf (sorted) order */
    uint32_t *tri_of_leaf;  /* n_leaves */
    uint32_t *nodes;        /* 16*(n_leaves-1) words */
    uint32_t *tris;         /* 12*n_leaves words */
    uint32_t max_depth;     /* edges from root to deepest leaf */
} or_bvh;

int or_bvh_build(const or_scene *sc, or_bvh *bvh); /* allocates; 0 on success */
void or_bvh_free(or_bvh *bvh);

typedef struct {
    uint64_t node_visits;   /* internal nodes fetched */
    uint64_t tri_tests;     /* leaf triangles tested */
} or_trav_stats;

/* closest hit: smallest t, ties to the smallest triangle id.  rays: 6 floats
 * (origin, direction) per ray.  prim = -1 on miss. */
void or_trace_bvh(const or_bvh *bvh, uint32_t nrays, const float *rays,
                  float tmin, float tmax, int32_t *prim, float *t, float *b1,
                  float *b2, or_trav_stats *st);
void or_trace_brute(const or_scene *sc, uint32_t nrays, const float *rays,
                    float tmin, float tmax, int32_t *prim, float *t,
                    float *b1, float *b2, uint32_t *n_ties);

/* ---- forward photon transport (thermal_tracing_unified.glsl:247-601) ---- */
typedef struct {
    uint32_t threads_per_tri;   /* launch x extent (thermal_transport.cpp:620) */
    uint32_t rays_per_thread;   /* ubo.rayCount */
    uint32_t batch_seed;        /* ubo.batchSeed */
    uint32_t max_depth;         /* <= 512 */
    uint32_t bin_mode;          /* 0: vertex bins (reference), 1: triangle bins */
    float tiny_dist;            /* ubo.tinyDist, 1e-3 */
} or_params;

typedef struct {
    uint64_t photons, rays, absorbed, miss, too_close, backface, degenerate,
        forced_absorb;
    uint64_t node_visits, tri_tests;
} or_stats;

/* Per-triangle photon weight w_t (FP32, as the shader computes it) and its
 * fixed-point form q_t = rint(w_t / Q) with Q = max_t w_t * 2^-s,
 * s = min(40, 62 - ceil(log2(total photons))).  Returns Q. */
double or_photon_weights(const or_scene *sc, const float *area,
                         const or_params *p, float *w, uint64_t *q);

/* Traces all photons of source triangles tri_begin, tri_begin + tri_stride,
 * ... below tri_end (tri_stride 0 means 1).
 * acc (may be NULL): dense nb*nb uint64 row-major, acc[abs*nb + src] += q_t.
 * rec_abs/rec_src (may be NULL): one record per absorbed photon, up to rec_cap
 * (records beyond the cap are dropped from the list, not from acc), in no
 * particular order; rec_tri (may be NULL) receives each record's source
 * triangle, so that sum of q[rec_tri] per (rec_abs, rec_src) is acc.
 * n_threads: OpenMP threads (0 = default). */
int or_transport(const or_scene *sc, const or_params *p, uint32_t tri_begin,
                 uint32_t tri_end, uint32_t tri_stride, uint64_t *acc, uint32_t *rec_abs,
                 uint32_t *rec_src, uint32_t *rec_tri, uint64_t rec_cap, uint64_t *rec_count,
                 int n_threads, or_stats *st);

/* ---- adjoint photon trace (thermal_tracing_unified.glsl:292-599), irtt_oracle_adjoint.c ---- */
#define OR_ADJ_AREA 1u
#define OR_ADJ_CONTINUOUS 2u
#define OR_ADJ_EDGE 4u
#define OR_ADJ_REFERENCE_DROPPED 8u /* dropped photons keep absorber bin 0, as the shader does */
#define OR_ADJ_EDGE_GAIN 16u        /* extension: occluding side of the edge term */
#define OR_ADJ_HIT_AREA 32u         /* extension: area Jacobian of every hit triangle */
#define OR_ADJ_EDGE_PROJECTED 64u   /* extension: edge displacement projected along the ray */
#define OR_ADJ_COLLIMATED 128u      /* extension: beam transfer on a collimated first segment */

typedef struct {
    float edge_ratio;   /* ubo.gradEdgeRatio, 0.05 */
    float tiny_eps;     /* ubo.tinyEps, 1e-2 */
    uint32_t flags;     /* OR_ADJ_* */
} or_adjoint_params;

typedef struct {
    uint64_t photons, absorbed, segments, continuous, edge_hits, occluded, rays;
} or_adjoint_stats;

/* dA/da, dA/db, dA/dc of A = |(b-a) x (c-a)| / 2 */
void or_area_gradient(const float *a, const float *b, const float *c, float *ga, float *gb,
                      float *gc);
/* G = -(N_i.d)(N_j.d) / (|N_i||N_j|(r+eps)^4), d = p_j - p_i, and
 * d ln|G| / d{a_i,b_i,c_i,a_j,b_j,c_j} (18 floats).  Returns G. */
float or_geom_term_grad(const float *ai, const float *bi, const float *ci, const float *aj,
                        const float *bj, const float *cj, float bi1, float bi2, float bj1,
                        float bj2, float eps, float *out);
/* grad (3*nv doubles, accumulated) += sum over photons of the adjoint terms,
 * weighted by lambda[bin] * x4[bin] (one history step). */
int or_adjoint(const or_scene *sc, const or_params *p, const or_adjoint_params *ap,
               const double *lambda, const double *x4, double *grad, int n_threads,
               or_adjoint_stats *st);

#ifdef __cplusplus
}
#endif
#endif
