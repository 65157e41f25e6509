/*
 * irtt.h -- C ABI of the MI355X radiative-thermal core (libirtt.so).
 *
 * Plain C: opaque handles, plain pointers and sizes.  Every function returns
 * an int status: 0 = OK, negative = error; irtt_last_error() then returns a
 * thread-local message.  The caller owns every array it passes; the library
 * copies inputs into device memory (HBM) and writes outputs into caller
 * buffers.  A handle may be used from one host thread at a time; different
 * handles may be used concurrently.
 *
 * The reference's seams that these entry points replace (paths relative to the
 * reference repository root):
 *   - src/implementations/thermal_renderer/thermal_api.hpp:10-62 (external C ABI,
 *     compatibility shims declared in thermal_api_compat.h)
 *   - ThermalTransport::compute / computeAdjoint
 *     (src/implementations/thermal_renderer/thermal_transport.hpp:56-57)
 *   - ThermalSolver::residual / jacobian / solve
 *     (src/implementations/thermal_renderer/thermal_solver.hpp:18-20)
 */
#ifndef IRTT_H
#define IRTT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IRTT_OK 0
#define IRTT_ERR_ARG -1
#define IRTT_ERR_HIP -2
#define IRTT_ERR_STATE -3
#define IRTT_ERR_NOMEM -4
#define IRTT_ERR_LIMIT -5
#define IRTT_ERR_NOCONV -6

#define IRTT_BIN_VERTEX 0u   /* reference binning: argmax-barycentric corner */
#define IRTT_BIN_TRIANGLE 1u /* one bin per triangle */

#define IRTT_MAT_DIFFUSE_EMISSION 1u
#define IRTT_MAT_TRACEABLE 2u

typedef struct irtt_scene_s *irtt_scene;

/* Per-triangle surface description.  The reference keeps these per object
 * (ObjectProperties, thermal_api.cpp:25-45; glTF extras, thermal_scene.cpp:76-108);
 * per-triangle is a superset. */
typedef struct irtt_material {
    float diffuse_reflectance;  /* rho_d */
    float specular_reflectance; /* rho_s; emissivity = 1 - rho_d - rho_s */
    float thickness;            /* m */
    float density;              /* kg/m^3 */
    float heat_capacity;        /* J/(kg K) */
    float heat_conductivity;    /* W/(m K) */
    float heat_flux;            /* absorbed external flux, W/m^2 (extension) */
    uint32_t flags;             /* IRTT_MAT_* */
} irtt_material;

typedef struct irtt_scene_info {
    uint32_t n_vertices, n_triangles, n_bins, n_traceable;
    uint32_t bvh_max_depth;
    uint32_t bin_mode;
    uint64_t device_bytes;      /* HBM held by this scene */
} irtt_scene_info;

/* ---------------------------------------------------------------- library */
int irtt_version(void);
const char *irtt_last_error(void);
/* name of device 0 and its compute-unit count */
int irtt_device_info(char *name, size_t name_len, int *cu_count);

/* ------------------------------------------------------------------ scene */
/* Replaces the scene assembly of ThermalScene::load / ThermalData::init,load
 * (thermal_scene.cpp:9-110, thermal_data.cpp:7-173): areas, emissivities,
 * lumped mass, cotangent conduction Laplacian.  xyz: 3*nv floats, tri: 3*nt
 * vertex indices, mat: nt materials.  Uploads geometry to HBM. */
int irtt_scene_create(const float *xyz, uint32_t nv, const uint32_t *tri,
                      uint32_t nt, const irtt_material *mat, uint32_t bin_mode,
                      irtt_scene *out);
int irtt_scene_destroy(irtt_scene h);
int irtt_scene_info_get(irtt_scene h, irtt_scene_info *info);
/* Initial / fixed temperatures per vertex (K); fixed != 0 pins the value
 * (ThermalData::load fixedVarsVector, initialValueVector, thermal_data.cpp:101-102).
 * In triangle bin mode a bin takes the mean of its corners and is fixed when
 * all three corners are. */
int irtt_scene_set_vertex_state(irtt_scene h, const float *kelvin,
                                const uint8_t *fixed);
/* Per-bin state directly (n_bins entries). */
int irtt_scene_set_bin_state(irtt_scene h, const float *kelvin,
                             const uint8_t *fixed);
/* Replace the external absorbed power per bin (n_bins, W), e.g. a sun that
 * moves between time steps (the reference updates its sky emitters instead,
 * ThermalRenderer::updateSkyValues, thermal_renderer.cpp:499-512).  The
 * transport operator is unaffected. */
int irtt_scene_set_source(irtt_scene h, const double *source);
/* Replace vertex positions (same topology) and rebuild areas, mass, Laplacian
 * and the BVH (ThermalRenderer::updateGeometryWithParameters,
 * thermal_renderer.cpp:1327-1337). */
int irtt_scene_update_vertices(irtt_scene h, const float *xyz);
/* Per-bin physics as assembled: area, emissivity*area, lumped mass (n_bins each). */
int irtt_scene_get_bin_data(irtt_scene h, double *area, double *eps_area,
                            double *mass);

/* -------------------------------------------------------------------- BVH */
/* Replaces the BLAS/TLAS build (thermal_transport.cpp:66-139): on-device LBVH
 * over the traceable triangles (30-bit Morton codes, LDS radix sort, Karras
 * hierarchy, bottom-up boxes). */
int irtt_build_bvh(irtt_scene h);
/* Copies the LBVH out: nodes 16 words each (lo0 hi0 lo1 hi1 child0 child1 0 0),
 * triangle records 12 words each (p0 id p1 0 p2 0), Morton codes in leaf order.
 * Any pointer may be NULL. */
int irtt_get_bvh(irtt_scene h, uint32_t *nodes, uint32_t *tris,
                 uint32_t *morton, uint32_t *n_leaves, uint32_t *max_depth);
/* Closest hit for n rays (6 floats each: origin, direction); the hook for the
 * bit-exact traversal test (replaces traceRayEXT + transport_ray.rchit,
 * thermal_tracing_unified.glsl:362, transport_ray.rchit:12-16).  prim = -1 on
 * miss; b1, b2 = barycentric weights of corners 1 and 2. */
int irtt_trace_rays(irtt_scene h, const float *rays, uint32_t n, float tmin,
                    float tmax, int32_t *prim, float *t, float *b1, float *b2);

/* irtt_trace_rays without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The rays and answers
 * are staged through one statically sized device array per device (up to
 * irtt_trace_rays_noalloc_max() = 262144 rays per call, 10 MiB, held until
 * the process ends); more rays fail with IRTT_ERR_LIMIT.  The BVH must have
 * been built (irtt_build_bvh or a transport build): this call does not build
 * it.  Results equal irtt_trace_rays's bit for bit.  Calls that overlap on
 * one device share the array and race: serialise them. */
int irtt_trace_rays_noalloc(irtt_scene h, const float *rays, uint32_t n, float tmin,
                            float tmax, int32_t *prim, float *t, float *b1, float *b2);
uint32_t irtt_trace_rays_noalloc_max(void);

/* -------------------------------------------------------------- transport */
typedef struct irtt_transport_params {
    uint32_t threads_per_tri;   /* ThermalVars_s::threadsPerTriangle (32768) */
    uint32_t rays_per_thread;   /* ThermalVars_s::raysPerThread (8) */
    uint32_t batch_seed;        /* ThermalVars_s::randSeed (1234) */
    uint32_t max_depth;         /* 0 -> 512 (MAX_MAX_DEPTH) */
    float tiny_dist;            /* ThermalVars_s::tinyDist (1e-3) */
    uint32_t layout;            /* IRTT_LAYOUT_*: how the operator is stored */
    uint32_t flags;             /* IRTT_TRANSPORT_* */
    uint32_t reserved;
} irtt_transport_params;

/* Keep the absorbing bin of every photon (4 B per photon, up to 2^31
 * photons) so irtt_adjoint_trace replays each photon once instead of twice
 * (the reference's adjoint shader traces a first pass to find it,
 * thermal_tracing_unified.glsl:292-599).  Costs the forward kernel a few
 * percent; set it when an adjoint trace will follow. */
#define IRTT_TRANSPORT_KEEP_ABSORBERS 1u

/* Keep the paths as well (implies KEEP_ABSORBERS; up to 2^31 photons): the
 * build's photon kernel writes the records irtt_adjoint_trace's path replay
 * would write -- 48 B per accepted segment, 64 B per hit within the default
 * edge ratio (0.05) of an edge -- so a following adjoint trace skips its
 * replay (city block, 2^26 photons: build 14.4 -> 15.9 ms, trace 19.2 ->
 * 10.7 ms, same gradient bits).  A trace with another edge ratio replays; the
 * records that do not fit the memory the build sets aside (at most 30 % of
 * what is free) are replayed too.  No counterpart in the reference, whose
 * adjoint shader traces every path twice (thermal_tracing_unified.glsl:292-599). */
#define IRTT_TRANSPORT_KEEP_PATHS 2u

/* Storage of the operator built by irtt_build_transport.  DENSE is the
 * reference's n_bins x n_bins FP32 matrix (thermal_transport.cpp:612-660).
 * SPARSE keeps only the entries some photon reached (plus the diagonal) in
 * compressed rows and compressed columns: the entries are bit-identical to
 * the dense ones, the memory is 8 B per stored entry instead of 4 n_bins^2
 * (a Monte-Carlo operator is sparse whenever photons per bin are fewer than
 * bins), and it has no n_bins limit (the reference caps dense indexing at
 * 46 340 bins, thermal_transport.cpp:143).  Photon hits are combined in a
 * device hash table with exact integer sums, then sorted by (row, column), so
 * the stored operator is bit-reproducible. */
#define IRTT_LAYOUT_DENSE 0u
#define IRTT_LAYOUT_SPARSE 1u

typedef struct irtt_transport_stats {
    uint64_t photons, rays, absorbed, miss, too_close, backface, degenerate,
        forced_absorb;
    double quantum;             /* W per fixed-point unit */
    float trace_ms;             /* photon kernel, HIP events */
    float finalize_ms;          /* fixed point -> FP32 dense + diagonal */
    float bvh_ms;               /* last LBVH build */
    uint32_t layout;            /* IRTT_LAYOUT_* of the built operator */
    uint64_t node_visits;       /* internal BVH nodes fetched over all rays */
    uint64_t tri_tests;         /* leaf triangles tested over all rays */
    uint64_t nnz;               /* stored entries (diagonal included) */
    /* Energy conservation, FP64 on the device, in W per unit emission factor
     * (x^4 = 1 K^4 in every bin): emitted = sum sigma eps A; absorbed (the
     * operator's off-diagonal sums), escaped (misses) and dropped (too close,
     * backface) from the photons' fixed-point weights.  energy_rel =
     * |emitted - absorbed - escaped - dropped| / emitted: the FP32 rounding of
     * the photon weights sigma eps A / photons (below 3e-7); energy_exact = 1
     * when the integer identity
     * photons * weights = absorbed + escaped + dropped holds exactly (no
     * photon lost). */
    double emitted_w, absorbed_w, escaped_w, dropped_w, energy_rel;
    uint32_t energy_exact, reserved2;
} irtt_transport_stats;

/* Replaces ThermalTransport::compute (thermal_transport.cpp:599-738):
 * photon tracing into a dense n_bins x n_bins operator, row = absorbing bin,
 * column = emitting bin, off-diagonal -sum(photon weights), diagonal
 * + sigma*eps*A.  Accumulation is exact fixed point, so the result is
 * bit-reproducible. Builds the BVH first if needed. */
int irtt_build_transport(irtt_scene h, const irtt_transport_params *p,
                         irtt_transport_stats *st);
int irtt_get_transport(irtt_scene h, float *dense);
/* Raw fixed-point accumulators (n_bins^2, row-major) and the quantum (dense
 * layout only). */
int irtt_get_transport_fixed(irtt_scene h, uint64_t *acc, double *quantum);
/* The stored operator in compressed rows: *nnz receives the entry count;
 * when rowptr, col and val are all non-NULL they receive n_bins + 1 row
 * offsets, then the column and FP32 value of every entry, rows in order and
 * columns ascending within a row.  Either layout (the dense one is
 * compressed on the fly, zeros dropped, the diagonal kept).  No reference
 * counterpart: the reference keeps only the dense host matrix
 * (thermal_transport.cpp:612-660). */
int irtt_get_transport_sparse(irtt_scene h, uint64_t *nnz, uint32_t *rowptr,
                              uint32_t *col, float *val);
/* Column sums of the operator, exact from the fixed-point accumulators (FP64). */
int irtt_get_column_sums(irtt_scene h, double *colsum);
/* One line of the stored operator (either layout) into out (n_bins FP32):
 * column = 1: T[:, bin], where the power emitted by `bin` is absorbed (the
 * reference's "Emission to all" view); column = 0: T[bin, :], where the power
 * absorbed by `bin` comes from ("Absorption from all").  Replaces the
 * transport-buffer reads of simple_ray.rgen:413-420 (visualization modes 0
 * and 1, thermal_gui.cpp:48). */
int irtt_get_transport_line(irtt_scene h, uint32_t bin, uint32_t column, float *out);

/* --------------------------------------------------------- operator forms */
#define IRTT_OP_DENSE 0u      /* FP32 dense operator of irtt_build_transport */
#define IRTT_OP_ACA 1u        /* U V + diag(C) of irtt_compress_transport */
#define IRTT_OP_DENSE_F16 2u  /* off-diagonal FP16 with an FP32 scale per row, FP64
                                 diagonal; Newton residuals keep the FP32 operator */
#define IRTT_OP_SPARSE 3u     /* FP32 compressed rows / columns (IRTT_LAYOUT_SPARSE) */
#define IRTT_OP_HMATRIX 4u    /* hierarchical blocks of irtt_compress_hierarchical (sparse
                                 layout): Krylov products with the low-rank far blocks +
                                 sparse near field, Newton residuals with the stored entries */

/* Operator used by irtt_solve / irtt_residual / irtt_adjoint_solve (the
 * reference's transportCompressed switch, thermal_solver.cpp:126-131).
 * IRTT_OP_DENSE_F16 quantises the dense operator on first use;
 * IRTT_OP_ACA needs irtt_compress_transport.  A new irtt_build_transport
 * selects IRTT_OP_DENSE (dense layout) or IRTT_OP_SPARSE (sparse layout) and
 * drops the other forms; the sparse layout offers no other form. */
int irtt_select_operator(irtt_scene h, uint32_t op);

/* Y = T X (transposed: T^T X) for n_rhs vectors with the selected operator
 * form (X, Y: n_rhs x n_bins, row-major, FP64).  IRTT_OP_DENSE_F16 runs one
 * FP16 MFMA GEMM per 64 vectors (each vector quantised to FP16 with one FP32
 * scale, FP32 accumulation); the other forms one product per vector.  The
 * reference's single product x = T x (solver mode 1, thermal_solver.cpp:141-150)
 * for many vectors at once. */
int irtt_apply_operator(irtt_scene h, uint32_t n_rhs, const double *X, double *Y,
                        uint32_t transposed);

#define IRTT_ACA_FULL 0u    /* simpleAdaptiveCrossApproximation, thermal_transport.cpp:360-386 */
#define IRTT_ACA_PARTIAL 1u   /* completeAdaptiveCrossApproximation (reference default), :388-461 */
#define IRTT_ACA_PLUS 2u      /* adaptiveCrossApproximationPlus, :522-597 */
#define IRTT_STORE_F32 0u
#define IRTT_STORE_F16 1u     /* FP16 with one FP32 scale per stored row */

typedef struct irtt_aca_params {
    uint32_t algorithm;             /* IRTT_ACA_* (ThermalVars_s::compressionAlgorithm = 1) */
    uint32_t rank;                  /* ThermalVars_s::compressionMaxIters (50); 0 -> 50 */
    uint32_t column_sum_correction; /* ThermalVars_s::applyCollumnSumCorrection (0) */
    uint32_t seed;                  /* ACA+ reference-row draws: C runtime rand() state, 0 -> 1 */
    uint32_t precision;             /* IRTT_STORE_*: factors read by the solver */
    uint32_t reserved[3];
} irtt_aca_params;

typedef struct irtt_aca_stats {
    double err_inf, err_frob;         /* ||T_off - U V||, max-abs and Frobenius */
    double err_inf_rel, err_frob_rel; /* relative to ||T_off|| */
    float compress_ms;
    uint32_t rank;
    uint32_t n_corrected;             /* columns raised by the column-sum correction */
    uint32_t reserved;
} irtt_aca_stats;

/* Replaces ThermalTransport::compressTransportMatrix + collumnSumCorrection
 * (thermal_transport.cpp:281-358, :634-656): rank-k cross approximation of
 * the photon part T_off of the operator (the dense operator without its
 * emission diagonal), T ~ U V + diag(C), C = sigma eps A (+ correction).
 * Factors are built in FP64 on the device; selects IRTT_OP_ACA. */
int irtt_compress_transport(irtt_scene h, const irtt_aca_params *p, irtt_aca_stats *st);
/* U: n_bins x rank row-major, V: rank x n_bins row-major (FP64 as built),
 * cdiag: n_bins, piv_rows / piv_cols: rank pivots each.  Any pointer may be
 * NULL. */
int irtt_get_aca(irtt_scene h, double *U, double *V, double *cdiag, int32_t *piv_rows,
                 int32_t *piv_cols);

/* Hierarchical block compression of the sparse-layout operator: bins in
 * Morton order, a cluster tree of halved index ranges, far blocks where
 * min(diam) <= eta dist, each far block cross-approximated (partial-pivot
 * ACA in FP64 with the reference's residual sums and pivot rule, stop at a
 * quarter of tol, rank <= max_rank) and kept in low rank when its exact
 * error |A - U V|_F <= tol |A|_F (and, unless allow_larger, its factors are
 * smaller than its entries); all other entries stay in a sparse near field.
 * Replaces the reference's single global ACA (thermal_transport.cpp:281-597,
 * which stops at a fixed rank: its tolerance test, :439-451, is commented
 * out) for scenes whose operator is too large for one rank-k form.  Selects
 * IRTT_OP_HMATRIX. */
#define IRTT_HBLOCK_ZERO 0u      /* far block without entries */
#define IRTT_HBLOCK_NOCONV 1u    /* ACA reached max_rank above the tolerance: sparse */
#define IRTT_HBLOCK_REJECTED 2u  /* exact error above tol, or factors not smaller: sparse */
#define IRTT_HBLOCK_LOWRANK 3u   /* stored in low rank */

typedef struct irtt_hmatrix_params {
    double tol;             /* relative Frobenius tolerance per far block (0 -> 1e-4) */
    double eta;             /* admissibility min(diam) <= eta dist (0 -> 1) */
    uint32_t max_rank;      /* 1..32 (0 -> 32) */
    uint32_t leaf_size;     /* cluster leaf size, 8..4096 (0 -> 64) */
    uint32_t precision;     /* IRTT_STORE_F32 or IRTT_STORE_F16 (one FP32 scale per row) */
    uint32_t allow_larger;  /* 1: keep converged blocks in low rank even when larger */
    uint32_t keep_factors;  /* 1: keep the FP64 factors for irtt_get_hmatrix_factors */
    uint32_t workspace_mib; /* FP64 work space of the cross approximations, taken in batches
                               of blocks (0: min(8 GiB, a quarter of free HBM), at least
                               1 GiB; all blocks at once with keep_factors) */
} irtt_hmatrix_params;

typedef struct irtt_hmatrix_stats {
    uint32_t n_clusters, n_leaves, n_far, n_near;   /* tree and block counts */
    uint32_t n_lowrank, n_zero, n_noconv, n_reject;  /* far blocks by outcome */
    uint32_t max_rank_used, reserved;
    double mean_rank;                 /* over the low-rank blocks */
    uint64_t nnz, nnz_near, nnz_lowrank;  /* entries: all, near field, inside low-rank blocks */
    uint64_t bytes_sparse, bytes_hmatrix, bytes_lowrank;
    uint64_t build_bytes;             /* peak FP64 work space of the cross approximations */
    double err_frob_rel;              /* |T - H|_F / |T|_F with the FP64 factors */
    double err_frob_rel_stored;       /* ... with the stored (FP16 / FP32) factors */
    double fro_far_rel;               /* |T restricted to far blocks|_F / |T|_F */
    float build_ms, reserved2;
} irtt_hmatrix_stats;

int irtt_compress_hierarchical(irtt_scene h, const irtt_hmatrix_params *p,
                               irtt_hmatrix_stats *st);
/* Block structure: perm (n_bins: Morton position -> bin), *n_far, and per
 * far block 6 words (r0 r1 c0 c1 in Morton positions, rank, IRTT_HBLOCK_*),
 * its relative error and its pivots (64 words: 32 rows then 32 columns,
 * block-local, -1 past the rank; only with keep_factors).  Any pointer may
 * be NULL. */
int irtt_get_hmatrix(irtt_scene h, uint32_t *perm, uint32_t *n_far, uint32_t *blocks,
                     double *err_rel, int32_t *pivots);
/* FP64 factors of far block b as built (keep_factors): U rank x m, V rank x n,
 * rank-major (u_l at U + l m). */
int irtt_get_hmatrix_factors(irtt_scene h, uint32_t b, double *U, double *V);

/* ------------------------------------------------------------------ solve */
#define IRTT_JAC_EXACT 0u      /* J = T diag(4x^3) + L (+ M/dt) */
#define IRTT_JAC_REFERENCE 1u   /* thermal_solver.cpp:109-166: diag(4x^3) T + L + I/dt */

typedef struct irtt_solver_params {
    uint32_t steady;            /* 1 steady state, 0 one implicit-Euler step */
    uint32_t jacobian;          /* IRTT_JAC_* */
    uint32_t max_newton;        /* reference max_iter = 100 */
    uint32_t max_linear;        /* BiCGSTAB iterations per Newton step */
    double dt;                  /* s; reference step_size = 300 */
    double ambient_kelvin;      /* < 0: no ambient exchange (reference) */
    double x_tol;               /* stop when max |dx| < x_tol (K) */
    double f_tol;               /* stop when ||r||^2 < f_tol (reference 1e-6) */
    double lin_rtol;            /* Krylov relative tolerance */
    uint32_t krylov;            /* IRTT_KRYLOV_*: linear solver of the Newton steps */
    uint32_t gmres_restart;     /* GMRES basis size before a restart; 0 -> 30 */
} irtt_solver_params;

/* Krylov method of irtt_solve / irtt_adjoint_solve (the batched solves use
 * BiCGSTAB).  BiCGSTAB is the reference's (Eigen::BiCGSTAB,
 * thermal_solver.hpp:55-56); GMRES(m) with the same right Jacobi
 * preconditioner and stop rule is the alternative (the Jacobian is
 * nonsymmetric, so CG does not apply). */
#define IRTT_KRYLOV_BICGSTAB 0u
#define IRTT_KRYLOV_GMRES 1u

typedef struct irtt_solve_stats {
    uint32_t newton_iters, linear_iters, line_search_halvings, converged;
    double residual_sq;         /* final ||r||^2 (W^2) */
    double max_dx;              /* last Newton update, K */
    float solve_ms;
    uint32_t reserved;
    /* Global power balance at the solution over the free bins, FP64 on the
     * device (irtt_solve only): power_in = external and ambient power
     * absorbed, power_out = power radiated and conducted away plus, for an
     * implicit-Euler step, the rate of heat stored; balance_rel =
     * |power_in - power_out| / power_in. */
    double power_in, power_out, balance_rel;
} irtt_solve_stats;

/* Replaces ThermalSolver::solve (thermal_solver.cpp:119-229) driven by
 * ThermalRenderer::thermalTimestep (thermal_renderer.cpp:542-660): one
 * steady solve or one implicit-Euler step from the current state.  The
 * result becomes the current state. */
int irtt_solve(irtt_scene h, const irtt_solver_params *p, irtt_solve_stats *st);
int irtt_get_temperatures(irtt_scene h, double *kelvin);   /* n_bins */
int irtt_set_temperatures(irtt_scene h, const double *kelvin); /* current state */
/* residual r(x) of the assembled system at x (n_bins), FP64 */
int irtt_residual(irtt_scene h, const irtt_solver_params *p, const double *x,
                  double *r);

/* Forward driver (ThermalRenderer::runForwardSim, thermal_renderer.cpp:958-1026):
 * `steps` implicit-Euler steps of p->dt from the current state (one solve when
 * p->steady), each step's temperatures stored in history (steps x n_bins).
 * The last step becomes the current state. */
int irtt_simulate(irtt_scene h, const irtt_solver_params *p, uint32_t steps, double *history,
                  irtt_solve_stats *st);

/* n_rhs steady solves that share the operator, conduction, fixed bins and
 * ambient, each with its own external absorbed power (source: n_rhs x n_bins,
 * W per bin; NULL = the scene's), in lockstep: every Krylov product is one
 * multi-vector product (an FP16 MFMA GEMM with IRTT_OP_DENSE_F16, whose Newton
 * iteration then finishes on residuals of the FP32 operator).  Each column is
 * irtt_solve's steady solve (exact Jacobian), started from the current state.
 * kelvin: n_rhs x n_bins out.  The scene state is not changed. */
int irtt_solve_batch(irtt_scene h, const irtt_solver_params *p, uint32_t n_rhs,
                     const double *source, double *kelvin, irtt_solve_stats *st);
/* n_rhs adjoint solves J(x_c)^T lambda_c = rhs_c (steady, exact Jacobian) in
 * lockstep; x, rhs, lambda: n_rhs x n_bins. */
int irtt_adjoint_solve_batch(irtt_scene h, const irtt_solver_params *p, uint32_t n_rhs,
                             const double *x, const double *rhs, double *lambda,
                             irtt_solve_stats *st);

/* ---------------------------------------------------------------- adjoint */
/* Replaces ThermalRenderer::adjointTimestep (thermal_renderer.cpp:932-956):
 * solves J(x)^T lambda = rhs with the transposed Jacobian of irtt_solve's
 * system (p->jacobian selects the exact or the reference form; p->steady and
 * p->dt select the steady or implicit-Euler system).  Fixed bins: rhs and
 * lambda are zero there.  x: state to linearise at (n_bins), NULL = current
 * state.  lambda: n_bins out. */
int irtt_adjoint_solve(irtt_scene h, const irtt_solver_params *p, const double *x,
                       const double *rhs, double *lambda, irtt_solve_stats *st);

#define IRTT_ADJ_AREA 1u               /* source-area terms */
#define IRTT_ADJ_CONTINUOUS 2u         /* geometric-term derivatives */
#define IRTT_ADJ_EDGE 4u               /* approximate edge sampling */
#define IRTT_ADJ_REFERENCE_DROPPED 8u  /* dropped photons still contribute with absorber bin 0
                                          (the reference shader's behaviour) */
#define IRTT_ADJ_EDGE_GAIN 16u         /* extension, off by default: the edge term also credits
                                          the occluding side (the photon's absorbing bin) with
                                          the power the moving edge takes from the shadow side
                                          (or from escape); the reference's edge term counts
                                          the shadow side only (thermal_tracing_unified.glsl:
                                          537-567, whose commented-out follow-up names the
                                          missing term) */
#define IRTT_ADJ_HIT_AREA 32u          /* extension, off by default: every hit point is a fixed
                                          barycentric point of its triangle, so a path's weight
                                          also carries w dA_j/dv / A_j of each hit triangle j
                                          (the reference keeps this factor for the emitting
                                          triangle only, thermal_tracing_unified.glsl:327-351,
                                          588-592): the occluder's own gain when its boundary
                                          triangles stretch */
#define IRTT_ADJ_EDGE_PROJECTED 64u    /* extension, off by default: the edge term's irradiance
                                          derivative divided by |da . n_edge| instead of
                                          multiplied (thermal_tracing_unified.glsl:486): the
                                          area an edge displacement sweeps on the occluder's
                                          plane, seen along the ray; equal at normal incidence,
                                          the reference's factor is (da . n_edge)^2 too small
                                          at grazing incidence */
#define IRTT_ADJ_COLLIMATED 128u       /* extension, off by default: on the first segment of a
                                          photon from a collimated (non-diffuse) emitter, such
                                          as the sun quad, the continuous term differentiates
                                          the beam's transfer, the cosine at the receiver
                                          -(N_j . n_i) / |N_j|, instead of the point-to-point
                                          geometric term (cos cos / r^2,
                                          thermal_tracing_unified.glsl:391-448), whose distance
                                          and emitter-cosine factors a parallel beam does not
                                          have */

typedef struct irtt_adjoint_params {
    float edge_ratio;   /* ThermalVars_s::gradEdgeRatio (0.05), thermal_common.hpp:124 */
    float tiny_eps;     /* ThermalVars_s::tinyEps (1e-2), thermal_common.hpp:96 */
    uint32_t flags;     /* IRTT_ADJ_*; 0 = area | continuous | edge */
    uint32_t reserved;
} irtt_adjoint_params;

typedef struct irtt_adjoint_stats {
    uint64_t photons, absorbed, segments, continuous, edge_hits, occluded, rays;
    float trace_ms;
    uint32_t kept_paths; /* 0: every photon replayed; 1: the build's kept paths
                            (IRTT_TRANSPORT_KEEP_PATHS) stood for the replay;
                            2: kept paths up to a photon, the rest replayed */
} irtt_adjoint_stats;

/* Replaces ThermalTransport::computeAdjoint (thermal_transport.cpp:740-832):
 * replays the photons of the last irtt_build_transport (same seeds and
 * counts) and accumulates into grad_x (3 * n_vertices doubles, overwritten)
 *     grad_x = - sum_h lambda_h^T (dT/dv) x4_h,
 * the vertex gradient of the objective through the operator: each photon
 * term is weighted by sum_h lambda_h[absorbing bin] * x4_h[emitting bin]
 * (lambda and x4 are n_hist x n_bins, history-major; x4 is the emission
 * factor x^4 per bin, as the reference uploads it, or any other weights). */
int irtt_adjoint_trace(irtt_scene h, const irtt_adjoint_params *ap, uint32_t n_hist,
                       const double *lambda, const double *x4, double *grad_x,
                       irtt_adjoint_stats *st);

/* Replaces Parametrization::projectGradient (thermal_params.cpp:130-156):
 * dp[param[k]] += dxdp[3k..3k+2] . grad_x[3 vertex[k] ..] for k < nnz, where
 * dxdp holds d(vertex position)/d(parameter) (world space).  dp: n_params,
 * accumulated.  Host-side; needs no device. */
int irtt_project_gradient(const double *grad_x, uint32_t n_vertices, uint32_t nnz,
                          const uint32_t *vertex, const uint32_t *param, const float *dxdp,
                          uint32_t n_params, double *dp);

/* grad_x (3 n_vertices, accumulated) += lambda^T d b / d vertex for the
 * external power b assembled from the materials' heat_flux (flux times
 * triangle area, a third to each corner bin in vertex mode): the geometry
 * derivative of the source term, which the reference's fixed-temperature
 * emitters carry inside T instead.  A source set per bin with
 * irtt_scene_set_source is a fixed power and adds nothing.  Host-side. */
int irtt_source_gradient(irtt_scene h, const double *lambda, double *grad_x);

/* One design gradient at the current (solved) state: lambda from
 * J^T lambda = dL/dx (irtt_adjoint_solve with p), the adjoint trace with the
 * rows (s lambda, x^4) and, with an ambient temperature, (-s Ta^4, lambda)
 * (s = 1 steady, dt for one implicit-Euler step), the source term
 * (irtt_source_gradient with s lambda), then dp = projection through dxdp
 * (irtt_project_gradient; dp overwritten).  grad_x (3 n_vertices) may be
 * NULL; lambda (n_bins) may be NULL.  The reference's gradient step
 * (ThermalRenderer::runAdjointSim + Parametrization::projectGradient,
 * thermal_renderer.cpp:1028-1133, thermal_params.cpp:130-156) for one state. */
int irtt_design_gradient(irtt_scene h, const irtt_solver_params *p, const double *dLdx,
                         const irtt_adjoint_params *ap, uint32_t nnz, const uint32_t *vertex,
                         const uint32_t *param, const float *dxdp, uint32_t n_params,
                         double *dp, double *grad_x, double *lambda, irtt_solve_stats *st);

/* Adjoint over a forward history (ThermalRenderer::runAdjointSim,
 * thermal_renderer.cpp:1028-1133, objective of thermal_objective.cpp:47-57):
 *   L = s * sum_t 0.5 sum_i w_i (x_t,i - target_t,i)^2,  s = 1/steps (transient) or 1,
 * backwards in time  lambda_t = J(x_t)^-T (dt g_t + M lambda_{t+1})  (steady: J^-T g),
 * g_t = w (x_t - target_t), fixed bins zero; lambda_history[t] = s lambda_t.
 * weights: n_bins (objective weight times bin area, as the reference forms it);
 * targets: steps x n_bins, or n_bins shared by every step (target_stride 0).
 * If grad_x is not NULL the adjoint trace runs too (ap NULL: reference
 * defaults) and grad_x (3 n_vertices) receives dL/dvertex (history rows
 * (lambda_t, x_t^4) and, with an ambient, (-Ta^4, sum_t lambda_t); plus the
 * source term of irtt_source_gradient with sum_t lambda_t). */
int irtt_adjoint_history(irtt_scene h, const irtt_solver_params *p, uint32_t steps,
                         const double *history, const double *weights, const double *targets,
                         uint32_t target_stride, double *lambda_history, double *objective,
                         const irtt_adjoint_params *ap, double *grad_x, irtt_solve_stats *st);

/* ---------------------------------------------------------------- serving */
/* Request loop over design variants (no counterpart inside the reference,
 * whose ThermalRenderer runs one scene at a time, thermal_renderer.cpp:958-1133).
 * Each request is one design: scene -> LBVH -> transport -> steady (or one
 * transient) solve -> adjoint solve -> adjoint trace -> dL/dp.  Requests are
 * admitted in arrival order as a slot and their HBM footprint fit (first fit
 * over the queue: a small request may pass a large one waiting for memory),
 * and run concurrently, one HIP stream per slot, so work of different
 * requests overlaps on the device.  Every kernel is deterministic, so a
 * request's results do not depend on what runs beside it. */
typedef struct irtt_server_s *irtt_server;

typedef struct irtt_serve_config {
    uint32_t slots;             /* requests on the device at once (0 -> 4) */
    uint32_t reserved0;
    uint64_t memory_budget;     /* HBM bytes for requests (0 -> 90% of free memory) */
} irtt_serve_config;

/* Inputs are copied at submit; outputs are written when the request ends. */
typedef struct irtt_request {
    const float *xyz;           /* 3 * nv */
    uint32_t nv, nt;
    const uint32_t *tri;        /* 3 * nt */
    const irtt_material *mat;   /* nt */
    uint32_t bin_mode;
    uint32_t operator_kind;     /* IRTT_OP_DENSE or IRTT_OP_DENSE_F16 */
    const float *kelvin0;       /* per vertex (vertex bins) or per triangle; NULL = 0 K */
    const uint8_t *fixed;       /* same indexing; NULL = none fixed */
    irtt_transport_params transport;
    irtt_solver_params solver;
    /* objective L = 0.5 w sum_i A_i (x_i - target_i)^2 (thermal_objective.cpp:34,47-57);
       target NULL: forward solve only */
    const double *target;       /* n_bins */
    double objective_weight;
    /* design parameters: dp[param[k]] += dxdp[3k..3k+2] . dL/dvertex[vertex[k]] */
    uint32_t nnz, n_params;
    const uint32_t *vertex, *param;
    const float *dxdp;
    /* outputs (caller-owned) */
    double *kelvin;             /* n_bins, or NULL */
    double *dp;                 /* n_params, or NULL */
    double *objective;          /* 1, or NULL */
    /* zero-initialised: the defaults */
    irtt_adjoint_params adjoint;  /* terms of the adjoint trace (flags 0: area | continuous |
                                     edge; edge_ratio, tiny_eps 0: 0.05, 1e-2) */
    const double *source;       /* n_bins external power per bin (W), replacing the materials'
                                   heat_flux; NULL: from the materials */
    uint64_t group;             /* nonzero: requests queued with the same group whose
                                   geometry, materials, bin mode, kelvin0, fixed, transport,
                                   solver and operator are identical share one scene and
                                   transport build; their steady forward and adjoint solves
                                   run in lockstep (batched products) and each gets its own
                                   adjoint trace.  Such requests differ in source and target. */
} irtt_request;

#define IRTT_REQ_QUEUED 0
#define IRTT_REQ_RUNNING 1
#define IRTT_REQ_DONE 2        /* negative: failed with that status code */

typedef struct irtt_request_status {
    int32_t state;
    uint32_t n_bins;
    double queued_ms;           /* submit -> admission */
    double run_ms;              /* admission -> done */
    double latency_ms;          /* submit -> done */
    uint64_t footprint_bytes;   /* HBM reserved for the request */
} irtt_request_status;

int irtt_serve_create(const irtt_serve_config *cfg, irtt_server *out);
int irtt_serve_submit(irtt_server sv, const irtt_request *req, uint64_t *ticket);
/* non-blocking status; irtt_serve_wait blocks until the request has ended.
 * A request has ended once its outputs are written; its scene may still be
 * releasing device memory, which stays counted against the budget until then.
 * The status of an ended request stays readable until irtt_serve_destroy. */
int irtt_serve_poll(irtt_server sv, uint64_t ticket, irtt_request_status *st);
int irtt_serve_wait(irtt_server sv, uint64_t ticket, irtt_request_status *st);
/* waits for every submitted request, then releases the server */
int irtt_serve_destroy(irtt_server sv);

#ifdef __cplusplus
}
#endif
#endif
