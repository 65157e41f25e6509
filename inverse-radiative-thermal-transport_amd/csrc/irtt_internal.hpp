// irtt_internal.hpp -- scene state shared by the host entry points and the
// kernel launchers of libirtt.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/irtt.h"

namespace irtt {

// Error carrying an irtt status code; entry points translate it into the
// return value and the thread-local message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define IRTT_HIP(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::irtt::Error(IRTT_ERR_HIP, std::string(#call) + ": " +                \
                                                  hipGetErrorString(e_));                \
    } while (0)

constexpr float kSigma = 5.670374419e-8f;      // thermal_common.hpp:39
constexpr double kSigmaD = 5.670374419e-8;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kMaxLeaves = 1u << 26;  // traceable triangles per scene
constexpr uint32_t kMaxDepthDefault = 512;     // MAX_MAX_DEPTH, thermal_tracing_unified.glsl:7

// Stream on which the calling thread's device buffers are allocated and freed
// (stream-ordered hipMallocAsync / hipFreeAsync), set by every entry point
// that takes a scene to that scene's stream: scenes served concurrently from
// different threads allocate and free without synchronising the device.
// nullptr: plain hipMalloc / hipFree.
extern thread_local hipStream_t t_alloc_stream;

// Device memory through a process-wide cache (api.cpp): a freed block is kept
// with an event recorded on the freeing stream, and a later allocation of a
// similar size takes it back (its stream waits for that event on the device,
// never on the host).  The stream-ordered pool's own allocate / free calls
// cost milliseconds each at these sizes and stall concurrent scenes; the
// cache makes the steady state of a request loop or a design loop free of
// them.  Throws Error(IRTT_ERR_NOMEM) when the device is out of memory.
void *dev_alloc(size_t bytes, hipStream_t stream);
void dev_free(void *p, size_t bytes, hipStream_t stream);
// bytes held by the cache for reuse (free for this process's allocations)
size_t dev_cache_bytes();
// device memory this process can allocate: free on the device, plus what the
// stream-ordered pool and the cache hold unused
size_t dev_available();
// cache hits, fresh allocations, their bytes, nanoseconds spent allocating,
// blocks returned by trims and their bytes (6 counters since process start)
void dev_counters(uint64_t *out);
// streams are reused across scenes the same way (creating one can take
// tens of milliseconds)
hipStream_t stream_take();
void stream_give(hipStream_t st);

// Device buffer with RAII release.  alloc() keeps the allocation when the new
// count fits in what it holds (grow-only): repeated builds and traces of one
// scene whose sizes drift (a design loop) reuse their buffers instead of
// returning tens of GB to the pool and mapping them again.
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;              // elements in use
    size_t cap = 0;            // elements allocated
    hipStream_t st = nullptr;  // stream the buffer was allocated on (nullptr: hipMalloc)
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p) dev_free(p, cap * sizeof(T), st);
        p = nullptr;
        n = 0;
        cap = 0;
        st = nullptr;
    }
    void alloc(size_t count)
    {
        if (p && count <= cap) {
            n = count;
            return;
        }
        release();
        if (count == 0) return;
        p = reinterpret_cast<T *>(dev_alloc(count * sizeof(T), t_alloc_stream));
        st = t_alloc_stream;
        n = count;
        cap = count;
    }
    size_t bytes() const { return n * sizeof(T); }
    size_t held_bytes() const { return cap * sizeof(T); }
    // alloc(count), returning the pointer
    T *alloc_get(size_t count)
    {
        alloc(count);
        return p;
    }
    // alloc(v.size()) and an asynchronous copy of v on st (v must outlive
    // the copy: callers synchronise st before v goes out of scope)
    void upload(const std::vector<T> &v, hipStream_t st)
    {
        alloc(v.size());
        if (!v.empty())
            IRTT_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    }
};

// Sparse matrix in CSR, FP64 (conduction Laplacian).
struct Csr {
    uint32_t n = 0;
    std::vector<uint32_t> rowptr, col;
    std::vector<double> val;
};

struct Bvh {
    uint32_t n_leaves = 0, max_depth = 0;
    bool valid = false;
    DBuf<float4> nodes;       // 4 float4 per internal node
    DBuf<float4> tris;        // 3 float4 per leaf
    DBuf<uint32_t> morton;    // leaf order
    float build_ms = 0.f;
};

// Hash-table slot of the sparse build: key (absorbing bin << key_bits |
// emitting bin), exact u64 sum of the photons' fixed-point weights.
struct alignas(16) TabSlot {
    unsigned long long key, val;
};

// Where the photon kernel bins an absorbed photon: the dense accumulators,
// or (tab set) the 2^lg-slot hash table of the sparse layout.
// Dense accumulators are tracked in row chunks of 2^kTouchShift entries
// (128 B): the photon kernel marks every chunk it adds to, so the next build
// clears only marked chunks and the finalize pass reads only those.
constexpr uint32_t kTouchShift = 4;

struct PhotonSink {
    unsigned long long *acc = nullptr;
    uint8_t *touch = nullptr;  // one byte per accumulator chunk (dense layout)
    TabSlot *tab = nullptr;
    uint32_t lg = 0, key_bits = 0;
    uint32_t *overflow = nullptr;
};

// Operator in compressed rows and compressed columns (sparse.hip).  Every row
// and column holds its diagonal entry.
struct Sparse {
    bool valid = false;
    uint64_t nnz = 0;
    DBuf<uint32_t> rowptr, col;       // n_bins + 1 offsets, column of each entry
    DBuf<float> val;                  // FP32 entries, rows in order, columns ascending
    DBuf<uint32_t> colptr, row;       // the same entries by column, rows ascending
    DBuf<float> cval;
    DBuf<double> d_xt, d_yt;          // vector-interleaved operands of a batched product
    DBuf<TabSlot> tab;                // hash table of a build (kept for the next)
    DBuf<uint32_t> flag;
    // table -> rows / columns work space (kept between builds)
    DBuf<uint32_t> ccount;
    DBuf<unsigned long long> coff, ka, kb, va, vb;
    DBuf<uint8_t> sort_tmp;
    // drops the operator, keeps the buffers for the next build
    void invalidate()
    {
        valid = false;
        nnz = 0;
    }
    void reset()
    {
        valid = false;
        nnz = 0;
        for (DBuf<unsigned long long> *b : {&coff, &ka, &kb, &va, &vb}) b->release();
        ccount.release();
        sort_tmp.release();
        for (DBuf<uint32_t> *b : {&rowptr, &col, &colptr, &row, &flag}) b->release();
        val.release();
        cval.release();
        d_xt.release();
        d_yt.release();
        tab.release();
    }
};

// Adjoint trace work space (adjoint.hip), kept between traces of a scene:
// path records (uint4 words: 3 per segment record, 4 per edge record), their
// sort keys and orders, per-call inputs and the exact gradient sums.
struct AdjWork {
    DBuf<float> w, area;
    DBuf<double> lam, x4;
    DBuf<unsigned long long> grad, st, ctl;
    DBuf<uint32_t> abs_own, fills, keys, keys2, order, order2, n_big;
    DBuf<uint32_t> live, live_counts;  // photons with an absorbing bin; per-block counts,
                                       // offsets and their number
    DBuf<uint8_t> scan_tmp;
    DBuf<uint4> segs, edges;
    DBuf<uint8_t> sort_tmp;
    // bytes held by the record buffers (the part sized against free HBM)
    size_t record_bytes() const
    {
        return segs.held_bytes() + edges.held_bytes() + fills.held_bytes() +
               keys.held_bytes() + keys2.held_bytes() + order.held_bytes() +
               order2.held_bytes() + sort_tmp.held_bytes();
    }
};

// Path records kept by the last transport build (IRTT_TRANSPORT_KEEP_PATHS):
// segment records (3 uint4 words each) and raw edge records (4 words) of
// every photon in wave-owned chunks, for the adjoint trace instead of its
// path replay.  Valid from the build's end until the next build.
struct PathStore {
    bool valid = false;
    DBuf<uint4> segs, edges;
    DBuf<uint32_t> fills;             // per chunk: seg_chunks segment chunks, then edge chunks
    DBuf<unsigned long long> ctl;     // chunks taken (segments, edges), first dropped photon
    uint64_t seg_chunks = 0, edge_chunks = 0;  // capacity
    uint64_t sc = 0, ec = 0;          // chunks written
    uint64_t first_dropped = ~0ull;   // smallest photon with a record that did not fit
    uint64_t photons = 0;
    void release()
    {
        valid = false;
        segs.release();
        edges.release();
        fills.release();
        ctl.release();
        seg_chunks = edge_chunks = sc = ec = 0;
    }
    size_t bytes() const { return segs.bytes() + edges.bytes() + fills.bytes() + ctl.bytes(); }
};

constexpr uint32_t kAcaMaxRank = 256;

// ACA-compressed operator U V + diag(C) (aca.hip)
struct Aca {
    bool valid = false;
    uint32_t k = 0, precision = 0;
    DBuf<double> d_ut, d_v;           // k x ld: columns of U, rows of V (FP64)
    DBuf<float> d_uf, d_vf;           // FP32 copies read by the solver
    DBuf<uint16_t> d_uh, d_vh;        // FP16 copies (bit patterns) ...
    DBuf<float> d_us, d_vs;           // ... and their per-row scales
    DBuf<float> d_od;                 // off-diagonal operator's diagonal, FP32
    DBuf<double> d_u, d_w, d_part, d_mag, d_t;
    DBuf<int32_t> d_piv;
    DBuf<uint8_t> d_excl;
    std::vector<int32_t> prow, pcol;
    std::vector<double> cdiag;        // C: sigma eps A (+ column-sum correction)
    std::vector<double> opdiag;       // diag(U V) + C
    double max_uv = 0.0;
    void reset()
    {
        valid = false;
        k = 0;
        for (DBuf<double> *b : {&d_ut, &d_v, &d_u, &d_w, &d_part, &d_mag, &d_t}) b->release();
        for (DBuf<float> *b : {&d_uf, &d_vf, &d_us, &d_vs, &d_od}) b->release();
        d_uh.release();
        d_vh.release();
        d_piv.release();
        d_excl.release();
        prow.clear();
        pcol.clear();
        cdiag.clear();
        opdiag.clear();
    }
};

constexpr uint32_t kHMaxRank = 32;

// Hierarchical block form of the sparse-layout operator (hmatrix.hip):
// Morton-ordered compressed rows / columns, far blocks in low rank (U and
// V^T row-major, FP16 with a scale per row or FP32), the rest as a sparse
// near field.  Indices below are positions in the Morton order.
struct Hmat {
    bool valid = false, keep = false;
    uint32_t precision = 0, nlr = 0, nleaf = 0, n_itn = 0, n_itt = 0;
    uint64_t nnz_near = 0, bytes = 0;
    std::vector<uint32_t> perm;                 // position -> bin
    std::vector<uint4> far;                     // r0 r1 c0 c1 of every far block
    std::vector<uint32_t> status, rank;         // per far block (IRTT_HBLOCK_*)
    std::vector<double> err_rel;
    std::vector<int32_t> list_pos;              // far block -> ACA list entry (-1: none)
    std::vector<unsigned long long> list_uoff, list_voff;
    std::vector<int32_t> piv;                   // 2 * kHMaxRank per list entry
    DBuf<uint32_t> d_perm;
    DBuf<uint32_t> prp, pcol, pcp, prow;        // the whole operator, Morton order
    DBuf<float> pval, pcv;
    DBuf<uint32_t> nrp, ncol, ncp, nrow;        // near field
    DBuf<float> nval, ncv;
    DBuf<double> Ub, Vb;                        // FP64 factors as built (rank-major)
    DBuf<double> d_err;
    // accepted blocks: rank, first row / column, factor and scale offsets
    DBuf<uint32_t> d_rk, d_r0, d_c0;
    DBuf<unsigned long long> d_uo, d_vo, d_suo, d_svo;
    DBuf<uint16_t> uh, vh;                      // FP16 factors (bit patterns)
    DBuf<float> uf, vf, su, sv;                 // FP32 factors; FP16 row scales
    // product plan
    DBuf<uint32_t> d_leaf_beg, d_rptr, d_rblk, d_cptr, d_cblk, d_ipn, d_ipt;
    DBuf<uint4> d_itn, d_itt;
    DBuf<double> tpart, tq, xp, yp;
    void reset()
    {
        valid = keep = false;
        nlr = nleaf = n_itn = n_itt = 0;
        nnz_near = bytes = 0;
        perm.clear();
        far.clear();
        status.clear();
        rank.clear();
        err_rel.clear();
        list_pos.clear();
        list_uoff.clear();
        list_voff.clear();
        piv.clear();
        for (DBuf<uint32_t> *b : {&d_perm, &prp, &pcol, &pcp, &prow, &nrp, &ncol, &ncp, &nrow, &d_rk,
                                  &d_r0, &d_c0, &d_leaf_beg, &d_rptr, &d_rblk, &d_cptr, &d_cblk,
                                  &d_ipn, &d_ipt})
            b->release();
        for (DBuf<float> *b : {&pval, &pcv, &nval, &ncv, &uf, &vf, &su, &sv}) b->release();
        for (DBuf<double> *b : {&Ub, &Vb, &d_err, &tpart, &tq, &xp, &yp}) b->release();
        for (DBuf<unsigned long long> *b : {&d_uo, &d_vo, &d_suo, &d_svo}) b->release();
        uh.release();
        vh.release();
        d_itn.release();
        d_itt.release();
    }
};

// Pinned host blocks for read-backs, taken from and returned to a
// process-wide pool (pinned_take / pinned_give, api.cpp): freeing pinned
// memory synchronises the whole device, which would stall every other
// scene's stream each time a served request ends.
void *pinned_take(size_t bytes, size_t &got);
void pinned_give(void *p, size_t bytes);

struct HostPin {
    void *p = nullptr;
    size_t n = 0;
    HostPin() = default;
    HostPin(const HostPin &) = delete;
    HostPin &operator=(const HostPin &) = delete;
    ~HostPin()
    {
        if (p) pinned_give(p, n);
    }
    // a block of at least `bytes`; the caller has synchronised the stream
    // that last copied into the current block before asking for a larger one
    void *get(size_t bytes)
    {
        if (bytes <= n) return p;
        if (p) pinned_give(p, n);
        p = nullptr;
        n = 0;
        p = pinned_take(bytes, n);
        return p;
    }
};

struct Scene {
    uint32_t nv = 0, nt = 0, nb = 0, bin_mode = 0;
    uint32_t ld = 0;  // row pitch of the dense operator: nb rounded up to 64
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    HostPin h_pin;                     // transport build read-backs

    // host copies of the inputs
    std::vector<float> xyz;
    std::vector<uint32_t> idx;
    std::vector<irtt_material> mat;
    std::vector<uint8_t> flags;
    std::vector<float> rho_d, rho_s, eps, tri_area;  // FP32 as the shader sees them

    // per-bin physics (FP64), ThermalData::load / computeLumpedThermalMass /
    // computeConductiveLaplacian
    std::vector<double> bin_area, bin_eps_area, bin_mass, bin_source;
    bool source_explicit = false;     // bin_source set by irtt_scene_set_source
    std::vector<double> bin_kelvin0;   // initial / fixed values
    std::vector<uint8_t> bin_fixed;
    std::vector<double> x;             // current state (K)
    Csr lap;

    // device geometry
    DBuf<float> d_xyz;
    DBuf<uint32_t> d_idx;
    DBuf<float> d_rho_d, d_rho_s;
    DBuf<uint8_t> d_flags;
    DBuf<uint32_t> d_traceable;       // ids of traceable triangles
    Bvh bvh;

    // transport
    bool has_transport = false;
    uint32_t layout = IRTT_LAYOUT_DENSE;  // of the built operator
    irtt_transport_params tparams{};
    Sparse sp;                        // IRTT_LAYOUT_SPARSE
    AdjWork adj;                      // adjoint trace work space
    double quantum = 0.0;
    DBuf<uint64_t> d_q;               // per-triangle fixed-point weight
    std::vector<uint64_t> h_q;        // what d_q holds (uploads of the same weights are skipped)
    uint64_t assembly_gen = 0;        // bumped by every assembly (areas, emissivities, bins)
    uint64_t wq_gen = ~0ull, wq_total = 0;  // assembly and photon count h_q, h_diag were made for
    uint32_t wq_per = 0;
    double wq_quantum = 0.0;          // Q of h_q
    DBuf<unsigned long long> d_acc;   // nb*ld fixed-point accumulators
    DBuf<uint8_t> d_touch;            // nb*ld >> kTouchShift chunk marks
    bool acc_marked = false;          // d_acc is zero outside the chunks d_touch marks
    uint64_t t_fill_gen = ~0ull;      // d_T, d_tdiag outside marked chunks: zeros and the
                                      // FP32 diagonal of this assembly
    DBuf<float> d_T;                  // nb*ld FP32 operator, row-major
    DBuf<double> d_diag;              // sigma*eps*A per bin
    std::vector<double> h_diag;       // what d_diag holds
    DBuf<unsigned long long> d_colacc;  // column sums of d_acc
    DBuf<float> d_tdiag;              // diagonal of d_T
    std::vector<double> colsum;        // column sums of T (exact)
    std::vector<double> Tdiag;         // diagonal of the FP32 operator as stored
    DBuf<unsigned long long> d_stats;
    DBuf<double> d_energy;            // energy figures of the last build (k_energy)
    DBuf<uint32_t> d_absorber;        // absorbing bin of every photon (~0: none), for the adjoint
    uint64_t fwd_rays = 0;            // ray segments of the last transport build
    uint64_t fwd_segments = 0;        // ... of them accepted (not a miss, too close or backface)
    PathStore paths;                  // path records of the last build, if kept

    // operator forms (operator.hip, aca.hip)
    uint32_t op_kind = IRTT_OP_DENSE;
    bool f16_valid = false;
    DBuf<uint16_t> d_h16;             // nb*ld FP16 off-diagonal operator (bit patterns)
    DBuf<float> d_h16s;               // its per-row scales
    DBuf<uint16_t> d_h16f, d_h16ft;   // the same (and its transpose) in MFMA fragment order
    DBuf<float> d_u32;                // FP32 copy of a product's input vector
    DBuf<double> d_cd;                // diagonal outside the stored product
    DBuf<float> d_xs;                 // per-vector FP16 scales of a batched product
    DBuf<double> d_xsm;               // their per-slice maxima
    DBuf<uint16_t> d_xh;              // FP16 vectors of a batched product (bit patterns)
    DBuf<float> d_gpart;              // split-K partial sums of the MFMA GEMM
    DBuf<double> d_x64;               // padded vectors of the FP64 MFMA GEMM
    DBuf<double> d_gpart64;           // its split-K partial sums
    Aca aca;
    Hmat hm;

    // solver workspace
    DBuf<double> d_vec;               // 12 vectors of nb doubles
    DBuf<double> d_red;               // reduction scratch
    DBuf<double> d_tpart;             // row-chunk partial sums of T^T u
    DBuf<uint32_t> d_lap_rowptr, d_lap_col;
    DBuf<double> d_lap_val;

    uint64_t device_bytes() const;
    // bytes one product with the selected operator form reads
    size_t op_bytes() const
    {
        if (op_kind == IRTT_OP_HMATRIX) return (size_t)hm.bytes;
        return op_kind == IRTT_OP_SPARSE ? (size_t)sp.nnz * 8
                                         : (size_t)nb * ld * (op_kind == IRTT_OP_DENSE_F16 ? 2 : 4);
    }
};

// operator forms whose Krylov products approximate the stored operator (FP16
// entries, low-rank blocks): residuals and refinement use the stored one
inline bool op_refined(const Scene &s)
{
    return s.op_kind == IRTT_OP_DENSE_F16 || s.op_kind == IRTT_OP_HMATRIX;
}

// api.cpp: thread-local message returned by irtt_last_error()
void set_last_error(const std::string &msg);

// lbvh.hip
void bvh_build(Scene &s);
// trace.hip
void trace_rays(Scene &s, const float *rays, uint32_t n, float tmin, float tmax, int32_t *prim,
                float *t, float *b1, float *b2);
void transport_build(Scene &s, const irtt_transport_params &p, irtt_transport_stats &st);
void trace_rays_noalloc(Scene &s, const float *rays, uint32_t n, float tmin, float tmax,
                        int32_t *prim, float *t, float *b1, float *b2);
uint32_t trace_rays_noalloc_max();
// solver.hip
void solve(Scene &s, const irtt_solver_params &p, irtt_solve_stats &st);
void residual_host(Scene &s, const irtt_solver_params &p, const double *x, double *r);
void adjoint_solve(Scene &s, const irtt_solver_params &p, const double *x, const double *rhs,
                   double *lambda, irtt_solve_stats &st);
// operator.hip
void op_reset(Scene &s);
void op_select(Scene &s, uint32_t kind);
void op_apply(Scene &s, const double *w, double *y, bool transposed, bool refine);
void op_prepare(Scene &s, bool transposed);
void op_apply_batch(Scene &s, uint32_t R, const double *X, double *Y, bool transposed,
                    bool refine = false);
const std::vector<double> &op_diag_outside(const Scene &s, bool refine);
const std::vector<double> &op_diag_total(const Scene &s);
// history.cpp
void simulate_history(Scene &s, const irtt_solver_params &p, uint32_t steps, double *history,
                      irtt_solve_stats &st);
void adjoint_history(Scene &s, const irtt_solver_params &p, uint32_t steps, const double *hist,
                     const double *w, const double *targets, uint32_t target_stride,
                     double *lam_hist, double *objective, const irtt_adjoint_params *ap,
                     double *grad_x, irtt_solve_stats &st);
// design.cpp
void source_gradient(const Scene &s, const double *lam, double *grad);
void design_gradient(Scene &s, const irtt_solver_params &p, const double *dLdx,
                     const irtt_adjoint_params &ap, uint32_t nnz, const uint32_t *vertex,
                     const uint32_t *param, const float *dxdp, uint32_t n_params, double *dp,
                     double *grad_out, double *lam_out, irtt_solve_stats &st);
// the adjoint trace, source term and projection of design_gradient for a
// given adjoint state lam at state x (design.cpp)
void gradient_from_adjoint(Scene &s, const irtt_solver_params &p, const double *x,
                           const double *lam, const irtt_adjoint_params &ap, uint32_t nnz,
                           const uint32_t *vertex, const uint32_t *param, const float *dxdp,
                           uint32_t n_params, double *dp, double *grad_out);
// batch.hip
void solve_batch(Scene &s, const irtt_solver_params &p, uint32_t R, const double *source,
                 double *kelvin, irtt_solve_stats &st);
void adjoint_solve_batch(Scene &s, const irtt_solver_params &p, uint32_t R, const double *x,
                         const double *rhs, double *lambda, irtt_solve_stats &st);
// sparse.hip
void sparse_table_begin(Scene &s, uint64_t photons, PhotonSink &sink);
void sparse_table_finish(Scene &s, const PhotonSink &sink);
void sp_apply(Scene &s, const double *w, double *y, bool transposed);
void sp_apply_batch(Scene &s, uint32_t R, const double *X, double *Y, bool transposed);
void sp_download(Scene &s, uint32_t *rowptr, uint32_t *col, float *val);
// aca.hip
void aca_compress(Scene &s, const irtt_aca_params &p, irtt_aca_stats &st);
void aca_apply(Scene &s, const double *w, double *y, bool transposed);
void aca_get(Scene &s, double *U, double *V, double *cdiag, int32_t *prow, int32_t *pcol);
// hmatrix.hip
void hm_compress(Scene &s, const irtt_hmatrix_params &p, irtt_hmatrix_stats &st);
void hm_apply(Scene &s, const double *w, double *y, bool transposed);
void hm_get(Scene &s, uint32_t *perm, uint32_t *n_far, uint32_t *blocks, double *err_rel,
            int32_t *pivots);
void hm_get_factors(Scene &s, uint32_t b, double *U, double *V);
// trace.hip: a forward pass that only records each photon's absorbing bin
// (absorber: one word per photon, ~0 where none); returns the ray segments
// traced and of them the accepted ones
void absorber_pass(Scene &s, uint32_t *absorber, uint64_t &rays, uint64_t &segments);
// device bytes of the path records a first build with IRTT_TRANSPORT_KEEP_PATHS
// sets aside for `photons` photons (before its 30 %-of-free cap)
uint64_t kept_path_bytes(uint64_t photons);
// adjoint.hip
uint64_t adjoint_record_bytes(uint64_t photons, uint64_t segments);
void adjoint_trace(Scene &s, const irtt_adjoint_params &ap, uint32_t n_hist, const double *lambda,
                   const double *x, double *grad, irtt_adjoint_stats &st);

}  // namespace irtt
