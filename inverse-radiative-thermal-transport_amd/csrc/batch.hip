// batch.hip -- many steady solves (and adjoint solves) that share one
// operator, in lockstep, so every Krylov product is one multi-vector product
// (an FP16 MFMA GEMM in the FP16 form) instead of one pass over the operator
// per solve.
//
// Each column is the solve of irtt_solve / irtt_adjoint_solve for its own
// external source (or right-hand side): the same residual, exact Jacobian,
// linear y = x^4 predictor, BiCGSTAB with right Jacobi preconditioning and
// damped line search (thermal_solver.cpp:119-229), with every scalar of the
// iteration kept per column.  With the FP16 form the predictor and every
// Krylov product use FP16 GEMMs while the Newton residuals use the FP32
// operator (FP64 accumulation): iterative refinement, the correction solves
// staying on FP16 products.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "irtt_internal.hpp"

namespace irtt {
namespace {

constexpr int kBRed = 64;  // reduction blocks per column

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

struct BSys {
    uint32_t nb, ld;
    const uint32_t *lrow, *lcol;
    const double *lval;
    const double *kfix, *tdiag, *ldiag;
    const double *cdr, *cdj;  // diagonal outside the product: residual / Krylov side
    const uint8_t *fixed;
    int predict;               // Jacobian of the y-space predictor (the operator alone)
    int transposed;            // adjoint: products with J^T
};

__device__ __forceinline__ double blap(const BSys &s, uint32_t i, const double *v)
{
    if (!s.lrow) return 0.0;
    double a = 0.0;
    for (uint32_t k = s.lrow[i]; k < s.lrow[i + 1]; k++) a += s.lval[k] * v[s.lcol[k]];
    return a;
}

// column c of an R x ld array
#define COL(p) ((p) + (size_t)blockIdx.y * s.ld)

__global__ void kb_pow4(BSys s, const double *__restrict__ x, double *__restrict__ xe,
                        double *__restrict__ y)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nb) return;
    double v = s.fixed[i] ? s.kfix[i] : COL(x)[i];
    COL(xe)[i] = v;
    double v2 = v * v;
    COL(y)[i] = v2 * v2;
}

// r = Ty + cdr y + L xe - b (free bins), part[c][block] = sum r^2
__global__ void kb_residual(BSys s, const double *__restrict__ Ty, const double *__restrict__ y,
                            const double *__restrict__ xe, const double *__restrict__ b,
                            double *__restrict__ r, double *__restrict__ part)
{
    __shared__ double sm[256];
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < s.nb; i += gridDim.x * blockDim.x) {
        double v = 0.0;
        if (!s.fixed[i])
            v = COL(Ty)[i] + (s.cdr ? s.cdr[i] * COL(y)[i] : 0.0) + blap(s, i, COL(xe)) - COL(b)[i];
        COL(r)[i] = v;
        acc += v * v;
    }
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.y * kBRed + blockIdx.x] = sm[0];
}

// w = fixed ? 0 : (D before the product ? 4 x^3 v : v); exact Jacobian T D,
// its transpose D T^T
__global__ void kb_jv_pre(BSys s, const double *__restrict__ xe, const double *__restrict__ v,
                          double *__restrict__ w)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nb) return;
    double vi = s.fixed[i] ? 0.0 : COL(v)[i];
    if (!s.predict && !s.transposed) {
        double x = COL(xe)[i];
        vi = 4.0 * x * x * x * vi;
    }
    COL(w)[i] = vi;
}

__global__ void kb_jv_post(BSys s, const double *__restrict__ xe, const double *__restrict__ v,
                           const double *__restrict__ w, const double *__restrict__ Tw,
                           double *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nb) return;
    if (s.fixed[i]) {
        COL(out)[i] = COL(v)[i];
        return;
    }
    double rad = COL(Tw)[i] + (s.cdj ? s.cdj[i] * COL(w)[i] : 0.0);
    if (s.predict) {
        COL(out)[i] = rad;
        return;
    }
    if (s.transposed) {
        double x = COL(xe)[i];
        rad = 4.0 * x * x * x * rad;
    }
    COL(out)[i] = rad + blap(s, i, COL(v));
}

__global__ void kb_jdiag(BSys s, const double *__restrict__ xe, double *__restrict__ dinv)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nb) return;
    double d;
    if (s.fixed[i]) d = 1.0;
    else if (s.predict) d = s.tdiag[i];
    else {
        double x = COL(xe)[i];
        d = 4.0 * x * x * x * s.tdiag[i] + s.ldiag[i];
    }
    COL(dinv)[i] = (d != 0.0 && isfinite(d)) ? 1.0 / d : 1.0;
}

// m dot products per column: part[(k * R + c) * kBRed + block]
__global__ void kb_dots(uint32_t n, uint32_t ld, int m, const double *a0, const double *b0,
                        const double *a1, const double *b1, double *part)
{
    __shared__ double sm[2][256];
    const size_t o = (size_t)blockIdx.y * ld;
    double acc0 = 0.0, acc1 = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        acc0 += a0[o + i] * b0[o + i];
        if (m > 1) acc1 += a1[o + i] * b1[o + i];
    }
    sm[0][threadIdx.x] = acc0;
    sm[1][threadIdx.x] = acc1;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + off];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < m; k++)
            part[((size_t)k * gridDim.y + blockIdx.y) * kBRed + blockIdx.x] = sm[k][0];
}

// out[j] = sum of kBRed partials j (fixed order)
__global__ void kb_sum(const double *__restrict__ part, uint32_t count, double *__restrict__ out)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    double a = 0.0;
    for (int b = 0; b < kBRed; b++) a += part[(size_t)j * kBRed + b];
    out[j] = a;
}

// per-column coefficients: coef[c] = {live, alpha, beta, omega}
__global__ void kb_p_update(uint32_t n, uint32_t ld, const double *__restrict__ coef,
                            const double *__restrict__ r, const double *__restrict__ v,
                            double *__restrict__ p)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double *cf = coef + 4 * blockIdx.y;
    if (i >= n || cf[0] == 0.0) return;
    size_t o = (size_t)blockIdx.y * ld + i;
    p[o] = r[o] + cf[2] * (p[o] - cf[3] * v[o]);
}

__global__ void kb_scale(uint32_t n, uint32_t ld, const double *__restrict__ coef,
                         const double *__restrict__ dinv, const double *__restrict__ x,
                         double *__restrict__ y)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || coef[4 * blockIdx.y] == 0.0) return;
    size_t o = (size_t)blockIdx.y * ld + i;
    y[o] = dinv[o] * x[o];
}

// out = x + sgn * alpha_c * y (sgn = +1 or -1)
__global__ void kb_axpy(uint32_t n, uint32_t ld, const double *__restrict__ coef, double sgn,
                        const double *__restrict__ x, const double *__restrict__ y,
                        double *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double *cf = coef + 4 * blockIdx.y;
    if (i >= n || cf[0] == 0.0) return;
    size_t o = (size_t)blockIdx.y * ld + i;
    out[o] = x[o] + sgn * cf[1] * y[o];
}

__global__ void kb_step(uint32_t n, uint32_t ld, const double *__restrict__ coef,
                        const double *__restrict__ ph, const double *__restrict__ sh,
                        const double *__restrict__ sv, const double *__restrict__ t,
                        double *__restrict__ x, double *__restrict__ r)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double *cf = coef + 4 * blockIdx.y;
    if (i >= n || cf[0] == 0.0) return;
    size_t o = (size_t)blockIdx.y * ld + i;
    x[o] += cf[1] * ph[o] + cf[3] * sh[o];
    r[o] = sv[o] - cf[3] * t[o];
}

// x_out = x + a_c dx (line search); a per column in coef[4c + 1]
__global__ void kb_trial(uint32_t n, uint32_t ld, const double *__restrict__ coef,
                         const double *__restrict__ x, const double *__restrict__ dx,
                         double *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    size_t o = (size_t)blockIdx.y * ld + i;
    out[o] = x[o] + coef[4 * blockIdx.y + 1] * dx[o];
}

__global__ void kb_y_update(BSys s, const double *__restrict__ xe, const double *__restrict__ dy,
                            double *__restrict__ x)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nb) return;
    double xg = COL(xe)[i];
    if (s.fixed[i]) {
        COL(x)[i] = xg;
        return;
    }
    double y = xg * xg * xg * xg + COL(dy)[i];
    COL(x)[i] = (y > 0.0 && isfinite(y)) ? sqrt(sqrt(y)) : xg;
}

__global__ void kb_neg(BSys s, const double *__restrict__ a, double *__restrict__ r)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < s.nb) COL(r)[i] = s.fixed[i] ? 0.0 : -COL(a)[i];
}

__global__ void kb_sub(BSys s, const double *__restrict__ a, const double *__restrict__ b,
                       double *__restrict__ r)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < s.nb) COL(r)[i] = s.fixed[i] ? 0.0 : COL(a)[i] - COL(b)[i];
}

#undef COL

// ---- per-column scalars of the lockstep BiCGSTAB in device memory ----
// bsc[c * kBS + field] per column, g[] wave-uniform fields; coef / coefh are
// the {live, alpha, beta, omega} rows the vector kernels read (coefh: the
// half step, live = the columns finishing with dx += alpha ph).  The
// recurrences are those of a host loop over the columns, in the same order
// and FP64 arithmetic.
enum { BS_RHO = 0, BS_RHO_NEW, BS_ALPHA, BS_OMEGA, BS_BETA, BS_LIVE, BS_STOP, kBS };
enum { BG_ITERS = 0, BG_FIRST, BG_ANY, kBG };

__device__ __forceinline__ void bs_coef(const double *b, double *coef, uint32_t c)
{
    coef[4 * c] = b[BS_LIVE];
    coef[4 * c + 1] = b[BS_ALPHA];
    coef[4 * c + 2] = b[BS_BETA];
    coef[4 * c + 3] = b[BS_OMEGA];
}

__global__ void kbs_init(uint32_t R, const double *__restrict__ stop, const uint8_t *__restrict__ live,
                         double *bsc, double *gs, double *coef)
{
    for (uint32_t c = threadIdx.x; c < R; c += blockDim.x) {
        double *b = bsc + (size_t)c * kBS;
        b[BS_RHO] = b[BS_ALPHA] = b[BS_OMEGA] = 1.0;
        b[BS_RHO_NEW] = b[BS_BETA] = 0.0;
        b[BS_STOP] = stop[c];
        b[BS_LIVE] = live[c] ? 1.0 : 0.0;
        bs_coef(b, coef, c);
    }
    if (threadIdx.x == 0) {
        gs[BG_ITERS] = 0.0;
        gs[BG_FIRST] = 1.0;
        gs[BG_ANY] = 1.0;
    }
}

// a stage's per-column step after its reduction (sums[k * R + c]);
// one block, threads over the columns
enum { BST_RHO = 0, BST_ALPHA, BST_HALF, BST_OMEGA, BST_END };
__global__ void kbs_step(uint32_t R, int stage, const double *__restrict__ sums, double *bsc,
                         double *gs, double *coef, double *coefh)
{
    __shared__ int any_live;
    if (threadIdx.x == 0) any_live = 0;
    __syncthreads();
    const bool first = gs[BG_FIRST] != 0.0;
    for (uint32_t c = threadIdx.x; c < R; c += blockDim.x) {
        double *b = bsc + (size_t)c * kBS;
        const bool live = b[BS_LIVE] != 0.0;
        if (stage == BST_RHO) {
            if (live) {
                atomicOr(&any_live, 1);
                const double rn = sums[c];
                b[BS_RHO_NEW] = rn;
                if (rn == 0.0 || !isfinite(rn)) b[BS_LIVE] = 0.0;
                b[BS_BETA] = first ? 0.0 : (rn / b[BS_RHO]) * (b[BS_ALPHA] / b[BS_OMEGA]);
            }
        } else if (stage == BST_ALPHA) {
            if (live) {
                const double d = sums[c];
                if (d == 0.0 || !isfinite(d)) b[BS_LIVE] = 0.0;
                else b[BS_ALPHA] = b[BS_RHO_NEW] / d;
            }
        } else if (stage == BST_HALF) {
            // columns whose s is small enough finish with dx += alpha ph
            const bool fin = live && sums[c] <= b[BS_STOP];
            coefh[4 * c] = fin ? 1.0 : 0.0;
            coefh[4 * c + 1] = b[BS_ALPHA];
            coefh[4 * c + 2] = b[BS_BETA];
            coefh[4 * c + 3] = b[BS_OMEGA];
            if (fin) b[BS_LIVE] = 0.0;
        } else if (stage == BST_OMEGA) {
            if (live) {
                const double tt = sums[R + c];
                if (tt == 0.0) b[BS_LIVE] = 0.0;
                else b[BS_OMEGA] = sums[c] / tt;
            }
        } else {
            if (live) {
                b[BS_RHO] = b[BS_RHO_NEW];
                if (sums[c] <= b[BS_STOP] || b[BS_OMEGA] == 0.0) b[BS_LIVE] = 0.0;
            }
            if (b[BS_LIVE] != 0.0) atomicOr(&any_live, 1);
        }
        bs_coef(b, coef, c);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (stage == BST_RHO) {
            // an iteration counts when some column was live at its start
            if (any_live) gs[BG_ITERS] += 1.0;
            gs[BG_FIRST] = 0.0;
        } else if (stage == BST_END) {
            gs[BG_ANY] = any_live ? 1.0 : 0.0;
        }
    }
}

enum {
    B_X = 0, B_XE, B_Y, B_TY, B_R, B_DX, B_XT, B_P, B_V, B_S, B_T, B_PH, B_SH, B_RHAT, B_W, B_TW,
    B_DINV, B_B, B_RL, B_LAM, B_AUX, B_COUNT
};

struct BCtx {
    Scene &s;
    uint32_t nb, ld, R;
    hipStream_t st;
    BSys sys{};
    DBuf<double> vec, part, sums, coef, coefh, bsc, gs;
    DBuf<double> dstop;
    DBuf<uint8_t> dlive;
    DBuf<double> kfix, tdiag, ldiag;
    DBuf<uint8_t> fixed;
    std::vector<double> hsum, hcoef;
    double *v[B_COUNT];
    dim3 g;      // elementwise grid over all columns
    dim3 gr;     // reduction grid
    BCtx(Scene &sc, uint32_t r) : s(sc), nb(sc.nb), ld(sc.ld), R(r), st(sc.stream)
    {
        g = dim3(cdiv(nb, 256), R);
        gr = dim3(kBRed, R);
    }
    double *col(int slot, uint32_t c) { return v[slot] + (size_t)c * ld; }

    // per-column sums of m dot products -> hsum[k * R + c]
    void dots(int m, const double *a0, const double *b0, const double *a1 = nullptr,
              const double *b1 = nullptr)
    {
        kb_dots<<<gr, 256, 0, st>>>(nb, ld, m, a0, b0, a1, b1, part.p);
        kb_sum<<<cdiv((size_t)m * R, 256), 256, 0, st>>>(part.p, m * R, sums.p);
        hsum.resize((size_t)m * R);
        IRTT_HIP(hipMemcpyAsync(hsum.data(), sums.p, (size_t)m * R * sizeof(double),
                                hipMemcpyDeviceToHost, st));
        IRTT_HIP(hipStreamSynchronize(st));
    }
    // per-column sums of m dot products into sums[k * R + c], on the device
    void dots_dev(int m, const double *a0, const double *b0, const double *a1 = nullptr,
                  const double *b1 = nullptr)
    {
        kb_dots<<<gr, 256, 0, st>>>(nb, ld, m, a0, b0, a1, b1, part.p);
        kb_sum<<<cdiv((size_t)m * R, 256), 256, 0, st>>>(part.p, m * R, sums.p);
    }
    void upload_coef()
    {
        IRTT_HIP(hipMemcpyAsync(coef.p, hcoef.data(), hcoef.size() * sizeof(double),
                                hipMemcpyHostToDevice, st));
    }
    // residual at x (all columns) into B_R; per-column ||r||^2 in rn
    void residual(const double *x, bool refine, std::vector<double> &rn)
    {
        kb_pow4<<<g, 256, 0, st>>>(sys, x, v[B_XE], v[B_Y]);
        op_apply_batch(s, R, v[B_Y], v[B_TY], false, refine);
        kb_residual<<<gr, 256, 0, st>>>(sys, v[B_TY], v[B_Y], v[B_XE], v[B_B], v[B_R], part.p);
        kb_sum<<<cdiv(R, 256), 256, 0, st>>>(part.p, R, sums.p);
        rn.resize(R);
        IRTT_HIP(hipMemcpyAsync(rn.data(), sums.p, R * sizeof(double), hipMemcpyDeviceToHost, st));
        IRTT_HIP(hipStreamSynchronize(st));
    }
    // out = J v (all columns), Krylov side
    void jv(const double *vin, double *out)
    {
        kb_jv_pre<<<g, 256, 0, st>>>(sys, v[B_XE], vin, v[B_W]);
        op_apply_batch(s, R, v[B_W], v[B_TW], sys.transposed != 0, false);
        kb_jv_post<<<g, 256, 0, st>>>(sys, v[B_XE], vin, v[B_W], v[B_TW], out);
    }
    // out = J v with the residual side's operator (refinement of the adjoint)
    void jv_refine(const double *vin, double *out)
    {
        kb_jv_pre<<<g, 256, 0, st>>>(sys, v[B_XE], vin, v[B_W]);
        op_apply_batch(s, R, v[B_W], v[B_TW], sys.transposed != 0, true);
        kb_jv_post<<<g, 256, 0, st>>>(sys, v[B_XE], vin, v[B_W], v[B_TW], out);
    }
};

void setup_batch(BCtx &c, const irtt_solver_params &p)
{
    Scene &s = c.s;
    const uint32_t nb = c.nb, ld = c.ld, R = c.R;
    hipStream_t st = c.st;
    c.vec.alloc((size_t)B_COUNT * R * ld);
    IRTT_HIP(hipMemsetAsync(c.vec.p, 0, c.vec.bytes(), st));
    for (int k = 0; k < B_COUNT; k++) c.v[k] = c.vec.p + (size_t)k * R * ld;
    c.part.alloc((size_t)2 * R * kBRed);
    c.sums.alloc((size_t)2 * R);
    c.coef.alloc((size_t)4 * R);
    c.coefh.alloc((size_t)4 * R);
    c.bsc.alloc((size_t)kBS * R);
    c.gs.alloc(kBG);
    c.dstop.alloc(R);
    c.dlive.alloc(R);
    c.hcoef.assign((size_t)4 * R, 0.0);
    auto up = [&](DBuf<double> &d, const std::vector<double> &h) {
        d.alloc(ld);
        IRTT_HIP(hipMemsetAsync(d.p, 0, d.bytes(), st));
        IRTT_HIP(hipMemcpyAsync(d.p, h.data(), nb * sizeof(double), hipMemcpyHostToDevice, st));
    };
    std::vector<double> ldiag(nb, 0.0);
    if (!s.lap.rowptr.empty())
        for (uint32_t i = 0; i < nb; i++)
            for (uint32_t k = s.lap.rowptr[i]; k < s.lap.rowptr[i + 1]; k++)
                if (s.lap.col[k] == i) ldiag[i] += s.lap.val[k];
    up(c.kfix, s.bin_kelvin0);
    up(c.tdiag, op_diag_total(s));
    up(c.ldiag, ldiag);
    c.fixed.alloc(nb);
    IRTT_HIP(hipMemcpyAsync(c.fixed.p, s.bin_fixed.data(), nb, hipMemcpyHostToDevice, st));
    const bool lap = !s.lap.rowptr.empty() && !s.lap.val.empty();
    if (lap && s.d_lap_rowptr.n == 0) {
        s.d_lap_rowptr.alloc(s.lap.rowptr.size());
        s.d_lap_col.alloc(s.lap.col.size());
        s.d_lap_val.alloc(s.lap.val.size());
        IRTT_HIP(hipMemcpyAsync(s.d_lap_rowptr.p, s.lap.rowptr.data(), s.lap.rowptr.size() * 4,
                                hipMemcpyHostToDevice, st));
        IRTT_HIP(hipMemcpyAsync(s.d_lap_col.p, s.lap.col.data(), s.lap.col.size() * 4,
                                hipMemcpyHostToDevice, st));
        IRTT_HIP(hipMemcpyAsync(s.d_lap_val.p, s.lap.val.data(), s.lap.val.size() * 8,
                                hipMemcpyHostToDevice, st));
    }
    BSys &y = c.sys;
    y.nb = nb;
    y.ld = ld;
    y.lrow = lap ? s.d_lap_rowptr.p : nullptr;
    y.lcol = lap ? s.d_lap_col.p : nullptr;
    y.lval = lap ? s.d_lap_val.p : nullptr;
    y.kfix = c.kfix.p;
    y.tdiag = c.tdiag.p;
    y.ldiag = c.ldiag.p;
    // op_apply_batch returns the whole operator product, diagonal included
    y.cdr = nullptr;
    y.cdj = nullptr;
    y.fixed = c.fixed.p;
    y.predict = 0;
    y.transposed = 0;
    (void)p;
}

// BiCGSTAB on all live columns: J dx = -r (B_R), dx in B_DX.  live[c] = 0
// columns keep dx = 0.  Returns the number of iterations in which some column
// was live.  The per-column scalars stay in device memory (kbs_step after each
// reduction), so the host checks once per iteration, or once per eight for
// operators up to 64 MiB (above that, the products of iterations run after
// every column has stopped would cost more than the checks they save).
uint32_t bicgstab_batch(BCtx &c, std::vector<uint8_t> live, uint32_t max_lin, double rtol,
                        const std::vector<double> &rn)
{
    const uint32_t R = c.R, nb = c.nb, ld = c.ld;
    hipStream_t st = c.st;
    kb_jdiag<<<c.g, 256, 0, st>>>(c.sys, c.v[B_XE], c.v[B_DINV]);
    for (int slot : {B_DX, B_P, B_V})
        IRTT_HIP(hipMemsetAsync(c.v[slot], 0, (size_t)R * ld * sizeof(double), st));
    kb_neg<<<c.g, 256, 0, st>>>(c.sys, c.v[B_R], c.v[B_RL]);
    IRTT_HIP(hipMemcpyAsync(c.v[B_RHAT], c.v[B_RL], (size_t)R * ld * sizeof(double),
                            hipMemcpyDeviceToDevice, st));
    std::vector<double> stop(R);
    for (uint32_t k = 0; k < R; k++) {
        stop[k] = rtol * rtol * rn[k];
        if (!(rn[k] > 0.0)) live[k] = 0;
    }
    bool any = false;
    for (uint32_t k = 0; k < R; k++) any |= live[k] != 0;
    if (!any || max_lin == 0) return 0;
    IRTT_HIP(hipMemcpyAsync(c.dstop.p, stop.data(), R * sizeof(double), hipMemcpyHostToDevice, st));
    IRTT_HIP(hipMemcpyAsync(c.dlive.p, live.data(), R, hipMemcpyHostToDevice, st));
    const uint32_t sthreads = std::min<uint32_t>(1024, (R + 63) / 64 * 64);
    kbs_init<<<1, sthreads, 0, st>>>(R, c.dstop.p, c.dlive.p, c.bsc.p, c.gs.p, c.coef.p);
    auto step = [&](int stage) {
        kbs_step<<<1, sthreads, 0, st>>>(R, stage, c.sums.p, c.bsc.p, c.gs.p, c.coef.p, c.coefh.p);
    };
    auto iteration = [&]() {
        c.dots_dev(1, c.v[B_RHAT], c.v[B_RL]);
        step(BST_RHO);
        kb_p_update<<<c.g, 256, 0, st>>>(nb, ld, c.coef.p, c.v[B_RL], c.v[B_V], c.v[B_P]);
        kb_scale<<<c.g, 256, 0, st>>>(nb, ld, c.coef.p, c.v[B_DINV], c.v[B_P], c.v[B_PH]);
        c.jv(c.v[B_PH], c.v[B_V]);
        c.dots_dev(1, c.v[B_RHAT], c.v[B_V]);
        step(BST_ALPHA);
        kb_axpy<<<c.g, 256, 0, st>>>(nb, ld, c.coef.p, -1.0, c.v[B_RL], c.v[B_V], c.v[B_S]);
        c.dots_dev(1, c.v[B_S], c.v[B_S]);
        step(BST_HALF);
        kb_axpy<<<c.g, 256, 0, st>>>(nb, ld, c.coefh.p, 1.0, c.v[B_DX], c.v[B_PH], c.v[B_DX]);
        kb_scale<<<c.g, 256, 0, st>>>(nb, ld, c.coef.p, c.v[B_DINV], c.v[B_S], c.v[B_SH]);
        c.jv(c.v[B_SH], c.v[B_T]);
        c.dots_dev(2, c.v[B_T], c.v[B_S], c.v[B_T], c.v[B_T]);
        step(BST_OMEGA);
        kb_step<<<c.g, 256, 0, st>>>(nb, ld, c.coef.p, c.v[B_PH], c.v[B_SH], c.v[B_S], c.v[B_T],
                                     c.v[B_DX], c.v[B_RL]);
        c.dots_dev(1, c.v[B_RL], c.v[B_RL]);
        step(BST_END);
    };
    const uint32_t K = c.s.op_bytes() > (64u << 20) ? 1u : 8u;
    double hg[kBG] = {0.0, 0.0, 0.0};
    uint32_t launched = 0;
    while (any && launched < max_lin) {
        const uint32_t kk = std::min<uint32_t>(K, max_lin - launched);
        for (uint32_t k = 0; k < kk; k++) iteration();
        launched += kk;
        IRTT_HIP(hipMemcpyAsync(hg, c.gs.p, sizeof(hg), hipMemcpyDeviceToHost, st));
        IRTT_HIP(hipStreamSynchronize(st));
        any = hg[BG_ANY] != 0.0;
    }
    return (uint32_t)hg[BG_ITERS];
}

}  // namespace

void solve_batch(Scene &s, const irtt_solver_params &p, uint32_t R, const double *source,
                 double *kelvin, irtt_solve_stats &stt)
{
    if (!s.has_transport) throw Error(IRTT_ERR_STATE, "transport operator not built");
    if (!p.steady) throw Error(IRTT_ERR_ARG, "batched solves are steady-state");
    if (p.jacobian != IRTT_JAC_EXACT) throw Error(IRTT_ERR_ARG, "batched solves use the exact Jacobian");
    BCtx c(s, R);
    setup_batch(c, p);
    const uint32_t nb = s.nb, ld = s.ld;
    hipStream_t st = s.stream;
    IRTT_HIP(hipEventRecord(s.ev0, st));
    // b_c = source_c + Ta^4 colsum(T); x0 from the scene state
    std::vector<double> hb((size_t)R * ld, 0.0), hx((size_t)R * ld, 0.0);
    const double Ta = p.ambient_kelvin, Ta4 = Ta >= 0.0 ? Ta * Ta * Ta * Ta : 0.0;
    for (uint32_t k = 0; k < R; k++)
        for (uint32_t i = 0; i < nb; i++) {
            double src = source ? source[(size_t)k * nb + i] : s.bin_source[i];
            hb[(size_t)k * ld + i] = src + Ta4 * s.colsum[i];
            double x0 = s.bin_fixed[i] ? s.bin_kelvin0[i] : s.x[i];
            hx[(size_t)k * ld + i] = (x0 == 0.0 && !s.bin_fixed[i]) ? 1.0 : x0;
        }
    IRTT_HIP(hipMemcpyAsync(c.v[B_B], hb.data(), hb.size() * 8, hipMemcpyHostToDevice, st));
    IRTT_HIP(hipMemcpyAsync(c.v[B_X], hx.data(), hx.size() * 8, hipMemcpyHostToDevice, st));
    const uint32_t max_newton = p.max_newton ? p.max_newton : 100;
    const uint32_t max_lin = p.max_linear ? p.max_linear : 200;
    const double f_tol = p.f_tol > 0 ? p.f_tol : 1e-6;
    const double x_tol = p.x_tol > 0 ? p.x_tol : 1e-4;
    const double lin_rtol = p.lin_rtol > 0 ? p.lin_rtol : 1e-10;
    const bool f16 = op_refined(s);
    const double krtol = f16 ? std::max(lin_rtol, 1e-6) : lin_rtol;
    stt = irtt_solve_stats{};
    std::vector<double> rn, rn_new;
    // the y-space predictor runs on the Krylov side; every Newton residual
    // after it on the residual side (the FP32 operator for the FP16 form:
    // FP16 residuals carry the 2^-11 rounding of x^4 and cannot converge)
    c.residual(c.v[B_X], !f16, rn);
    std::vector<uint8_t> live(R, 1);
    c.sys.predict = 1;
    stt.linear_iters += bicgstab_batch(c, live, max_lin, krtol, rn);
    kb_y_update<<<c.g, 256, 0, st>>>(c.sys, c.v[B_XE], c.v[B_DX], c.v[B_X]);
    c.sys.predict = 0;
    const bool refine = true;
    c.residual(c.v[B_X], refine, rn);
    std::vector<uint8_t> done(R, 0);
    std::vector<double> a(R), mdx(R, 0.0);
    std::vector<double> hdx((size_t)R * ld);
    uint32_t it = 0;
    for (; it < max_newton; it++) {
        for (uint32_t k = 0; k < R; k++)
            if (rn[k] < f_tol) done[k] = 1;
        bool all = true;
        for (uint32_t k = 0; k < R; k++) all &= done[k] != 0;
        if (all) break;
        for (uint32_t k = 0; k < R; k++) live[k] = !done[k];
        // converged columns take no step
        for (uint32_t k = 0; k < R; k++)
            if (done[k])
                IRTT_HIP(hipMemsetAsync(c.col(B_R, k), 0, ld * sizeof(double), st));
        stt.linear_iters += bicgstab_batch(c, live, max_lin, krtol, rn);
        // damped line search per column (thermal_solver.cpp:192-203)
        std::fill(a.begin(), a.end(), 1.0);
        std::vector<uint8_t> accepted(R, 0);
        for (uint32_t ls = 0; ls <= 20; ls++) {
            for (uint32_t k = 0; k < R; k++) c.hcoef[4 * k + 1] = done[k] ? 0.0 : a[k];
            c.upload_coef();
            kb_trial<<<c.g, 256, 0, st>>>(nb, ld, c.coef.p, c.v[B_X], c.v[B_DX], c.v[B_XT]);
            c.residual(c.v[B_XT], refine, rn_new);
            bool pending = false;
            for (uint32_t k = 0; k < R; k++) {
                if (done[k] || accepted[k]) continue;
                if ((rn_new[k] <= rn[k] && std::isfinite(rn_new[k])) || ls == 20) accepted[k] = 1;
                else {
                    a[k] *= 0.5;
                    pending = true;
                    stt.line_search_halvings++;
                }
            }
            if (!pending) break;
        }
        // x = x + a dx with the accepted steps (the last trial holds them)
        for (uint32_t k = 0; k < R; k++) c.hcoef[4 * k + 1] = done[k] ? 0.0 : a[k];
        c.upload_coef();
        kb_trial<<<c.g, 256, 0, st>>>(nb, ld, c.coef.p, c.v[B_X], c.v[B_DX], c.v[B_XT]);
        IRTT_HIP(hipMemcpyAsync(c.v[B_X], c.v[B_XT], (size_t)R * ld * 8, hipMemcpyDeviceToDevice, st));
        IRTT_HIP(hipMemcpyAsync(hdx.data(), c.v[B_DX], hdx.size() * 8, hipMemcpyDeviceToHost, st));
        c.residual(c.v[B_X], refine, rn);
        for (uint32_t k = 0; k < R; k++) {
            if (done[k]) continue;
            double m = 0.0;
            for (uint32_t i = 0; i < nb; i++)
                if (!s.bin_fixed[i]) m = std::max(m, std::fabs(a[k] * hdx[(size_t)k * ld + i]));
            mdx[k] = m;
            if (rn[k] < f_tol || m < x_tol || a[k] < 1e-6) done[k] = 1;
        }
    }
    stt.newton_iters = it + 1;
    // results: V_XE holds x with fixed values applied (last residual call)
    std::vector<double> hxe((size_t)R * ld);
    IRTT_HIP(hipMemcpyAsync(hxe.data(), c.v[B_XE], hxe.size() * 8, hipMemcpyDeviceToHost, st));
    IRTT_HIP(hipEventRecord(s.ev1, st));
    IRTT_HIP(hipStreamSynchronize(st));
    IRTT_HIP(hipEventElapsedTime(&stt.solve_ms, s.ev0, s.ev1));
    stt.converged = 1;
    stt.residual_sq = 0.0;
    stt.max_dx = 0.0;
    for (uint32_t k = 0; k < R; k++) {
        stt.residual_sq = std::max(stt.residual_sq, rn[k]);
        stt.max_dx = std::max(stt.max_dx, mdx[k]);
        if (!(rn[k] < f_tol || mdx[k] < x_tol)) stt.converged = 0;
        std::copy(hxe.begin() + (size_t)k * ld, hxe.begin() + (size_t)k * ld + nb,
                  kelvin + (size_t)k * nb);
    }
}

void adjoint_solve_batch(Scene &s, const irtt_solver_params &p, uint32_t R, const double *x,
                         const double *rhs, double *lambda, irtt_solve_stats &stt)
{
    if (!s.has_transport) throw Error(IRTT_ERR_STATE, "transport operator not built");
    if (!p.steady || p.jacobian != IRTT_JAC_EXACT)
        throw Error(IRTT_ERR_ARG, "batched adjoint solves are steady-state, exact Jacobian");
    BCtx c(s, R);
    setup_batch(c, p);
    c.sys.transposed = 1;
    const uint32_t nb = s.nb, ld = s.ld;
    hipStream_t st = s.stream;
    IRTT_HIP(hipEventRecord(s.ev0, st));
    std::vector<double> hx((size_t)R * ld, 0.0), hr((size_t)R * ld, 0.0);
    for (uint32_t k = 0; k < R; k++)
        for (uint32_t i = 0; i < nb; i++) {
            hx[(size_t)k * ld + i] = x[(size_t)k * nb + i];
            hr[(size_t)k * ld + i] = rhs[(size_t)k * nb + i];
        }
    IRTT_HIP(hipMemcpyAsync(c.v[B_X], hx.data(), hx.size() * 8, hipMemcpyHostToDevice, st));
    IRTT_HIP(hipMemcpyAsync(c.v[B_AUX], hr.data(), hr.size() * 8, hipMemcpyHostToDevice, st));
    kb_pow4<<<c.g, 256, 0, st>>>(c.sys, c.v[B_X], c.v[B_XE], c.v[B_Y]);
    // B_R = -rhs on free bins; J^T lambda = rhs  <=>  J^T dx = -r
    kb_neg<<<c.g, 256, 0, st>>>(c.sys, c.v[B_AUX], c.v[B_R]);
    c.dots(1, c.v[B_R], c.v[B_R]);
    std::vector<double> rn(c.hsum.begin(), c.hsum.begin() + R);
    const uint32_t max_lin = p.max_linear ? p.max_linear : 200;
    const double lin_rtol = p.lin_rtol > 0 ? p.lin_rtol : 1e-10;
    const bool f16 = op_refined(s);
    const double krtol = f16 ? std::max(lin_rtol, 1e-6) : lin_rtol;
    stt = irtt_solve_stats{};
    std::vector<uint8_t> live(R, 1);
    stt.linear_iters = bicgstab_batch(c, live, max_lin, krtol, rn);
    IRTT_HIP(hipMemcpyAsync(c.v[B_LAM], c.v[B_DX], (size_t)R * ld * 8, hipMemcpyDeviceToDevice, st));
    if (f16) {
        // refinement with the FP32 operator: r = J^T lambda - rhs
        std::vector<double> r2;
        for (int pass = 0; pass < 20; pass++) {
            c.jv_refine(c.v[B_LAM], c.v[B_TY]);
            kb_sub<<<c.g, 256, 0, st>>>(c.sys, c.v[B_TY], c.v[B_AUX], c.v[B_R]);
            c.dots(1, c.v[B_R], c.v[B_R]);
            r2.assign(c.hsum.begin(), c.hsum.begin() + R);
            bool ok = true;
            for (uint32_t k = 0; k < R; k++) {
                live[k] = r2[k] > lin_rtol * lin_rtol * rn[k] && std::isfinite(r2[k]);
                ok &= !live[k];
            }
            stt.newton_iters = pass + 1;
            if (ok) break;
            stt.linear_iters += bicgstab_batch(c, live, max_lin, krtol, r2);
            for (uint32_t k = 0; k < R; k++) c.hcoef[4 * k] = 1.0, c.hcoef[4 * k + 1] = 1.0;
            c.upload_coef();
            kb_axpy<<<c.g, 256, 0, st>>>(nb, ld, c.coef.p, 1.0, c.v[B_LAM], c.v[B_DX], c.v[B_LAM]);
        }
    }
    std::vector<double> hl((size_t)R * ld);
    IRTT_HIP(hipMemcpyAsync(hl.data(), c.v[B_LAM], hl.size() * 8, hipMemcpyDeviceToHost, st));
    IRTT_HIP(hipEventRecord(s.ev1, st));
    IRTT_HIP(hipStreamSynchronize(st));
    IRTT_HIP(hipEventElapsedTime(&stt.solve_ms, s.ev0, s.ev1));
    stt.converged = 1;
    for (uint32_t k = 0; k < R; k++)
        for (uint32_t i = 0; i < nb; i++)
            lambda[(size_t)k * nb + i] = s.bin_fixed[i] ? 0.0 : hl[(size_t)k * ld + i];
    stt.residual_sq = *std::max_element(rn.begin(), rn.end());
}

}  // namespace irtt
