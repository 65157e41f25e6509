// solver.hip -- Newton solve of the radiative + conductive energy balance on
// the device.
//
// Replaces ThermalSolver::residual / jacobian / solve
// (src/implementations/thermal_renderer/thermal_solver.cpp:13-229), which run
// on the host with Eigen.  Residual (thermal_solver.cpp:20-31, plus the
// source and ambient terms of this library):
//   steady:    r = T x^4 + L x - b
//   transient: r = M (x - x_prev) + dt (T x^4 + L x - b)
//   b = external absorbed power + T_amb^4 * colsum(T)   (ambient exchange)
// Fixed bins hold their value and have r = 0.  Newton steps solve J dx = -r
// with BiCGSTAB or restarted GMRES (right Jacobi preconditioning) using
// matrix-free products;
// the Jacobian is the exact one or the reference's row-scaled form
// (thermal_solver.cpp:39-96).  A damped line search halves the step while the
// residual norm grows (thermal_solver.cpp:192-203).  Operator products go
// through op_apply (dense FP32, dense FP16 + row scales, or ACA); with the
// FP16 form the Krylov products are FP16 while every Newton residual uses
// the FP32 operator with FP64 accumulation, so Newton is the refinement loop.
// Every vector is FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "irtt_internal.hpp"

namespace irtt {
namespace {

constexpr int kRedBlocks = 256;
constexpr int kRedThreads = 256;

struct Sys {
    uint32_t nb;
    const uint32_t *lrow, *lcol;
    const double *lval;
    const double *mass, *b, *kfix, *xprev;
    const uint8_t *fixed;
    double dt;
    int steady, jmode;  // jmode: J_EXACT, J_REFERENCE or J_PREDICT
    int transposed;     // products with J^T (adjoint solve)
    // operator diagonal kept outside the stored product (NULL = none): in the
    // residual (cdr) and in the Krylov products (cdj)
    const double *cdr, *cdj;
    int c_unscaled;     // reference ACA Jacobian: D U V + C, C not scaled by D
                        // (thermal_solver.cpp:44-75, thermal_aca_matrix.hpp:128)
};

__device__ __forceinline__ double lap_row(const Sys &s, uint32_t i, const double *v)
{
    if (!s.lrow) return 0.0;
    double a = 0.0;
    for (uint32_t k = s.lrow[i]; k < s.lrow[i + 1]; k++) a += s.lval[k] * v[s.lcol[k]];
    return a;
}

// y = (fixed ? kfix : x)^4 ; xe = (fixed ? kfix : x)
__global__ void k_pow4(Sys s, const double *__restrict__ x, double *__restrict__ xe,
                       double *__restrict__ y)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nb) return;
    double v = s.fixed[i] ? s.kfix[i] : x[i];
    xe[i] = v;
    double v2 = v * v;
    y[i] = v2 * v2;
}

// r from Ty (= P x^4), y (= x^4) and xe; part[block] = sum r^2
__global__ void k_residual(Sys s, const double *__restrict__ Ty, const double *__restrict__ y,
                           const double *__restrict__ xe, double *__restrict__ r, double *part)
{
    __shared__ double sm[kRedThreads];
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < s.nb;
         i += gridDim.x * blockDim.x) {
        double v = 0.0;
        if (!s.fixed[i]) {
            double ty = Ty[i] + (s.cdr ? s.cdr[i] * y[i] : 0.0);
            double rad = ty + lap_row(s, i, xe) - s.b[i];
            v = s.steady ? rad : s.mass[i] * (xe[i] - s.xprev[i]) + s.dt * rad;
        }
        r[i] = v;
        acc += v * v;
    }
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int off = kRedThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// Power balance at x over the free bins (from the last residual's Ty, y,
// xe): part[0..] = sum b (in), part[kRedBlocks..] = sum (T x^4 + L x) plus
// the storage rate M (x - x_prev) / dt of a transient step (out).
__global__ void k_balance(Sys s, const double *__restrict__ Ty, const double *__restrict__ y,
                          const double *__restrict__ xe, double *part)
{
    __shared__ double sm[2][kRedThreads];
    double a_in = 0.0, a_out = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < s.nb;
         i += gridDim.x * blockDim.x) {
        if (s.fixed[i]) continue;
        const double ty = Ty[i] + (s.cdr ? s.cdr[i] * y[i] : 0.0);
        a_in += s.b[i];
        a_out += ty + lap_row(s, i, xe) +
                 (s.steady ? 0.0 : s.mass[i] * (xe[i] - s.xprev[i]) / s.dt);
    }
    sm[0][threadIdx.x] = a_in;
    sm[1][threadIdx.x] = a_out;
    __syncthreads();
    for (int off = kRedThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + off];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sm[0][0];
        part[kRedBlocks + blockIdx.x] = sm[1][0];
    }
}

// Jacobian-vector product modes
enum { J_EXACT = 0, J_REFERENCE = 1, J_PREDICT = 2 };

// The radiative block is T D (exact) or D T (reference), D = diag(4x^3); its
// transpose swaps the two.  D is applied before the operator product when
// (exact != transposed), after it otherwise.
__device__ __forceinline__ bool d_before(const Sys &s)
{
    return (s.jmode == J_EXACT) != (s.transposed != 0);
}

// w = fixed ? 0 : (D before the product ? 4 x^3 v : v)
__device__ __forceinline__ double jv_pre_at(const Sys &s, const double *__restrict__ xe, double v,
                                            uint32_t i)
{
    double vi = s.fixed[i] ? 0.0 : v;
    if (s.jmode != J_PREDICT && d_before(s)) {
        double x = xe[i];
        vi = 4.0 * x * x * x * vi;
    }
    return vi;
}

__global__ void k_jv_pre(Sys s, const double *__restrict__ xe, const double *__restrict__ v,
                         double *__restrict__ w)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nb) return;
    w[i] = jv_pre_at(s, xe, v[i], i);
}

// (J v)_i from Tw (= T w).  J_PREDICT is the operator of the steady radiative
// balance in y = x^4 alone (T), used for the initial guess.
__device__ __forceinline__ double jv_post_at(const Sys &s, const double *__restrict__ xe,
                                             const double *__restrict__ v,
                                             const double *__restrict__ w,
                                             const double *__restrict__ Tw, uint32_t i)
{
    if (s.fixed[i]) return v[i];
    const double cd = s.cdj ? s.cdj[i] : 0.0;
    if (s.jmode == J_PREDICT) return Tw[i] + cd * w[i];
    double lv = lap_row(s, i, v);  // the cotangent Laplacian is symmetric
    double rad = Tw[i];
    if (!s.c_unscaled) rad += cd * w[i];
    if (!d_before(s)) {
        double x = xe[i];
        rad = 4.0 * x * x * x * rad;  // reference: diag(4x^3) T v, thermal_solver.cpp:149
    }
    if (s.c_unscaled) rad += cd * v[i];
    if (s.steady) {
        double o = rad + lv;
        if (s.jmode == J_REFERENCE) o += v[i] / s.dt;  // I/dt regulariser, thermal_solver.cpp:150
        return o;
    }
    return s.mass[i] * v[i] + s.dt * (rad + lv);
}

__global__ void k_jv_post(Sys s, const double *__restrict__ xe, const double *__restrict__ v,
                          const double *__restrict__ w, const double *__restrict__ Tw,
                          double *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nb) return;
    out[i] = jv_post_at(s, xe, v, w, Tw, i);
}

// deterministic multi-dot: part[k*kRedBlocks + block] = sum a_k * b_k
__global__ void k_dots(uint32_t n, int m, const double *a0, const double *b0, const double *a1,
                       const double *b1, const double *a2, const double *b2, double *part)
{
    __shared__ double sm[3][kRedThreads];
    double acc[3] = {0.0, 0.0, 0.0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        acc[0] += a0[i] * b0[i];
        if (m > 1) acc[1] += a1[i] * b1[i];
        if (m > 2) acc[2] += a2[i] * b2[i];
    }
    for (int k = 0; k < 3; k++) sm[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int off = kRedThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int k = 0; k < 3; k++) sm[k][threadIdx.x] += sm[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < m; k++) part[k * kRedBlocks + blockIdx.x] = sm[k][0];
}

__global__ void k_final_sum(const double *part, int m, int nparts, double *out)
{
    __shared__ double sm[kRedThreads];
    for (int k = 0; k < m; k++) {
        double a = 0.0;
        for (int i = threadIdx.x; i < nparts; i += blockDim.x) a += part[k * kRedBlocks + i];
        sm[threadIdx.x] = a;
        __syncthreads();
        for (int off = kRedThreads / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = sm[0];
        __syncthreads();
    }
}

// generic fused vector updates
// mode 0: p = r + beta (p - omega v)
// mode 1: y = a * x (pointwise Jacobi: y = x * dinv)
// mode 2: s = r - alpha v
// mode 3: x += alpha ph + omega sh ; r = s - omega t
__global__ void k_p_update(uint32_t n, double beta, double omega, const double *r, const double *v,
                           double *p)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = r[i] + beta * (p[i] - omega * v[i]);
}
__global__ void k_scale(uint32_t n, const double *dinv, const double *x, double *y)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = dinv[i] * x[i];
}
__global__ void k_axpy2(uint32_t n, double alpha, const double *x, const double *y, double *out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = x[i] + alpha * y[i];
}
__global__ void k_step(uint32_t n, double alpha, double omega, const double *ph, const double *sh,
                       const double *s, const double *t, double *x, double *r)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x[i] += alpha * ph[i] + omega * sh[i];
        r[i] = s[i] - omega * t[i];
    }
}
// Jacobi diagonal: dinv_i = 1 / J_ii
__global__ void k_jdiag(Sys s, const double *__restrict__ xe, const double *__restrict__ Tdiag,
                        const double *__restrict__ Ldiag, double *__restrict__ dinv)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nb) return;
    if (s.fixed[i]) {
        dinv[i] = 1.0;
        return;
    }
    double x = xe[i];
    double rad = 4.0 * x * x * x * Tdiag[i];
    if (s.c_unscaled) rad = 4.0 * x * x * x * (Tdiag[i] - s.cdj[i]) + s.cdj[i];
    double d;
    if (s.jmode == J_PREDICT) d = Tdiag[i];
    else if (s.steady) d = rad + Ldiag[i] + (s.jmode == J_REFERENCE ? 1.0 / s.dt : 0.0);
    else d = s.mass[i] + s.dt * (rad + Ldiag[i]);
    dinv[i] = (d != 0.0 && isfinite(d)) ? 1.0 / d : 1.0;
}

// after the predictor solve T dy = -r (at x_g): x = (x_g^4 + dy)^(1/4) on
// free bins whose updated y stays positive
__global__ void k_y_update(Sys s, const double *__restrict__ xe, const double *__restrict__ dy,
                           double *__restrict__ x)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nb) return;
    double xg = xe[i];
    if (s.fixed[i]) {
        x[i] = xg;
        return;
    }
    double y = xg * xg * xg * xg + dy[i];
    x[i] = (y > 0.0 && isfinite(y)) ? sqrt(sqrt(y)) : xg;
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// vector slots in d_vec (each ld doubles, padded with zeros)
enum {
    V_X = 0, V_XE, V_Y, V_TY, V_R, V_DX, V_XT, V_P, V_V, V_S, V_T, V_PH, V_SH, V_RHAT, V_W, V_TW,
    V_DINV, V_B, V_KFIX, V_XPREV, V_MASS, V_TDIAG, V_LDIAG, V_CDR, V_CDJ, V_LAM, V_RHS, V_AUX,
    V_COUNT
};

struct Ctx {
    Scene &s;
    hipStream_t st;
    uint32_t nb, ld;
    Sys sys;
    double *v[V_COUNT];
    double *part, *out;
    double *sc;  // device scalars of the Krylov iteration
    double hbuf[4];
    uint32_t krylov = IRTT_KRYLOV_BICGSTAB, restart = 30;  // linear solver
    explicit Ctx(Scene &sc) : s(sc), st(sc.stream), nb(sc.nb), ld(sc.ld) {}

    // operator product of the residual side
    void gemv_res(const double *u, double *y) { op_apply(s, u, y, false, true); }
    // returns sum of m dot products
    void dots(int m, const double *a0, const double *b0, const double *a1, const double *b1,
              const double *a2, const double *b2, double *res)
    {
        k_dots<<<kRedBlocks, kRedThreads, 0, st>>>(nb, m, a0, b0, a1, b1, a2, b2, part);
        k_final_sum<<<1, kRedThreads, 0, st>>>(part, m, kRedBlocks, out);
        IRTT_HIP(hipMemcpyAsync(hbuf, out, m * sizeof(double), hipMemcpyDeviceToHost, st));
        IRTT_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < m; k++) res[k] = hbuf[k];
    }
    // r(x) into V_R, returns ||r||^2; V_XE, V_Y, V_TY updated
    double residual(const double *x)
    {
        uint32_t g = cdiv(nb, 256);
        k_pow4<<<g, 256, 0, st>>>(sys, x, v[V_XE], v[V_Y]);
        gemv_res(v[V_Y], v[V_TY]);
        k_residual<<<kRedBlocks, kRedThreads, 0, st>>>(sys, v[V_TY], v[V_Y], v[V_XE], v[V_R],
                                                       part);
        k_final_sum<<<1, kRedThreads, 0, st>>>(part, 1, kRedBlocks, out);
        IRTT_HIP(hipMemcpyAsync(hbuf, out, sizeof(double), hipMemcpyDeviceToHost, st));
        IRTT_HIP(hipStreamSynchronize(st));
        return hbuf[0];
    }
    // out = J v (J^T v when transposed); refine: with the residual side's
    // operator (FP32 when the Krylov side is FP16)
    void jv(const double *vin, double *outv, bool refine = false)
    {
        uint32_t g = cdiv(nb, 256);
        Sys y = sys;
        if (refine) y.cdj = y.cdr;
        k_jv_pre<<<g, 256, 0, st>>>(y, v[V_XE], vin, v[V_W]);
        op_apply(s, v[V_W], v[V_TW], sys.transposed != 0, refine);
        k_jv_post<<<g, 256, 0, st>>>(y, v[V_XE], vin, v[V_W], v[V_TW], outv);
    }
};

void upload_static(Scene &s, Ctx &c, const irtt_solver_params &p)
{
    const uint32_t nb = s.nb, ld = s.ld;
    s.d_vec.alloc((size_t)V_COUNT * ld);
    s.d_red.alloc(4 * kRedBlocks + 8 + 32);
    IRTT_HIP(hipMemsetAsync(s.d_vec.p, 0, s.d_vec.bytes(), s.stream));
    for (int k = 0; k < V_COUNT; k++) c.v[k] = s.d_vec.p + (size_t)k * ld;
    c.part = s.d_red.p;
    c.out = s.d_red.p + 4 * kRedBlocks;
    c.sc = s.d_red.p + 4 * kRedBlocks + 8;
    std::vector<double> b(nb), kfix(nb), ldiag(nb, 0.0);
    const double Ta = p.ambient_kelvin;
    for (uint32_t i = 0; i < nb; i++) {
        double amb = Ta >= 0.0 ? Ta * Ta * Ta * Ta * s.colsum[i] : 0.0;
        b[i] = s.bin_source[i] + amb;
        kfix[i] = s.bin_kelvin0[i];
    }
    if (!s.lap.rowptr.empty())
        for (uint32_t i = 0; i < nb; i++)
            for (uint32_t k = s.lap.rowptr[i]; k < s.lap.rowptr[i + 1]; k++)
                if (s.lap.col[k] == i) ldiag[i] += s.lap.val[k];
    auto up = [&](int slot, const std::vector<double> &h) {
        IRTT_HIP(hipMemcpyAsync(c.v[slot], h.data(), nb * sizeof(double), hipMemcpyHostToDevice,
                                s.stream));
    };
    up(V_B, b);
    up(V_KFIX, kfix);
    up(V_MASS, s.bin_mass);
    up(V_TDIAG, op_diag_total(s));
    up(V_LDIAG, ldiag);
    const std::vector<double> &cdr = op_diag_outside(s, true), &cdj = op_diag_outside(s, false);
    if (!cdr.empty()) up(V_CDR, cdr);
    if (!cdj.empty()) up(V_CDJ, cdj);
    up(V_XPREV, s.x);
    if (!s.lap.rowptr.empty() && s.d_lap_rowptr.n == 0) {
        s.d_lap_rowptr.alloc(s.lap.rowptr.size());
        s.d_lap_col.alloc(std::max<size_t>(1, s.lap.col.size()));
        s.d_lap_val.alloc(std::max<size_t>(1, s.lap.val.size()));
        IRTT_HIP(hipMemcpyAsync(s.d_lap_rowptr.p, s.lap.rowptr.data(),
                                s.lap.rowptr.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                                s.stream));
        if (!s.lap.col.empty()) {
            IRTT_HIP(hipMemcpyAsync(s.d_lap_col.p, s.lap.col.data(), s.lap.col.size() * sizeof(uint32_t),
                                    hipMemcpyHostToDevice, s.stream));
            IRTT_HIP(hipMemcpyAsync(s.d_lap_val.p, s.lap.val.data(), s.lap.val.size() * sizeof(double),
                                    hipMemcpyHostToDevice, s.stream));
        }
    }
}

}  // namespace

// fixed flags live in a device byte array allocated per solve
struct FixedBuf {
    DBuf<uint8_t> b;
};

static void setup(Scene &s, Ctx &c, const irtt_solver_params &p, FixedBuf &fb)
{
    if (!s.has_transport) throw Error(IRTT_ERR_STATE, "transport operator not built");
    upload_static(s, c, p);
    fb.b.alloc(s.nb);
    IRTT_HIP(hipMemcpyAsync(fb.b.p, s.bin_fixed.data(), s.nb, hipMemcpyHostToDevice, s.stream));
    Sys &y = c.sys;
    y.nb = s.nb;
    bool lap = !s.lap.rowptr.empty() && s.lap.val.size() > 0;
    y.lrow = lap ? s.d_lap_rowptr.p : nullptr;
    y.lcol = lap ? s.d_lap_col.p : nullptr;
    y.lval = lap ? s.d_lap_val.p : nullptr;
    y.mass = c.v[V_MASS];
    y.b = c.v[V_B];
    y.kfix = c.v[V_KFIX];
    y.xprev = c.v[V_XPREV];
    y.fixed = fb.b.p;
    y.dt = p.dt > 0.0 ? p.dt : 300.0;
    y.steady = p.steady ? 1 : 0;
    y.jmode = p.jacobian == IRTT_JAC_REFERENCE ? J_REFERENCE : J_EXACT;
    y.transposed = 0;
    y.cdr = op_diag_outside(s, true).empty() ? nullptr : c.v[V_CDR];
    y.cdj = op_diag_outside(s, false).empty() ? nullptr : c.v[V_CDJ];
    y.c_unscaled = (s.op_kind == IRTT_OP_ACA && y.jmode == J_REFERENCE) ? 1 : 0;
    if (p.krylov > IRTT_KRYLOV_GMRES) throw Error(IRTT_ERR_ARG, "unknown Krylov method");
    c.krylov = p.krylov;
    c.restart = p.gmres_restart ? p.gmres_restart : 30;
}

void residual_host(Scene &s, const irtt_solver_params &p, const double *x, double *r)
{
    Ctx c(s);
    FixedBuf fb;
    setup(s, c, p, fb);
    IRTT_HIP(hipMemcpyAsync(c.v[V_XT], x, s.nb * sizeof(double), hipMemcpyHostToDevice, s.stream));
    c.residual(c.v[V_XT]);
    IRTT_HIP(hipMemcpyAsync(r, c.v[V_R], s.nb * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    IRTT_HIP(hipStreamSynchronize(s.stream));
}

// ---- BiCGSTAB with its scalars in device memory --------------------------
// The iteration's scalars (rho, alpha, omega, ...) live in a small device
// array: each reduction's last kernel writes into it and derives the next
// scalars, vector kernels read them, and a converged or broken-down
// iteration sets `done`, after which every update is a no-op.  No host round
// trip is needed inside an iteration, so K iterations are enqueued back to
// back and the host reads `done` once per K.  (Capturing the K iterations
// into a hipGraph per solve measured slower under concurrent serving: the
// instantiation costs more than the launches it saves.)
enum { B_RHO = 0, B_RHO_NEW, B_ALPHA, B_OMEGA, B_BETA, B_DONE, B_STOP, B_TMP, B_SS, B_TS, B_TT,
       B_RR, B_ITERS, B_FIRST, B_HALF, B_NSC };

__global__ void kd_init(double *sc, double stop)
{
    for (int k = 0; k < B_NSC; k++) sc[k] = 0.0;
    sc[B_RHO] = sc[B_ALPHA] = sc[B_OMEGA] = 1.0;
    sc[B_FIRST] = 1.0;
    sc[B_STOP] = stop;
}

// the scalar steps of the iteration (run by one thread after a reduction)
// rho_new from the last reduction; breakdown ends the solve
__device__ __forceinline__ void sc_rho(double *sc)
{
    if (sc[B_DONE] != 0.0) return;
    const double rn = sc[B_RHO_NEW];
    if (rn == 0.0 || !isfinite(rn)) {
        sc[B_DONE] = 1.0;
        return;
    }
    sc[B_BETA] = sc[B_FIRST] != 0.0 ? 0.0 : (rn / sc[B_RHO]) * (sc[B_ALPHA] / sc[B_OMEGA]);
    sc[B_FIRST] = 0.0;
}
__device__ __forceinline__ void sc_alpha(double *sc)
{
    if (sc[B_DONE] != 0.0) return;
    const double d = sc[B_TMP];
    if (d == 0.0 || !isfinite(d)) sc[B_DONE] = 1.0;
    else sc[B_ALPHA] = sc[B_RHO_NEW] / d;
}
// ||s||^2 small enough: finish with dx += alpha ph
__device__ __forceinline__ void sc_half(double *sc)
{
    sc[B_HALF] = (sc[B_DONE] == 0.0 && sc[B_SS] <= sc[B_STOP]) ? 1.0 : 0.0;
}
__device__ __forceinline__ void sc_half_done(double *sc)
{
    if (sc[B_HALF] != 0.0) {
        sc[B_DONE] = 1.0;
        sc[B_ITERS] += 1.0;
    }
}
__device__ __forceinline__ void sc_omega(double *sc)
{
    if (sc[B_DONE] != 0.0) return;
    if (sc[B_TT] == 0.0) sc[B_DONE] = 1.0;
    else sc[B_OMEGA] = sc[B_TS] / sc[B_TT];
}
__device__ __forceinline__ void sc_end(double *sc)
{
    if (sc[B_DONE] != 0.0) return;
    sc[B_RHO] = sc[B_RHO_NEW];
    if (sc[B_RR] <= sc[B_STOP]) {
        sc[B_ITERS] += 1.0;
        sc[B_DONE] = 1.0;
    } else if (sc[B_OMEGA] == 0.0) {
        sc[B_DONE] = 1.0;
    } else {
        sc[B_ITERS] += 1.0;
    }
}

// which scalar steps follow a reduction
enum { F_RHO = 0, F_ALPHA, F_HALF, F_OMEGA, F_END_RHO };

// k_final_sum of m partial sums, then the scalar steps of `stage`, which
// stores the sums in their slots of sc.
// F_END_RHO's second sum (R . r) is the next iteration's rho: it is stored
// only after the end step has used the previous one.
__global__ void kd_final(const double *part, int m, double *sc, int stage)
{
    __shared__ double sm[kRedThreads];
    double res[2] = {0.0, 0.0};
    for (int k = 0; k < m; k++) {
        double a = 0.0;
        for (int i = threadIdx.x; i < kRedBlocks; i += blockDim.x) a += part[k * kRedBlocks + i];
        sm[threadIdx.x] = a;
        __syncthreads();
        for (int off = kRedThreads / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
            __syncthreads();
        }
        res[k] = sm[0];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    switch (stage) {
    case F_RHO:
        sc[B_RHO_NEW] = res[0];
        sc_rho(sc);
        break;
    case F_ALPHA:
        sc[B_TMP] = res[0];
        sc_alpha(sc);
        break;
    case F_HALF:
        sc[B_SS] = res[0];
        sc_half(sc);
        break;
    case F_OMEGA:
        sc_half_done(sc);
        sc[B_TS] = res[0];
        sc[B_TT] = res[1];
        sc_omega(sc);
        break;
    default:
        sc[B_RR] = res[0];
        sc_end(sc);
        sc[B_RHO_NEW] = res[1];
        sc_rho(sc);
        break;
    }
}

// the reduction skeleton of k_dots: grid-stride over i, per-thread sums in
// index order, tree within the block, partials per block (the same order as
// k_dots, so the fused kernels below give the same sums)
template <int M>
__device__ __forceinline__ void block_partials(const double (&acc)[M], double *part)
{
    __shared__ double sm[M][kRedThreads];
    for (int k = 0; k < M; k++) sm[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int off = kRedThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int k = 0; k < M; k++) sm[k][threadIdx.x] += sm[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < M; k++) part[k * kRedBlocks + blockIdx.x] = sm[k][0];
}

// p = r + beta (p - omega v); ph = dinv p; w = jv_pre(ph)
__global__ void kd_p_pre(Sys sy, const double *__restrict__ sc, const double *__restrict__ xe,
                         const double *r, const double *v, double *p, const double *dinv,
                         double *ph, double *w)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sy.nb || sc[B_DONE] != 0.0) return;
    const double pi = r[i] + sc[B_BETA] * (p[i] - sc[B_OMEGA] * v[i]);
    p[i] = pi;
    const double phi = dinv[i] * pi;
    ph[i] = phi;
    w[i] = jv_pre_at(sy, xe, phi, i);
}

// out = J vin from Tw; partial sums of a0 . out (and, M = 2, out . out)
template <int M>
__global__ void kd_post_dot(Sys sy, const double *__restrict__ xe, const double *vin,
                            const double *w, const double *Tw, double *out, const double *a0,
                            double *part)
{
    double acc[M];
    for (int k = 0; k < M; k++) acc[k] = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < sy.nb;
         i += gridDim.x * blockDim.x) {
        const double o = jv_post_at(sy, xe, vin, w, Tw, i);
        out[i] = o;
        acc[0] += a0[i] * o;
        if (M > 1) acc[M - 1] += o * o;
    }
    block_partials<M>(acc, part);
}

// s = r - alpha v; partial sums of s . s
__global__ void kd_s_dot(uint32_t n, const double *__restrict__ sc, const double *r,
                         const double *v, double *s, double *part)
{
    const bool live = sc[B_DONE] == 0.0;
    double acc[1] = {0.0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double si;
        if (live) {
            si = r[i] + (-sc[B_ALPHA]) * v[i];
            s[i] = si;
        } else {
            si = s[i];
        }
        acc[0] += si * si;
    }
    block_partials<1>(acc, part);
}

// half step: dx += alpha ph; otherwise sh = dinv s; then w = jv_pre(sh)
__global__ void kd_sh_pre(Sys sy, const double *__restrict__ sc, const double *__restrict__ xe,
                          const double *ph, double *dx, const double *dinv, const double *s,
                          double *sh, double *w)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sy.nb) return;
    if (sc[B_HALF] != 0.0) {
        dx[i] = dx[i] + sc[B_ALPHA] * ph[i];
        return;
    }
    if (sc[B_DONE] != 0.0) return;
    const double shi = dinv[i] * s[i];
    sh[i] = shi;
    w[i] = jv_pre_at(sy, xe, shi, i);
}

// x += alpha ph + omega sh; r = s - omega t; partial sums of r . r and R . r
__global__ void kd_step_dot(uint32_t n, const double *__restrict__ sc, const double *ph,
                            const double *sh, const double *s, const double *t, double *x,
                            double *r, const double *R, double *part)
{
    const bool live = sc[B_DONE] == 0.0;
    double acc[2] = {0.0, 0.0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double ri;
        if (live) {
            x[i] += sc[B_ALPHA] * ph[i] + sc[B_OMEGA] * sh[i];
            ri = s[i] - sc[B_OMEGA] * t[i];
            r[i] = ri;
        } else {
            ri = r[i];
        }
        acc[0] += ri * ri;
        acc[1] += R[i] * ri;
    }
    block_partials<2>(acc, part);
}

// BiCGSTAB (van der Vorst 1992), right Jacobi preconditioning, on J dx = -r
// with J the current Jacobian mode at V_XE and r in V_R; dx0 = 0, result in
// V_DX.  Stops at ||r_lin||^2 <= lin_rtol^2 * rn (rn = ||r||^2).  Returns the
// iteration count.  An iteration is 12 launches: each vector update is fused
// with the preconditioner, the Jacobian's pre-scaling or the dot product that
// reads it, and each reduction's last step runs the scalar recurrences.
// Operators above 64 MiB run one iteration per host check (a product costs
// far more than the check); smaller ones enqueue eight iterations per check.
static uint32_t bicgstab(Ctx &c, uint32_t max_lin, double lin_rtol, double rn)
{
    Scene &s = c.s;
    const uint32_t nb = c.nb;
    const uint32_t g = cdiv(nb, 256);
    hipStream_t stm = c.st;
    double *sc = c.sc;
    k_jdiag<<<g, 256, 0, stm>>>(c.sys, c.v[V_XE], c.v[V_TDIAG], c.v[V_LDIAG], c.v[V_DINV]);
    double *R = c.v[V_RHAT];
    IRTT_HIP(hipMemsetAsync(c.v[V_DX], 0, nb * sizeof(double), stm));
    IRTT_HIP(hipMemsetAsync(c.v[V_P], 0, nb * sizeof(double), stm));
    IRTT_HIP(hipMemsetAsync(c.v[V_V], 0, nb * sizeof(double), stm));
    double *rl = c.v[V_XT];  // linear residual, starts at -r
    k_axpy2<<<g, 256, 0, stm>>>(nb, -2.0, c.v[V_R], c.v[V_R], rl);
    IRTT_HIP(hipMemcpyAsync(R, rl, nb * sizeof(double), hipMemcpyDeviceToDevice, stm));
    kd_init<<<1, 1, 0, stm>>>(sc, lin_rtol * lin_rtol * rn);
    op_prepare(s, c.sys.transposed != 0);
    const bool tr = c.sys.transposed != 0;
    // the first rho; later ones come with each iteration's last reduction
    k_dots<<<kRedBlocks, kRedThreads, 0, stm>>>(nb, 1, R, rl, nullptr, nullptr, nullptr, nullptr,
                                                c.part);
    kd_final<<<1, kRedThreads, 0, stm>>>(c.part, 1, sc, F_RHO);
    auto iteration = [&]() {
        kd_p_pre<<<g, 256, 0, stm>>>(c.sys, sc, c.v[V_XE], rl, c.v[V_V], c.v[V_P], c.v[V_DINV],
                                     c.v[V_PH], c.v[V_W]);
        op_apply(s, c.v[V_W], c.v[V_TW], tr, false);
        kd_post_dot<1><<<kRedBlocks, kRedThreads, 0, stm>>>(c.sys, c.v[V_XE], c.v[V_PH], c.v[V_W],
                                                           c.v[V_TW], c.v[V_V], R, c.part);
        kd_final<<<1, kRedThreads, 0, stm>>>(c.part, 1, sc, F_ALPHA);
        kd_s_dot<<<kRedBlocks, kRedThreads, 0, stm>>>(nb, sc, rl, c.v[V_V], c.v[V_S], c.part);
        kd_final<<<1, kRedThreads, 0, stm>>>(c.part, 1, sc, F_HALF);
        kd_sh_pre<<<g, 256, 0, stm>>>(c.sys, sc, c.v[V_XE], c.v[V_PH], c.v[V_DX], c.v[V_DINV],
                                      c.v[V_S], c.v[V_SH], c.v[V_W]);
        op_apply(s, c.v[V_W], c.v[V_TW], tr, false);
        kd_post_dot<2><<<kRedBlocks, kRedThreads, 0, stm>>>(c.sys, c.v[V_XE], c.v[V_SH], c.v[V_W],
                                                           c.v[V_TW], c.v[V_T], c.v[V_S], c.part);
        kd_final<<<1, kRedThreads, 0, stm>>>(c.part, 2, sc, F_OMEGA);
        kd_step_dot<<<kRedBlocks, kRedThreads, 0, stm>>>(nb, sc, c.v[V_PH], c.v[V_SH], c.v[V_S],
                                                         c.v[V_T], c.v[V_DX], rl, R, c.part);
        kd_final<<<1, kRedThreads, 0, stm>>>(c.part, 2, sc, F_END_RHO);
    };
    const uint32_t K = s.op_bytes() > (64u << 20) ? 1u : 8u;
    double hs[2] = {0.0, 0.0};
    uint32_t launched = 0;
    while (launched < max_lin) {
        const uint32_t kk = std::min(K, max_lin - launched);
        for (uint32_t k = 0; k < kk; k++) iteration();
        launched += kk;
        IRTT_HIP(hipMemcpyAsync(hs, sc + B_DONE, sizeof(double), hipMemcpyDeviceToHost, stm));
        IRTT_HIP(hipMemcpyAsync(hs + 1, sc + B_ITERS, sizeof(double), hipMemcpyDeviceToHost,
                                stm));
        IRTT_HIP(hipStreamSynchronize(stm));
        if (hs[0] != 0.0) break;
    }
    return std::min<uint32_t>((uint32_t)hs[1], max_lin);
}

// ---- restarted GMRES(m) ---------------------------------------------------
namespace {
// part[i * nparts + b] = partial (V_i . w) over block b's rows, i < nv
__global__ __launch_bounds__(kRedThreads) void k_gm_dots(uint32_t n, const double *__restrict__ V,
                                                         uint32_t ld, const double *__restrict__ w,
                                                         double *__restrict__ part)
{
    __shared__ double sm[kRedThreads];
    const double *v = V + (size_t)blockIdx.y * ld;
    double a = 0.0;
    for (uint32_t i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads)
        a += v[i] * w[i];
    sm[threadIdx.x] = a;
    __syncthreads();
    for (int off = kRedThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = sm[0];
}
// out[i] = sum_b part[i * nparts + b] in block order (deterministic)
__global__ void k_gm_sums(const double *__restrict__ part, uint32_t nparts, double *__restrict__ out)
{
    __shared__ double sm[kRedThreads];
    double a = 0.0;
    for (uint32_t b = threadIdx.x; b < nparts; b += kRedThreads)
        a += part[(size_t)blockIdx.x * nparts + b];
    sm[threadIdx.x] = a;
    __syncthreads();
    for (int off = kRedThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}
// w -= sum_{i < nv} h_i V_i
__global__ void k_gm_update(uint32_t n, const double *__restrict__ V, uint32_t ld, uint32_t nv,
                            const double *__restrict__ h, double *__restrict__ w)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double a = w[r];
    for (uint32_t i = 0; i < nv; i++) a -= h[i] * V[(size_t)i * ld + r];
    w[r] = a;
}
// out = a * s
__global__ void k_gm_scale(uint32_t n, const double *__restrict__ a, double s,
                           double *__restrict__ out)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[r] = a[r] * s;
}
// out = a .* b
__global__ void k_gm_mul(uint32_t n, const double *__restrict__ a, const double *__restrict__ b,
                         double *__restrict__ out)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[r] = a[r] * b[r];
}
// dx += dinv .* sum_{i < nv} y_i V_i
__global__ void k_gm_combine(uint32_t n, const double *__restrict__ V, uint32_t ld, uint32_t nv,
                             const double *__restrict__ y, const double *__restrict__ dinv,
                             double *__restrict__ dx)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double a = 0.0;
    for (uint32_t i = 0; i < nv; i++) a += y[i] * V[(size_t)i * ld + r];
    dx[r] += dinv[r] * a;
}
}  // namespace

// GMRES(m) (Saad & Schultz 1986) on J dx = -r with right Jacobi
// preconditioning, the same system, preconditioner, start (dx = 0) and stop
// rule (||r_lin||^2 <= lin_rtol^2 rn) as bicgstab.  Arnoldi by classical
// Gram-Schmidt applied twice (one batched dot kernel and one update kernel
// per pass, instead of one pair per basis vector); the small Hessenberg
// least-squares problem (Givens rotations) runs on the host.  Returns the
// iteration count (one operator product each).
static uint32_t gmres(Ctx &c, uint32_t max_lin, double lin_rtol, double rn, uint32_t m)
{
    Scene &s = c.s;
    const uint32_t nb = c.nb, ld = c.ld, g = cdiv(nb, 256);
    hipStream_t stm = c.st;
    m = std::max<uint32_t>(1, std::min<uint32_t>(m, 200));
    DBuf<double> V, part, dev;
    V.alloc((size_t)(m + 1) * ld);
    part.alloc((size_t)(m + 1) * kRedBlocks);
    dev.alloc(2 * (size_t)(m + 1) + 1 + m);
    double *hA = dev.p, *hB = dev.p + (m + 1), *nrm = dev.p + 2 * (m + 1), *ydev = nrm + 1;
    k_jdiag<<<g, 256, 0, stm>>>(c.sys, c.v[V_XE], c.v[V_TDIAG], c.v[V_LDIAG], c.v[V_DINV]);
    IRTT_HIP(hipMemsetAsync(c.v[V_DX], 0, nb * sizeof(double), stm));
    const double goal = lin_rtol * lin_rtol * rn;
    double *w = c.v[V_T], *z = c.v[V_PH], *rl = c.v[V_XT];
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), hbuf(2 * (m + 1) + 1);
    uint32_t it = 0;
    // one pass of Gram-Schmidt of w against V_0..V_{nv-1}; coefficients into h
    auto cgs = [&](uint32_t nv, double *h) {
        k_gm_dots<<<dim3(kRedBlocks, nv), kRedThreads, 0, stm>>>(nb, V.p, ld, w, part.p);
        k_gm_sums<<<nv, kRedThreads, 0, stm>>>(part.p, kRedBlocks, h);
        k_gm_update<<<g, 256, 0, stm>>>(nb, V.p, ld, nv, h, w);
    };
    auto norm2 = [&](const double *a, double *dst) {
        k_gm_dots<<<dim3(kRedBlocks, 1), kRedThreads, 0, stm>>>(nb, a, ld, a, part.p);
        k_gm_sums<<<1, kRedThreads, 0, stm>>>(part.p, kRedBlocks, dst);
    };
    while (it < max_lin) {
        // linear residual rl = -r - J dx
        if (it == 0) {
            k_gm_scale<<<g, 256, 0, stm>>>(nb, c.v[V_R], -1.0, rl);
        } else {
            c.jv(c.v[V_DX], w);
            k_axpy2<<<g, 256, 0, stm>>>(nb, 1.0, c.v[V_R], w, rl);  // r + J dx
            k_gm_scale<<<g, 256, 0, stm>>>(nb, rl, -1.0, rl);
        }
        norm2(rl, nrm);
        double b2 = 0.0;
        IRTT_HIP(hipMemcpyAsync(&b2, nrm, sizeof(double), hipMemcpyDeviceToHost, stm));
        IRTT_HIP(hipStreamSynchronize(stm));
        if (!(b2 > goal) || !std::isfinite(b2)) break;
        const double beta = std::sqrt(b2);
        k_gm_scale<<<g, 256, 0, stm>>>(nb, rl, 1.0 / beta, V.p);
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = beta;
        uint32_t k = 0;
        bool done = false;
        for (; k < m && it < max_lin; k++) {
            k_gm_mul<<<g, 256, 0, stm>>>(nb, c.v[V_DINV], V.p + (size_t)k * ld, z);
            c.jv(z, w);
            it++;
            cgs(k + 1, hA);
            cgs(k + 1, hB);
            norm2(w, nrm);
            IRTT_HIP(hipMemcpyAsync(hbuf.data(), dev.p, (2 * (m + 1) + 1) * sizeof(double),
                                    hipMemcpyDeviceToHost, stm));
            IRTT_HIP(hipStreamSynchronize(stm));
            double *col = H.data() + (size_t)k * (m + 1);
            for (uint32_t i = 0; i <= k; i++) col[i] = hbuf[i] + hbuf[(m + 1) + i];
            const double hn = std::sqrt(std::max(hbuf[2 * (m + 1)], 0.0));
            col[k + 1] = hn;
            for (uint32_t i = 0; i < k; i++) {
                const double a = cs[i] * col[i] + sn[i] * col[i + 1];
                col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1];
                col[i] = a;
            }
            const double den = std::hypot(col[k], col[k + 1]);
            cs[k] = den > 0.0 ? col[k] / den : 1.0;
            sn[k] = den > 0.0 ? col[k + 1] / den : 0.0;
            col[k] = den;
            col[k + 1] = 0.0;
            gv[k + 1] = -sn[k] * gv[k];
            gv[k] = cs[k] * gv[k];
            const double res2 = gv[k + 1] * gv[k + 1];
            if (res2 <= goal || !(hn > 0.0) || !std::isfinite(res2)) {
                k++;
                done = true;
                break;
            }
            k_gm_scale<<<g, 256, 0, stm>>>(nb, w, 1.0 / hn, V.p + (size_t)(k + 1) * ld);
        }
        // y = H^-1 g (upper triangular, k x k), dx += M^-1 V y
        std::vector<double> y(k, 0.0);
        for (int i = (int)k - 1; i >= 0; i--) {
            double a = gv[i];
            for (uint32_t j = i + 1; j < k; j++) a -= H[(size_t)j * (m + 1) + i] * y[j];
            const double d = H[(size_t)i * (m + 1) + i];
            y[i] = d != 0.0 ? a / d : 0.0;
        }
        if (k) {
            IRTT_HIP(hipMemcpyAsync(ydev, y.data(), k * sizeof(double), hipMemcpyHostToDevice, stm));
            k_gm_combine<<<g, 256, 0, stm>>>(nb, V.p, ld, k, ydev, c.v[V_DINV], c.v[V_DX]);
        }
        if (done) break;
    }
    IRTT_HIP(hipStreamSynchronize(stm));
    (void)s;
    return std::min(it, max_lin);
}

// the Krylov method of the solver parameters
static uint32_t linsolve(Ctx &c, uint32_t max_lin, double lin_rtol, double rn)
{
    if (c.krylov == IRTT_KRYLOV_GMRES) return gmres(c, max_lin, lin_rtol, rn, c.restart);
    return bicgstab(c, max_lin, lin_rtol, rn);
}

namespace {
// r = fixed ? 0 : -rhs
__global__ void k_neg_masked(Sys s, const double *__restrict__ rhs, double *__restrict__ r)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < s.nb) r[i] = s.fixed[i] ? 0.0 : -rhs[i];
}
// r = fixed ? 0 : a - b
__global__ void k_sub_masked(Sys s, const double *__restrict__ a, const double *__restrict__ b,
                             double *__restrict__ r)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < s.nb) r[i] = s.fixed[i] ? 0.0 : a[i] - b[i];
}
}  // namespace

// Krylov tolerance: with FP16 products a linear solve cannot go much below the
// quantisation level, the outer (FP32-residual) iteration refines instead
static double krylov_rtol(const Scene &s, double lin_rtol)
{
    return op_refined(s) ? std::max(lin_rtol, 1e-6) : lin_rtol;
}

// Adjoint solve J(x)^T lambda = rhs (thermal_renderer.cpp:932-956): the same
// Krylov solver on the transposed products; fixed bins get lambda = 0.
void adjoint_solve(Scene &s, const irtt_solver_params &p, const double *x, const double *rhs,
                   double *lambda, irtt_solve_stats &st)
{
    Ctx c(s);
    FixedBuf fb;
    setup(s, c, p, fb);
    c.sys.transposed = 1;
    const uint32_t nb = s.nb;
    const uint32_t g = cdiv(nb, 256);
    hipStream_t stm = s.stream;
    IRTT_HIP(hipEventRecord(s.ev0, stm));
    IRTT_HIP(hipMemcpyAsync(c.v[V_X], x ? x : s.x.data(), nb * sizeof(double), hipMemcpyHostToDevice,
                            stm));
    k_pow4<<<g, 256, 0, stm>>>(c.sys, c.v[V_X], c.v[V_XE], c.v[V_Y]);
    IRTT_HIP(hipMemcpyAsync(c.v[V_W], rhs, nb * sizeof(double), hipMemcpyHostToDevice, stm));
    k_neg_masked<<<g, 256, 0, stm>>>(c.sys, c.v[V_W], c.v[V_R]);
    double rn;
    c.dots(1, c.v[V_R], c.v[V_R], nullptr, nullptr, nullptr, nullptr, &rn);
    st = irtt_solve_stats{};
    const uint32_t max_lin = p.max_linear ? p.max_linear : 200;
    const double lin_rtol = p.lin_rtol > 0 ? p.lin_rtol : 1e-10;
    double *res_lam = c.v[V_DX];
    if (rn > 0.0) st.linear_iters = linsolve(c, max_lin, krylov_rtol(s, lin_rtol), rn);
    else IRTT_HIP(hipMemsetAsync(c.v[V_DX], 0, nb * sizeof(double), stm));
    if (rn > 0.0 && op_refined(s)) {
        // iterative refinement: r = J^T lambda - rhs with the FP32 operator,
        // correction from the FP16 Krylov solve
        IRTT_HIP(hipMemcpyAsync(c.v[V_RHS], rhs, nb * sizeof(double), hipMemcpyHostToDevice, stm));
        IRTT_HIP(hipMemcpyAsync(c.v[V_LAM], c.v[V_DX], nb * sizeof(double),
                                hipMemcpyDeviceToDevice, stm));
        const double goal = lin_rtol * lin_rtol * rn;
        for (int it = 0; it < 20; it++) {
            c.jv(c.v[V_LAM], c.v[V_AUX], true);
            k_sub_masked<<<g, 256, 0, stm>>>(c.sys, c.v[V_AUX], c.v[V_RHS], c.v[V_R]);
            double r2;
            c.dots(1, c.v[V_R], c.v[V_R], nullptr, nullptr, nullptr, nullptr, &r2);
            st.newton_iters = it + 1;
            if (r2 <= goal || !std::isfinite(r2)) break;
            st.linear_iters += linsolve(c, max_lin, krylov_rtol(s, lin_rtol), r2);
            k_axpy2<<<g, 256, 0, stm>>>(nb, 1.0, c.v[V_LAM], c.v[V_DX], c.v[V_LAM]);
        }
        res_lam = c.v[V_LAM];
    }
    std::vector<double> lam(nb);
    IRTT_HIP(hipMemcpyAsync(lam.data(), res_lam, nb * sizeof(double), hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipEventRecord(s.ev1, stm));
    IRTT_HIP(hipStreamSynchronize(stm));
    IRTT_HIP(hipEventElapsedTime(&st.solve_ms, s.ev0, s.ev1));
    for (uint32_t i = 0; i < nb; i++) lambda[i] = s.bin_fixed[i] ? 0.0 : lam[i];
    // newton_iters counts refinement passes (FP16 operator form), 0 otherwise
    st.converged = st.linear_iters < max_lin * (st.newton_iters + 1) ? 1 : 0;
    st.residual_sq = rn;
}

void solve(Scene &s, const irtt_solver_params &p, irtt_solve_stats &st)
{
    Ctx c(s);
    FixedBuf fb;
    setup(s, c, p, fb);
    const uint32_t nb = s.nb;
    const uint32_t g = cdiv(nb, 256);
    hipStream_t stm = s.stream;
    IRTT_HIP(hipEventRecord(s.ev0, stm));

    // initial guess, thermal_renderer.cpp:544-551: fixed -> initial value,
    // zeros -> 1 K in steady state
    std::vector<double> x0(s.x);
    for (uint32_t i = 0; i < nb; i++) {
        if (s.bin_fixed[i]) x0[i] = s.bin_kelvin0[i];
        else if (p.steady && x0[i] == 0.0) x0[i] = 1.0;
    }
    IRTT_HIP(hipMemcpyAsync(c.v[V_X], x0.data(), nb * sizeof(double), hipMemcpyHostToDevice, stm));

    const uint32_t max_newton = p.max_newton ? p.max_newton : 100;
    const uint32_t max_lin = p.max_linear ? p.max_linear : 200;
    const double f_tol = p.f_tol > 0 ? p.f_tol : 1e-6;
    const double x_tol = p.x_tol > 0 ? p.x_tol : 1e-4;
    const double lin_rtol = p.lin_rtol > 0 ? p.lin_rtol : 1e-10;
    st = irtt_solve_stats{};

    double rn = c.residual(c.v[V_X]);
    if (p.steady && p.jacobian != IRTT_JAC_REFERENCE) {
        // initial guess: the steady radiative balance is linear in y = x^4, so
        // solve T dy = -r(x_g) and take x = (x_g^4 + dy)^(1/4); exact without
        // conduction, a close start otherwise.  (The reference relies on its
        // I/dt regulariser instead, thermal_solver.cpp:150.)
        c.sys.jmode = J_PREDICT;
        st.linear_iters += linsolve(c, max_lin, krylov_rtol(s, lin_rtol), rn);
        k_y_update<<<g, 256, 0, stm>>>(c.sys, c.v[V_XE], c.v[V_DX], c.v[V_X]);
        c.sys.jmode = J_EXACT;
        rn = c.residual(c.v[V_X]);
    }
    for (uint32_t it = 0; it < max_newton; it++) {
        st.newton_iters = it + 1;
        if (rn < f_tol) {
            st.converged = 1;
            break;
        }
        st.linear_iters += linsolve(c, max_lin, krylov_rtol(s, lin_rtol), rn);
        // damped line search, thermal_solver.cpp:192-203
        double a = 1.0;
        k_axpy2<<<g, 256, 0, stm>>>(nb, a, c.v[V_X], c.v[V_DX], c.v[V_XT]);
        double rn_new = c.residual(c.v[V_XT]);
        uint32_t ls = 0;
        while ((rn_new > rn || !std::isfinite(rn_new)) && ls < 20) {
            ls++;
            a *= 0.5;
            k_axpy2<<<g, 256, 0, stm>>>(nb, a, c.v[V_X], c.v[V_DX], c.v[V_XT]);
            rn_new = c.residual(c.v[V_XT]);
        }
        st.line_search_halvings += ls;
        IRTT_HIP(hipMemcpyAsync(c.v[V_X], c.v[V_XT], nb * sizeof(double), hipMemcpyDeviceToDevice,
                                stm));
        // max |a dx| over free bins
        std::vector<double> dx(nb);
        IRTT_HIP(hipMemcpyAsync(dx.data(), c.v[V_DX], nb * sizeof(double), hipMemcpyDeviceToHost, stm));
        IRTT_HIP(hipStreamSynchronize(stm));
        double mdx = 0.0;
        for (uint32_t i = 0; i < nb; i++)
            if (!s.bin_fixed[i]) mdx = std::max(mdx, std::fabs(a * dx[i]));
        st.max_dx = mdx;
        rn = rn_new;
        if (!std::isfinite(rn)) break;
        if (rn < f_tol || mdx < x_tol || a < 1e-6) {
            st.converged = (rn < f_tol || mdx < x_tol) ? 1 : 0;
            break;
        }
    }
    st.residual_sq = rn;
    // the global power balance at the solution, FP64 on the device (V_TY,
    // V_Y and V_XE are those of the last residual, i.e. of the final x)
    {
        double *part = s.d_red.p;
        double *out = part + 2 * kRedBlocks;
        k_balance<<<kRedBlocks, kRedThreads, 0, stm>>>(c.sys, c.v[V_TY], c.v[V_Y], c.v[V_XE],
                                                       part);
        k_final_sum<<<1, kRedThreads, 0, stm>>>(part, 2, kRedBlocks, out);
        double pb[2];
        IRTT_HIP(hipMemcpyAsync(pb, out, sizeof(pb), hipMemcpyDeviceToHost, stm));
        IRTT_HIP(hipStreamSynchronize(stm));
        st.power_in = pb[0];
        st.power_out = pb[1];
        st.balance_rel = pb[0] != 0.0 ? std::fabs(pb[0] - pb[1]) / std::fabs(pb[0]) : 0.0;
    }
    std::vector<double> xr(nb);
    // V_XE holds the final x with fixed values applied (set by the last residual call)
    IRTT_HIP(hipMemcpyAsync(xr.data(), c.v[V_XE], nb * sizeof(double), hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipEventRecord(s.ev1, stm));
    IRTT_HIP(hipStreamSynchronize(stm));
    IRTT_HIP(hipEventElapsedTime(&st.solve_ms, s.ev0, s.ev1));
    s.x = xr;
}

}  // namespace irtt
