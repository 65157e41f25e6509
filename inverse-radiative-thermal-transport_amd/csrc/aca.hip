// aca.hip -- adaptive cross approximation (ACA) of the transport operator on
// the device, and the low-rank operator products used by the solver.
//
// Replaces ThermalTransport::compressTransportMatrix and its three variants
// (src/implementations/thermal_renderer/thermal_transport.cpp:281-597) plus
// collumnSumCorrection (:309-358), which run single-threaded on the host with
// Eigen, and the ACAMatrix product dst += D U (V x) + (A + C) x
// (thermal_aca_matrix.hpp:100-141).
//
// The compressed matrix is the off-diagonal part of the operator as the
// reference compresses it (before the emission diagonal is added,
// thermal_transport.cpp:634-660): T_off(r, c) = -acc(r, c) Q rounded to FP32,
// the same values as the dense FP32 operator off the diagonal.  The emission
// diagonal sigma eps A goes into C.  Factors are built in FP64 with every
// residual sum taken in the reference's order (l = 0 .. k-1, one rounding per
// product and per sum), so the pivots and factors are reproducible and equal
// to a sequential FP64 restatement bit for bit.
//
// Storage: Ut (k x ld) holds the columns of U, V (k x ld) the rows of V, both
// FP64 while building; the solver reads FP32 or FP16 copies (FP16 with one
// FP32 scale per stored row).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "irtt_internal.hpp"

namespace irtt {
namespace {

constexpr int kAcaThreads = 256;
constexpr int kAcaBlocks = 512;   // argmax partials per vector kernel

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// T_off(r, c): the FP32 operator off the diagonal, od[r] on it
__device__ __forceinline__ double toff(const float *__restrict__ T, uint32_t ld,
                                       const float *__restrict__ od, uint32_t r, uint32_t c)
{
    return r == c ? (double)od[r] : (double)T[(size_t)r * ld + c];
}

// argmax |v| over allowed indices, ties to the smallest index; only strictly
// positive magnitudes qualify and the default is index 0
// (findArgmaxNotInList, thermal_transport.cpp:463-474)
struct Best {
    double v;
    unsigned long long i;
};

__device__ __forceinline__ void best_merge(Best &a, double v, unsigned long long i)
{
    if (v > a.v || (v == a.v && v > 0.0 && i < a.i)) {
        a.v = v;
        a.i = i;
    }
}

__device__ __forceinline__ void block_best(Best b, Best *out)
{
    __shared__ double sv[kAcaThreads];
    __shared__ unsigned long long si[kAcaThreads];
    sv[threadIdx.x] = b.v;
    si[threadIdx.x] = b.i;
    __syncthreads();
    for (int off = kAcaThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            Best a{sv[threadIdx.x], si[threadIdx.x]};
            best_merge(a, sv[threadIdx.x + off], si[threadIdx.x + off]);
            sv[threadIdx.x] = a.v;
            si[threadIdx.x] = a.i;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = Best{sv[0], si[0]};
}

// off-diagonal part of the operator on the diagonal: FP32(0 - acc_rr Q)
__global__ void k_offdiag(const unsigned long long *__restrict__ acc, uint32_t nb, uint32_t ld,
                          double Q, float *__restrict__ od)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nb) od[r] = (float)(0.0 - (double)acc[(size_t)r * ld + r] * Q);
}

// residual row (calculateResidualRow, thermal_transport.cpp:476-487):
// out[c] = T_off(i, c) - sum_{l<k} U(i, l) V(l, c); argmax over columns with
// excl[c] == 0
__global__ __launch_bounds__(kAcaThreads) void k_res_row(const float *__restrict__ T, uint32_t ld,
                                                         const float *__restrict__ od, uint32_t n,
                                                         const double *__restrict__ Ut,
                                                         const double *__restrict__ V, uint32_t k,
                                                         const int32_t *__restrict__ ip,
                                                         const uint8_t *__restrict__ excl,
                                                         double *__restrict__ out,
                                                         Best *__restrict__ part)
{
    __shared__ double ui[kAcaMaxRank];
    const uint32_t i = (uint32_t)*ip;
    for (uint32_t l = threadIdx.x; l < k; l += blockDim.x) ui[l] = Ut[(size_t)l * ld + i];
    __syncthreads();
    Best b{0.0, 0ull};
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        double a = 0.0;
        for (uint32_t l = 0; l < k; l++) a = a + ui[l] * V[(size_t)l * ld + c];
        double val = toff(T, ld, od, i, c) - a;
        out[c] = val;
        if (!excl[c]) best_merge(b, fabs(val), c);
    }
    block_best(b, part + blockIdx.x);
}

// residual column (calculateResidualCol, thermal_transport.cpp:489-500):
// out[r] = T_off(r, j) - sum_{l<k} V(l, j) U(r, l); argmax over rows with
// excl[r] == 0
__global__ __launch_bounds__(kAcaThreads) void k_res_col(const float *__restrict__ T, uint32_t ld,
                                                         const float *__restrict__ od, uint32_t n,
                                                         const double *__restrict__ Ut,
                                                         const double *__restrict__ V, uint32_t k,
                                                         const int32_t *__restrict__ jp,
                                                         const uint8_t *__restrict__ excl,
                                                         double *__restrict__ out,
                                                         Best *__restrict__ part)
{
    __shared__ double vj[kAcaMaxRank];
    const uint32_t j = (uint32_t)*jp;
    for (uint32_t l = threadIdx.x; l < k; l += blockDim.x) vj[l] = V[(size_t)l * ld + j];
    __syncthreads();
    Best b{0.0, 0ull};
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        double a = 0.0;
        for (uint32_t l = 0; l < k; l++) a = a + vj[l] * Ut[(size_t)l * ld + r];
        double val = toff(T, ld, od, r, j) - a;
        out[r] = val;
        if (!excl[r]) best_merge(b, fabs(val), r);
    }
    block_best(b, part + blockIdx.x);
}

// reduces the partials of one vector kernel: *idx = argmax; excl[idx] = 1 if
// mark; *mag = |vec[idx]| (the reference reads the entry at the returned
// index, also when no entry qualified and the index is 0)
__global__ void k_pick(const Best *__restrict__ part, int nparts, int32_t *__restrict__ idx,
                       uint8_t *__restrict__ excl, int mark, const double *__restrict__ vec,
                       double *__restrict__ mag)
{
    Best b{0.0, 0ull};
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) best_merge(b, part[p].v, part[p].i);
    __shared__ Best r;
    block_best(b, &r);
    if (threadIdx.x == 0) {
        *idx = (int32_t)r.i;
        if (mark) excl[r.i] = 1;
        if (mag) *mag = fabs(vec[r.i]);
    }
}

// Ut[k] = u / (unorm ? u[*ip] : 1), V[k] = v / (vnorm ? v[*jp] : 1), a zero
// pivot leaving the vector unscaled (thermal_transport.cpp:421-426,588-592)
__global__ void k_commit(uint32_t n, uint32_t ld, const double *__restrict__ u,
                         const double *__restrict__ v, const int32_t *__restrict__ ip,
                         const int32_t *__restrict__ jp, int unorm, int vnorm,
                         double *__restrict__ ut_k, double *__restrict__ v_k)
{
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double du = unorm ? u[*ip] : 0.0, dv = vnorm ? v[*jp] : 0.0;
    ut_k[c] = du != 0.0 ? u[c] / du : u[c];
    v_k[c] = dv != 0.0 ? v[c] / dv : v[c];
}

// full pivoting (simpleAdaptiveCrossApproximation, thermal_transport.cpp:360-386):
// A += u v^T (when upd), then the argmax of |T_off - A| over the whole matrix,
// row-major index, ties to the first in row-major order
__global__ __launch_bounds__(kAcaThreads) void k_full_update(const float *__restrict__ T,
                                                             uint32_t ld,
                                                             const float *__restrict__ od,
                                                             uint32_t n, double *__restrict__ A,
                                                             const double *__restrict__ u,
                                                             const double *__restrict__ v, int upd,
                                                             Best *__restrict__ part)
{
    Best b{0.0, 0ull};
    for (uint32_t r = blockIdx.x; r < n; r += gridDim.x) {
        const double ur = upd ? u[r] : 0.0;
        double *Ar = A + (size_t)r * ld;
        for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
            double a = Ar[c];
            if (upd) {
                a = a + ur * v[c];
                Ar[c] = a;
            }
            double R = toff(T, ld, od, r, c) - a;
            best_merge(b, fabs(R), (unsigned long long)r * n + c);
        }
    }
    block_best(b, part + blockIdx.x);
}

// full pivoting step: Ut[k] = R(:, pc) / R(pr, pc), V[k] = R(pr, :) with
// R = T_off - A; pivot from the packed row-major index
__global__ void k_full_extract(const float *__restrict__ T, uint32_t ld,
                               const float *__restrict__ od, uint32_t n,
                               const double *__restrict__ A, const Best *__restrict__ pick,
                               double *__restrict__ ut_k, double *__restrict__ v_k,
                               int32_t *__restrict__ prow, int32_t *__restrict__ pcol)
{
    const unsigned long long p = pick->i;
    const uint32_t pr = (uint32_t)(p / n), pc = (uint32_t)(p % n);
    const double piv = toff(T, ld, od, pr, pc) - A[(size_t)pr * ld + pc];
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        *prow = (int32_t)pr;
        *pcol = (int32_t)pc;
    }
    if (t >= n) return;
    double col = toff(T, ld, od, t, pc) - A[(size_t)t * ld + pc];
    ut_k[t] = piv != 0.0 ? col / piv : col;
    v_k[t] = toff(T, ld, od, pr, t) - A[(size_t)pr * ld + t];
}

__global__ void k_pick_full(const Best *__restrict__ part, int nparts, Best *__restrict__ out)
{
    Best b{0.0, 0ull};
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) best_merge(b, part[p].v, part[p].i);
    block_best(b, out);
}

// Low-rank product statistics over 64 x 64 tiles: UV(r, c) = sum_l Ut[l,r] V[l,c]
// in l order, e = T_off - UV.  stats per block: max|UV|, max|e|, sum e^2,
// max|T_off|, sum T_off^2 (compressTransportMatrix, thermal_transport.cpp:301-306;
// collumnSumCorrection's maximum, :322-323).
constexpr int kTile = 64;
__global__ __launch_bounds__(256) void k_lowrank_stats(const float *__restrict__ T, uint32_t ld,
                                                       const float *__restrict__ od, uint32_t n,
                                                       const double *__restrict__ Ut,
                                                       const double *__restrict__ V, uint32_t k,
                                                       double *__restrict__ stats)
{
    __shared__ double su[kTile], sv[kTile];
    __shared__ double red[5][256];
    const uint32_t r0 = blockIdx.y * kTile, c0 = blockIdx.x * kTile;
    const uint32_t tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // 16 x 16 threads, 4 x 4 each
    double acc[4][4] = {};
    for (uint32_t l = 0; l < k; l++) {
        __syncthreads();
        if (threadIdx.x < kTile) {
            uint32_t r = r0 + threadIdx.x;
            su[threadIdx.x] = r < n ? Ut[(size_t)l * ld + r] : 0.0;
        } else if (threadIdx.x < 2 * kTile) {
            uint32_t c = c0 + threadIdx.x - kTile;
            sv[threadIdx.x - kTile] = c < n ? V[(size_t)l * ld + c] : 0.0;
        }
        __syncthreads();
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) acc[a][b] = acc[a][b] + su[ty + 16 * a] * sv[tx + 16 * b];
    }
    double muv = 0.0, me = 0.0, se = 0.0, mt = 0.0, st = 0.0;
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
            uint32_t r = r0 + ty + 16 * a, c = c0 + tx + 16 * b;
            if (r >= n || c >= n) continue;
            double t = toff(T, ld, od, r, c), uv = acc[a][b], e = t - uv;
            muv = fmax(muv, fabs(uv));
            me = fmax(me, fabs(e));
            se += e * e;
            mt = fmax(mt, fabs(t));
            st += t * t;
        }
    red[0][threadIdx.x] = muv;
    red[1][threadIdx.x] = me;
    red[2][threadIdx.x] = se;
    red[3][threadIdx.x] = mt;
    red[4][threadIdx.x] = st;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + off]);
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + off]);
            red[2][threadIdx.x] += red[2][threadIdx.x + off];
            red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + off]);
            red[4][threadIdx.x] += red[4][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double *o = stats + 5 * ((size_t)blockIdx.y * gridDim.x + blockIdx.x);
        for (int q = 0; q < 5; q++) o[q] = red[q][0];
    }
}

// FP64 factors -> stored operator factors: FP32, or FP16 with one FP32 scale
// per stored row (scale = max|row| / 2^14, so every entry is a normal FP16)
__global__ void k_to_f32(const double *__restrict__ a, size_t n, float *__restrict__ o)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = (float)a[i];
}

__global__ void k_row_absmax(const double *__restrict__ a, uint32_t n, uint32_t ld,
                             float *__restrict__ scale)
{
    __shared__ double sm[kAcaThreads];
    const double *row = a + (size_t)blockIdx.x * ld;
    double m = 0.0;
    for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) m = fmax(m, fabs(row[c]));
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int off = kAcaThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) scale[blockIdx.x] = (float)(sm[0] * 0x1p-14);
}

__global__ void k_to_f16(const double *__restrict__ a, uint32_t rows, uint32_t ld,
                         const float *__restrict__ scale, __half *__restrict__ o)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * ld) return;
    float s = scale[i / ld];
    o[i] = __float2half(s > 0.0f ? (float)(a[i] / (double)s) : 0.0f);
}

// ---- low-rank products for the solver ----------------------------------
// t[l] = sum_c A[l, c] w[c] over the stored factor rows (one block per l,
// fixed-order reduction)
template <typename E>
__device__ __forceinline__ double ld_elem(const E *p, size_t i);
template <>
__device__ __forceinline__ double ld_elem<float>(const float *p, size_t i)
{
    return (double)p[i];
}
template <>
__device__ __forceinline__ double ld_elem<__half>(const __half *p, size_t i)
{
    return (double)__half2float(p[i]);
}

template <typename E>
__global__ __launch_bounds__(kAcaThreads) void k_lr_dots(const E *__restrict__ A,
                                                         const float *__restrict__ scale,
                                                         uint32_t n, uint32_t ld,
                                                         const double *__restrict__ w,
                                                         double *__restrict__ t)
{
    __shared__ double sm[kAcaThreads];
    const uint32_t l = blockIdx.x;
    const E *row = A + (size_t)l * ld;
    double a = 0.0;
    for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) a += ld_elem(row, c) * w[c];
    sm[threadIdx.x] = a;
    __syncthreads();
    for (int off = kAcaThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) t[l] = sm[0] * (scale ? (double)scale[l] : 1.0);
}

// y[c] = sum_l B[l, c] t[l]
template <typename E>
__global__ __launch_bounds__(kAcaThreads) void k_lr_expand(const E *__restrict__ B,
                                                           const float *__restrict__ scale,
                                                           uint32_t n, uint32_t ld, uint32_t k,
                                                           const double *__restrict__ t,
                                                           double *__restrict__ y)
{
    __shared__ double st[kAcaMaxRank];
    for (uint32_t l = threadIdx.x; l < k; l += blockDim.x)
        st[l] = t[l] * (scale ? (double)scale[l] : 1.0);
    __syncthreads();
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ld) return;
    double a = 0.0;
    if (c < n)
        for (uint32_t l = 0; l < k; l++) a += ld_elem(B, (size_t)l * ld + c) * st[l];
    y[c] = a;
}

}  // namespace

// draws of the ACA+ reference rows and columns: the MSVC C runtime's rand()
// (x = 214013 x + 2531011, output bits 16..30), the runtime the reference
// targets, mapped as (double)rand() / RAND_MAX * (n - 1)
// (resetReferenceRow / resetReferenceCol, thermal_transport.cpp:502-520)
struct CrtRand {
    uint32_t x;
    int next()
    {
        x = x * 214013u + 2531011u;
        return (int)((x >> 16) & 0x7fff);
    }
    uint32_t index(uint32_t n) { return (uint32_t)((double)next() / 32767.0 * (double)(n - 1)); }
};

static uint32_t reset_reference(CrtRand &rng, uint32_t n, const std::vector<uint8_t> &used)
{
    uint32_t r = 0;
    for (uint32_t t = 1; t < n; t++) {
        r = rng.index(n);
        if (!used[r]) break;
    }
    return r;
}

void aca_compress(Scene &s, const irtt_aca_params &p, irtt_aca_stats &st)
{
    if (!s.has_transport) throw Error(IRTT_ERR_STATE, "transport operator not built");
    if (s.layout != IRTT_LAYOUT_DENSE)
        throw Error(IRTT_ERR_STATE, "global ACA compresses the dense transport layout");
    const uint32_t n = s.nb, ld = s.ld;
    const uint32_t k = p.rank ? p.rank : 50;
    if (k > kAcaMaxRank || k > n)
        throw Error(IRTT_ERR_ARG, "ACA rank must be in [1, min(n_bins, " +
                                      std::to_string(kAcaMaxRank) + ")]");
    if (p.algorithm > IRTT_ACA_PLUS) throw Error(IRTT_ERR_ARG, "unknown ACA algorithm");
    if (p.precision > IRTT_STORE_F16) throw Error(IRTT_ERR_ARG, "unknown storage precision");
    hipStream_t stm = s.stream;
    Aca &a = s.aca;
    a.valid = false;
    a.k = k;
    a.d_ut.alloc((size_t)k * ld);
    a.d_v.alloc((size_t)k * ld);
    a.d_od.alloc(n);
    a.d_u.alloc(ld);
    a.d_w.alloc(ld);
    a.d_part.alloc(sizeof(Best) * kAcaBlocks / sizeof(double) + 1);
    a.d_piv.alloc(2 * (size_t)k + 2);
    a.d_excl.alloc(2 * (size_t)ld);
    a.d_mag.alloc(2);
    IRTT_HIP(hipMemsetAsync(a.d_ut.p, 0, a.d_ut.bytes(), stm));
    IRTT_HIP(hipMemsetAsync(a.d_v.p, 0, a.d_v.bytes(), stm));
    IRTT_HIP(hipMemsetAsync(a.d_piv.p, 0, a.d_piv.bytes(), stm));
    IRTT_HIP(hipMemsetAsync(a.d_excl.p, 0, a.d_excl.bytes(), stm));
    IRTT_HIP(hipEventRecord(s.ev0, stm));
    k_offdiag<<<cdiv(n, 256), 256, 0, stm>>>(s.d_acc.p, n, ld, s.quantum, a.d_od.p);

    Best *part = reinterpret_cast<Best *>(a.d_part.p);
    int32_t *prow = a.d_piv.p, *pcol = a.d_piv.p + k + 1;
    uint8_t *xrow = a.d_excl.p, *xcol = a.d_excl.p + ld;
    const uint32_t gv = std::min<uint32_t>(kAcaBlocks, cdiv(n, kAcaThreads));
    const uint32_t gc = cdiv(n, 256);
    const float *T = s.d_T.p;
    const float *od = a.d_od.p;
    double *Ut = a.d_ut.p, *V = a.d_v.p, *uu = a.d_u.p, *vv = a.d_w.p;

    if (p.algorithm == IRTT_ACA_PARTIAL) {
        // completeAdaptiveCrossApproximation (thermal_transport.cpp:388-461):
        // start at row 0; V rows are normalised by the pivot, U columns are not;
        // the second pivot row is not entered into the excluded set
        // (only the rows chosen inside the loop are)
        IRTT_HIP(hipMemsetAsync(xrow, 1, 1, stm));
        for (uint32_t kk = 0; kk < k; kk++) {
            k_res_row<<<gv, kAcaThreads, 0, stm>>>(T, ld, od, n, Ut, V, kk, prow + kk, xcol, vv,
                                                   part);
            k_pick<<<1, kAcaThreads, 0, stm>>>(part, gv, pcol + kk, xcol, 1, vv, nullptr);
            k_res_col<<<gv, kAcaThreads, 0, stm>>>(T, ld, od, n, Ut, V, kk, pcol + kk, xrow, uu,
                                                   part);
            k_pick<<<1, kAcaThreads, 0, stm>>>(part, gv, prow + kk + 1, xrow, kk > 0 ? 1 : 0,
                                               uu, nullptr);
            k_commit<<<gc, 256, 0, stm>>>(n, ld, uu, vv, prow + kk, pcol + kk, 0, 1,
                                          Ut + (size_t)kk * ld, V + (size_t)kk * ld);
        }
    } else if (p.algorithm == IRTT_ACA_FULL) {
        size_t need = (size_t)n * ld * sizeof(double);
        const size_t fr = dev_available();
        if (need + (64u << 20) > fr)
            throw Error(IRTT_ERR_NOMEM, "full-pivot ACA needs an n_bins^2 FP64 work matrix (" +
                                            std::to_string(need) + " bytes)");
        DBuf<double> A;
        A.alloc((size_t)n * ld);
        IRTT_HIP(hipMemsetAsync(A.p, 0, A.bytes(), stm));
        DBuf<Best> pick;
        pick.alloc(1);
        const uint32_t gf = std::min<uint32_t>(2048, n);
        DBuf<Best> fpart;
        fpart.alloc(gf);
        for (uint32_t kk = 0; kk < k; kk++) {
            const double *up = kk ? Ut + (size_t)(kk - 1) * ld : Ut;
            const double *vp = kk ? V + (size_t)(kk - 1) * ld : V;
            k_full_update<<<gf, kAcaThreads, 0, stm>>>(T, ld, od, n, A.p, up, vp, kk ? 1 : 0,
                                                       fpart.p);
            k_pick_full<<<1, kAcaThreads, 0, stm>>>(fpart.p, gf, pick.p);
            k_full_extract<<<gc, 256, 0, stm>>>(T, ld, od, n, A.p, pick.p, Ut + (size_t)kk * ld,
                                                V + (size_t)kk * ld, prow + kk, pcol + kk);
        }
        IRTT_HIP(hipStreamSynchronize(stm));
    } else {
        // adaptiveCrossApproximationPlus (thermal_transport.cpp:522-597), host
        // driven: the branch on the two pivot magnitudes needs them on the host
        CrtRand rng{p.seed ? p.seed : 1u};
        std::vector<uint8_t> hrow(n, 0), hcol(n, 0);
        uint32_t r = reset_reference(rng, n, hrow);
        uint32_t c = reset_reference(rng, n, hrow);  // the reference passes the row set here
        DBuf<int32_t> dref;
        dref.alloc(4);
        int32_t h4[4];
        double mags[2];
        for (uint32_t kk = 0; kk < k; kk++) {
            h4[0] = (int32_t)r;
            h4[1] = (int32_t)c;
            IRTT_HIP(hipMemcpyAsync(dref.p, h4, 2 * sizeof(int32_t), hipMemcpyHostToDevice, stm));
            k_res_row<<<gv, kAcaThreads, 0, stm>>>(T, ld, od, n, Ut, V, kk, dref.p, xcol, vv,
                                                   part);
            k_pick<<<1, kAcaThreads, 0, stm>>>(part, gv, pcol + kk, xcol, 0, vv, a.d_mag.p + 1);
            k_res_col<<<gv, kAcaThreads, 0, stm>>>(T, ld, od, n, Ut, V, kk, dref.p + 1, xrow, uu,
                                                   part);
            k_pick<<<1, kAcaThreads, 0, stm>>>(part, gv, prow + kk, xrow, 0, uu, a.d_mag.p);
            IRTT_HIP(hipMemcpyAsync(mags, a.d_mag.p, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                    stm));
            IRTT_HIP(hipStreamSynchronize(stm));
            if (mags[0] > mags[1]) {
                k_res_row<<<gv, kAcaThreads, 0, stm>>>(T, ld, od, n, Ut, V, kk, prow + kk, xcol,
                                                       vv, part);
                k_pick<<<1, kAcaThreads, 0, stm>>>(part, gv, pcol + kk, xcol, 0, vv, nullptr);
                k_res_col<<<gv, kAcaThreads, 0, stm>>>(T, ld, od, n, Ut, V, kk, pcol + kk, xrow,
                                                       uu, part);
            } else {
                k_res_col<<<gv, kAcaThreads, 0, stm>>>(T, ld, od, n, Ut, V, kk, pcol + kk, xrow,
                                                       uu, part);
                k_pick<<<1, kAcaThreads, 0, stm>>>(part, gv, prow + kk, xrow, 0, uu, nullptr);
                k_res_row<<<gv, kAcaThreads, 0, stm>>>(T, ld, od, n, Ut, V, kk, prow + kk, xcol,
                                                       vv, part);
            }
            k_commit<<<gc, 256, 0, stm>>>(n, ld, uu, vv, prow + kk, pcol + kk, 1, 0,
                                          Ut + (size_t)kk * ld, V + (size_t)kk * ld);
            int32_t ij[2];
            IRTT_HIP(hipMemcpyAsync(&ij[0], prow + kk, sizeof(int32_t), hipMemcpyDeviceToHost, stm));
            IRTT_HIP(hipMemcpyAsync(&ij[1], pcol + kk, sizeof(int32_t), hipMemcpyDeviceToHost, stm));
            IRTT_HIP(hipStreamSynchronize(stm));
            hrow[ij[0]] = 1;
            hcol[ij[1]] = 1;
            uint8_t one = 1;
            IRTT_HIP(hipMemcpyAsync(xrow + ij[0], &one, 1, hipMemcpyHostToDevice, stm));
            IRTT_HIP(hipMemcpyAsync(xcol + ij[1], &one, 1, hipMemcpyHostToDevice, stm));
            if ((uint32_t)ij[0] == r) r = reset_reference(rng, n, hrow);
            else if ((uint32_t)ij[1] == c) c = reset_reference(rng, n, hcol);
            IRTT_HIP(hipStreamSynchronize(stm));
        }
    }
    IRTT_HIP(hipGetLastError());
    IRTT_HIP(hipEventRecord(s.ev1, stm));

    // pivots and the low-rank statistics
    a.prow.assign(k, 0);
    a.pcol.assign(k, 0);
    IRTT_HIP(hipMemcpyAsync(a.prow.data(), prow, k * sizeof(int32_t), hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipMemcpyAsync(a.pcol.data(), pcol, k * sizeof(int32_t), hipMemcpyDeviceToHost, stm));
    const uint32_t nt = cdiv(n, kTile);
    DBuf<double> dstats;
    dstats.alloc((size_t)5 * nt * nt);
    k_lowrank_stats<<<dim3(nt, nt), 256, 0, stm>>>(T, ld, od, n, Ut, V, k, dstats.p);
    std::vector<double> hst((size_t)5 * nt * nt);
    IRTT_HIP(hipMemcpyAsync(hst.data(), dstats.p, dstats.bytes(), hipMemcpyDeviceToHost, stm));
    // U column sums in row order for the column-sum correction (Ut is k x ld)
    std::vector<double> hut((size_t)k * ld), hv((size_t)k * ld);
    IRTT_HIP(hipMemcpyAsync(hut.data(), Ut, hut.size() * sizeof(double), hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipMemcpyAsync(hv.data(), V, hv.size() * sizeof(double), hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipStreamSynchronize(stm));
    IRTT_HIP(hipEventElapsedTime(&st.compress_ms, s.ev0, s.ev1));
    double muv = 0.0, me = 0.0, se = 0.0, mt = 0.0, sT = 0.0;
    for (size_t b = 0; b < (size_t)nt * nt; b++) {
        const double *o = &hst[5 * b];
        muv = std::max(muv, o[0]);
        me = std::max(me, o[1]);
        se += o[2];
        mt = std::max(mt, o[3]);
        sT += o[4];
    }
    st.err_inf = me;
    st.err_frob = std::sqrt(se);
    st.err_inf_rel = mt > 0 ? me / mt : 0.0;
    st.err_frob_rel = sT > 0 ? std::sqrt(se) / std::sqrt(sT) : 0.0;
    st.rank = k;

    // C = diag(sigma eps A) (+ column-sum correction, thermal_transport.cpp:309-358)
    a.cdiag.assign(n, 0.0);
    for (uint32_t i = 0; i < n; i++) a.cdiag[i] = kSigmaD * s.bin_eps_area[i];
    st.n_corrected = 0;
    if (p.column_sum_correction) {
        std::vector<double> us(k, 0.0);
        for (uint32_t l = 0; l < k; l++) {
            double acc = 0.0;
            for (uint32_t r = 0; r < n; r++) acc = acc + hut[(size_t)l * ld + r];
            us[l] = acc;
        }
        for (uint32_t c = 0; c < n; c++) {
            double cs = 0.0;
            for (uint32_t l = 0; l < k; l++) cs = cs + us[l] * hv[(size_t)l * ld + c];
            cs = cs + a.cdiag[c];
            if (cs < 0.5 * muv) {
                a.cdiag[c] += 0.5 * muv - cs;
                st.n_corrected++;
            }
        }
    }
    a.max_uv = muv;
    // diagonal of U V + C for the Jacobi preconditioner
    a.opdiag.assign(n, 0.0);
    for (uint32_t i = 0; i < n; i++) {
        double d = 0.0;
        for (uint32_t l = 0; l < k; l++) d = d + hut[(size_t)l * ld + i] * hv[(size_t)l * ld + i];
        a.opdiag[i] = d + a.cdiag[i];
    }

    // stored factors
    a.precision = p.precision;
    const size_t kn = (size_t)k * ld;
    if (p.precision == IRTT_STORE_F16) {
        a.d_uh.alloc(kn);
        a.d_vh.alloc(kn);
        a.d_us.alloc(k);
        a.d_vs.alloc(k);
        k_row_absmax<<<k, kAcaThreads, 0, stm>>>(Ut, n, ld, a.d_us.p);
        k_row_absmax<<<k, kAcaThreads, 0, stm>>>(V, n, ld, a.d_vs.p);
        k_to_f16<<<cdiv(kn, 256), 256, 0, stm>>>(Ut, k, ld, a.d_us.p,
                                                 reinterpret_cast<__half *>(a.d_uh.p));
        k_to_f16<<<cdiv(kn, 256), 256, 0, stm>>>(V, k, ld, a.d_vs.p,
                                                 reinterpret_cast<__half *>(a.d_vh.p));
        a.d_uf.release();
        a.d_vf.release();
    } else {
        a.d_uf.alloc(kn);
        a.d_vf.alloc(kn);
        k_to_f32<<<cdiv(kn, 256), 256, 0, stm>>>(Ut, kn, a.d_uf.p);
        k_to_f32<<<cdiv(kn, 256), 256, 0, stm>>>(V, kn, a.d_vf.p);
        a.d_uh.release();
        a.d_vh.release();
    }
    a.d_t.alloc(kAcaMaxRank);
    IRTT_HIP(hipGetLastError());
    IRTT_HIP(hipStreamSynchronize(stm));
    a.valid = true;
    s.op_kind = IRTT_OP_ACA;
}

// y = (U V) w, or (U V)^T w when transposed, with the stored factors
void aca_apply(Scene &s, const double *w, double *y, bool transposed)
{
    Aca &a = s.aca;
    const uint32_t n = s.nb, ld = s.ld, k = a.k;
    hipStream_t stm = s.stream;
    const uint32_t ge = cdiv(ld, kAcaThreads);
    if (a.precision == IRTT_STORE_F16) {
        const __half *uh = reinterpret_cast<const __half *>(a.d_uh.p);
        const __half *vh = reinterpret_cast<const __half *>(a.d_vh.p);
        const __half *first = transposed ? uh : vh;
        const __half *second = transposed ? vh : uh;
        const float *s1 = transposed ? a.d_us.p : a.d_vs.p;
        const float *s2 = transposed ? a.d_vs.p : a.d_us.p;
        k_lr_dots<__half><<<k, kAcaThreads, 0, stm>>>(first, s1, n, ld, w, a.d_t.p);
        k_lr_expand<__half><<<ge, kAcaThreads, 0, stm>>>(second, s2, n, ld, k, a.d_t.p, y);
    } else {
        const float *first = transposed ? a.d_uf.p : a.d_vf.p;
        const float *second = transposed ? a.d_vf.p : a.d_uf.p;
        k_lr_dots<float><<<k, kAcaThreads, 0, stm>>>(first, nullptr, n, ld, w, a.d_t.p);
        k_lr_expand<float><<<ge, kAcaThreads, 0, stm>>>(second, nullptr, n, ld, k, a.d_t.p, y);
    }
}

void aca_get(Scene &s, double *U, double *V, double *cdiag, int32_t *prow, int32_t *pcol)
{
    Aca &a = s.aca;
    if (!a.valid) throw Error(IRTT_ERR_STATE, "no ACA operator");
    const uint32_t n = s.nb, ld = s.ld, k = a.k;
    hipStream_t stm = s.stream;
    if (U) {
        // U is n x k row-major for the caller; the device keeps its columns
        std::vector<double> h((size_t)k * ld);
        IRTT_HIP(hipMemcpyAsync(h.data(), a.d_ut.p, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                                stm));
        IRTT_HIP(hipStreamSynchronize(stm));
        for (uint32_t r = 0; r < n; r++)
            for (uint32_t l = 0; l < k; l++) U[(size_t)r * k + l] = h[(size_t)l * ld + r];
    }
    if (V)
        IRTT_HIP(hipMemcpy2DAsync(V, (size_t)n * sizeof(double), a.d_v.p, (size_t)ld * sizeof(double),
                                  (size_t)n * sizeof(double), k, hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipStreamSynchronize(stm));
    if (cdiag) std::copy(a.cdiag.begin(), a.cdiag.end(), cdiag);
    if (prow) std::copy(a.prow.begin(), a.prow.end(), prow);
    if (pcol) std::copy(a.pcol.begin(), a.pcol.end(), pcol);
}

}  // namespace irtt
