// operator.hip -- products with the radiative operator in its stored forms.
//
// The reference multiplies with a host Eigen matrix (thermal_solver.cpp:20-23,
// 161-174) or with its ACAMatrix (thermal_aca_matrix.hpp:100-141).  Here the
// operator stays in HBM in one of three forms:
//   IRTT_OP_DENSE      FP32 T (row-major, pitch ld), products accumulate in FP64
//   IRTT_OP_DENSE_F16  T off the diagonal in FP16 with one FP32 scale per row
//                      (scale = max |row| / 2^14, every entry a normal FP16 down
//                      to 2^-28 of the row maximum), the diagonal in FP64
//   IRTT_OP_ACA        U V in FP32 or FP16 + the diagonal C (aca.hip)
// op_apply computes P w, where the operator is P + diag(cd) with cd from
// op_diag_*; the dense FP32 form has its diagonal inside P.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "irtt_internal.hpp"

namespace irtt {
namespace {

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Rows per wave in the row-major products: the wave reads each input-vector
// chunk once for kGemvRows rows, so the vector's cache traffic is a quarter
// of the matrix's.
#ifndef IRTT_GEMV_ROWS
#define IRTT_GEMV_ROWS 4
#endif
constexpr uint32_t kGemvRows = IRTT_GEMV_ROWS;
// Loads issued before any of their products are summed: k steps of a wave in
// the row-major products, rows of a thread in the column products (FP32 and
// FP16 forms).  The sums keep their order, so results do not depend on these
// counts.  Measured on the 36 475-bin city block operator (rows / rows
// unrolled / FP32 column unroll): 4/1/1 977 and 1005 us, 4/2/8 908 and
// 926 us, 8/2/4 940 and 938 us; FP16 column product 447 us unrolled 1,
// 453-476 us unrolled 2-8.
#ifndef IRTT_GEMV_UNROLL
#define IRTT_GEMV_UNROLL 2
#endif
#ifndef IRTT_GEMVT_UNROLL
#define IRTT_GEMVT_UNROLL 8
#endif
#ifndef IRTT_GEMV16T_UNROLL
#define IRTT_GEMV16T_UNROLL 1
#endif
constexpr uint32_t kGemvUnroll = IRTT_GEMV_UNROLL;
constexpr uint32_t kGemvTUnroll = IRTT_GEMVT_UNROLL;
constexpr uint32_t kGemv16TUnroll = IRTT_GEMV16T_UNROLL;

// y_i = sum_j T_ij u_j: kGemvRows rows per wave, float4 loads, FP64 accumulation.
__global__ __launch_bounds__(256) void k_gemv_rows(const float *__restrict__ T, uint32_t nb,
                                                   uint32_t ld, const double *__restrict__ u,
                                                   double *__restrict__ y)
{
    const uint32_t row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * kGemvRows;
    const uint32_t lane = threadIdx.x & 63;
    if (row0 >= nb) return;
    const float4 *r[kGemvRows];
#pragma unroll
    for (uint32_t q = 0; q < kGemvRows; q++)
        r[q] = reinterpret_cast<const float4 *>(T + (size_t)min(row0 + q, nb - 1) * ld);
    const double4 *u4 = reinterpret_cast<const double4 *>(u);
    double acc[kGemvRows] = {};
    const uint32_t n4 = ld >> 2;
    for (uint32_t k0 = lane; k0 < n4; k0 += 64 * kGemvUnroll) {
        float4 v[kGemvUnroll][kGemvRows];
        double4 uu[kGemvUnroll];
#pragma unroll
        for (uint32_t s = 0; s < kGemvUnroll; s++) {
            const uint32_t k = k0 + 64 * s;
            if (k < n4) {
                uu[s] = u4[k];
#pragma unroll
                for (uint32_t q = 0; q < kGemvRows; q++) v[s][q] = r[q][k];
            }
        }
#pragma unroll
        for (uint32_t s = 0; s < kGemvUnroll; s++) {
            if (k0 + 64 * s < n4) {
#pragma unroll
                for (uint32_t q = 0; q < kGemvRows; q++) {
                    // the padded tail of each row is zero and u is padded with zeros
                    const float4 w = v[s][q];
                    acc[q] += (double)w.x * uu[s].x + (double)w.y * uu[s].y +
                              (double)w.z * uu[s].z + (double)w.w * uu[s].w;
                }
            }
        }
    }
#pragma unroll
    for (uint32_t q = 0; q < kGemvRows; q++) {
        double a = acc[q];
        for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
        if (lane == 0 && row0 + q < nb) y[row0 + q] = a;
    }
}

// y = T^T u in two deterministic steps: partial column sums over row chunks
// of kTRows rows (threads along columns read coalesced rows), then a fixed-
// order sum over the chunks.
constexpr uint32_t kTRows = 512;
// four columns per thread (float4 loads)
__global__ __launch_bounds__(256) void k_gemv_t_partial(const float *__restrict__ T, uint32_t nb,
                                                        uint32_t ld, const double *__restrict__ u,
                                                        double *__restrict__ part)
{
    uint32_t j4 = blockIdx.x * 256 + threadIdx.x;
    if (4 * j4 >= ld) return;
    uint32_t r0 = blockIdx.y * kTRows, r1 = min(r0 + kTRows, nb);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (uint32_t i0 = r0; i0 < r1; i0 += kGemvTUnroll) {
        float4 v[kGemvTUnroll];
        double w[kGemvTUnroll];
#pragma unroll
        for (uint32_t s = 0; s < kGemvTUnroll; s++) {
            if (i0 + s < r1) {
                v[s] = reinterpret_cast<const float4 *>(T + (size_t)(i0 + s) * ld)[j4];
                w[s] = u[i0 + s];
            }
        }
#pragma unroll
        for (uint32_t s = 0; s < kGemvTUnroll; s++) {
            if (i0 + s < r1) {
                a0 += (double)v[s].x * w[s];
                a1 += (double)v[s].y * w[s];
                a2 += (double)v[s].z * w[s];
                a3 += (double)v[s].w * w[s];
            }
        }
    }
    double *o = part + (size_t)blockIdx.y * ld + 4 * j4;
    o[0] = a0;
    o[1] = a1;
    o[2] = a2;
    o[3] = a3;
}

__global__ void k_chunk_reduce(const double *__restrict__ part, uint32_t nchunk, uint32_t ld,
                               double *__restrict__ y)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ld) return;
    double a = 0.0;
    for (uint32_t c = 0; c < nchunk; c++) a += part[(size_t)c * ld + j];
    y[j] = a;
}

// ---- FP16 dense form ----------------------------------------------------
__global__ __launch_bounds__(256) void k_row_scale16(const float *__restrict__ T, uint32_t nb,
                                                     uint32_t ld, float *__restrict__ scale)
{
    __shared__ float sm[256];
    const uint32_t r = blockIdx.x;
    const float *row = T + (size_t)r * ld;
    float m = 0.0f;
    for (uint32_t c = threadIdx.x; c < nb; c += 256)
        if (c != r) m = fmaxf(m, fabsf(row[c]));
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) scale[r] = sm[0] * 0x1p-14f;
}

__global__ void k_quant16(const float *__restrict__ T, uint32_t nb, uint32_t ld,
                          const float *__restrict__ scale, __half *__restrict__ H)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nb * ld) return;
    uint32_t r = (uint32_t)(i / ld), c = (uint32_t)(i % ld);
    float s = scale[r];
    float v = (c < nb && c != r && s > 0.0f) ? T[i] / s : 0.0f;
    H[i] = __float2half(v);
}

__global__ void k_vec_f32(const double *__restrict__ u, uint32_t n, float *__restrict__ o)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = (float)u[i];
}

// y_r = scale_r * sum_c H_rc u_c: four rows per wave, 8 halves (16 B) per
// lane and row per step, FP32 products and partial sums, FP64 across lanes
__global__ __launch_bounds__(256) void k_gemv16_rows(const __half *__restrict__ H,
                                                     const float *__restrict__ scale, uint32_t nb,
                                                     uint32_t ld, const float *__restrict__ u,
                                                     double *__restrict__ y)
{
    const uint32_t row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * kGemvRows;
    const uint32_t lane = threadIdx.x & 63;
    if (row0 >= nb) return;
    const uint4 *r[kGemvRows];
#pragma unroll
    for (uint32_t q = 0; q < kGemvRows; q++)
        r[q] = reinterpret_cast<const uint4 *>(H + (size_t)min(row0 + q, nb - 1) * ld);
    const float4 *u4 = reinterpret_cast<const float4 *>(u);
    float acc[kGemvRows] = {};
    const uint32_t n8 = ld >> 3;
    for (uint32_t k0 = lane; k0 < n8; k0 += 64 * kGemvUnroll) {
        uint4 hv[kGemvUnroll][kGemvRows];
        float4 ua[kGemvUnroll], ub[kGemvUnroll];
#pragma unroll
        for (uint32_t s = 0; s < kGemvUnroll; s++) {
            const uint32_t k = k0 + 64 * s;
            if (k < n8) {
                ua[s] = u4[2 * k];
                ub[s] = u4[2 * k + 1];
#pragma unroll
                for (uint32_t q = 0; q < kGemvRows; q++) hv[s][q] = r[q][k];
            }
        }
#pragma unroll
        for (uint32_t s = 0; s < kGemvUnroll; s++) {
            if (k0 + 64 * s < n8) {
#pragma unroll
                for (uint32_t q = 0; q < kGemvRows; q++) {
                    const __half2 *hp = reinterpret_cast<const __half2 *>(&hv[s][q]);
                    float2 f0 = __half22float2(hp[0]), f1 = __half22float2(hp[1]);
                    float2 f2 = __half22float2(hp[2]), f3 = __half22float2(hp[3]);
                    acc[q] += f0.x * ua[s].x + f0.y * ua[s].y + f1.x * ua[s].z + f1.y * ua[s].w +
                              f2.x * ub[s].x + f2.y * ub[s].y + f3.x * ub[s].z + f3.y * ub[s].w;
                }
            }
        }
    }
#pragma unroll
    for (uint32_t q = 0; q < kGemvRows; q++) {
        double a = acc[q];
        for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
        if (lane == 0 && row0 + q < nb) y[row0 + q] = a * (double)scale[row0 + q];
    }
}

// part[chunk][c] = sum_{r in chunk} scale_r u_r H_rc, 8 columns per thread
__global__ __launch_bounds__(256) void k_gemv16_t_partial(const __half *__restrict__ H,
                                                          const float *__restrict__ scale,
                                                          uint32_t nb, uint32_t ld,
                                                          const double *__restrict__ u,
                                                          double *__restrict__ part)
{
    uint32_t c8 = blockIdx.x * 256 + threadIdx.x;
    if (8 * c8 >= ld) return;
    uint32_t r0 = blockIdx.y * kTRows, r1 = min(r0 + kTRows, nb);
    float acc[8] = {};
    for (uint32_t i0 = r0; i0 < r1; i0 += kGemv16TUnroll) {
        uint4 hv[kGemv16TUnroll];
        float w[kGemv16TUnroll];
#pragma unroll
        for (uint32_t s = 0; s < kGemv16TUnroll; s++) {
            if (i0 + s < r1) {
                w[s] = (float)((double)scale[i0 + s] * u[i0 + s]);
                hv[s] = reinterpret_cast<const uint4 *>(H + (size_t)(i0 + s) * ld)[c8];
            }
        }
#pragma unroll
        for (uint32_t s = 0; s < kGemv16TUnroll; s++) {
            if (i0 + s < r1) {
                const __half2 *hp = reinterpret_cast<const __half2 *>(&hv[s]);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    float2 f = __half22float2(hp[q]);
                    acc[2 * q] += f.x * w[s];
                    acc[2 * q + 1] += f.y * w[s];
                }
            }
        }
    }
    double *o = part + (size_t)blockIdx.y * ld + 8 * c8;
    for (int q = 0; q < 8; q++) o[q] = acc[q];
}

// ---- many right-hand sides: FP16 MFMA GEMM --------------------------------
// Y = P X for R vectors at once with v_mfma_f32_32x32x16_f16 (FP16 inputs,
// FP32 accumulation).  Each wave owns a 32-row (or, transposed, 32-column)
// tile of the operator and one or two 32-vector tiles of X; every lane
// supplies 32 consecutive k of its row for four MFMAs (the k order inside a
// dot product is free, and A and B use the same order), so the operator is
// read once, 64 B per lane per step.  K is split over grid.y for occupancy;
// the partial sums are reduced in a fixed order.
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr uint32_t kMfmaStep = 64;  // k per wave iteration

__device__ __forceinline__ half8_t as_half8(uint4 v)
{
    return __builtin_bit_cast(half8_t, v);
}

// xs[r] = max_i |rs_i x_r[i]| / 2^14 (rs = nullptr: 1), in two steps: block
// (r, s) takes the maximum over slice s of kRhsSlices of vector r, then one
// block takes the maximum of the slices (a maximum does not depend on the
// grouping).  One block per vector would leave most CUs idle.
constexpr uint32_t kRhsSlices = 16;
__global__ __launch_bounds__(256) void k_rhs_scale(const double *__restrict__ X, uint32_t n,
                                                   uint32_t ld, const float *__restrict__ rs,
                                                   double *__restrict__ smax)
{
    __shared__ double sm[256];
    const double *x = X + (size_t)blockIdx.x * ld;
    const uint32_t per = (n + kRhsSlices - 1) / kRhsSlices;
    const uint32_t i0 = blockIdx.y * per, i1 = min(i0 + per, n);
    double m = 0.0;
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += 256)
        m = fmax(m, fabs(x[i] * (rs ? (double)rs[i] : 1.0)));
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) smax[blockIdx.x * kRhsSlices + blockIdx.y] = sm[0];
}

__global__ void k_rhs_scale_fin(const double *__restrict__ smax, uint32_t R,
                                float *__restrict__ xs)
{
    const uint32_t r = threadIdx.x;
    if (r >= R) return;
    double m = 0.0;
    for (uint32_t k = 0; k < kRhsSlices; k++) m = fmax(m, smax[r * kRhsSlices + k]);
    xs[r] = (float)(m * 0x1p-14);
}

// Xh = FP16(rs_i x_r[i] / xs[r]), zero for r >= R and i >= n, stored in MFMA
// fragment order like the operator: vector tile nt (32 vectors) x k-block kb
// = [t][lane (hh, m)][8], lane holding vector 32 nt + m, k = 64 kb + 32 hh + 8 t
__global__ void k_rhs_quant(const double *__restrict__ X, uint32_t n, uint32_t ld, uint32_t R,
                            uint32_t rpad, const float *__restrict__ rs,
                            const float *__restrict__ xs, _Float16 *__restrict__ Xh)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)rpad * ld) return;
    uint32_t r = (uint32_t)(e / ld), i = (uint32_t)(e % ld);
    float v = 0.0f;
    if (r < R && i < n && xs[r] > 0.0f)
        v = (float)(X[(size_t)r * ld + i] * (rs ? (double)rs[i] : 1.0) / (double)xs[r]);
    const uint32_t KB = ld / kMfmaStep, nt = r >> 5, m = r & 31, kb = i / kMfmaStep;
    const uint32_t w = i % kMfmaStep, hh = w >> 5, t = (w & 31) >> 3, j = w & 7;
    Xh[(((size_t)nt * KB + kb) * 256 + 64 * t + 32 * hh + m) * 8 + j] = (_Float16)v;
}

// The FP16 operator in MFMA fragment order: tile (32 rows) x k-block (64 k)
// = 2048 halves stored as [t 0..3][lane 0..63][8], lane (hh, m) = 32 hh + m
// holding row 32 tile + m, k = 64 kb + 32 hh + 8 t + 0..7.  One load
// instruction of a wave then reads 1 KiB contiguously.  The transposed copy
// holds the same for H^T (rows = operator columns, k = operator rows).
__global__ void k_frag_pack(const __half *__restrict__ H, uint32_t nb, uint32_t ld,
                            uint32_t tiles, int transposed, uint4 *__restrict__ F)
{
    const uint32_t KB = ld / kMfmaStep;
    const size_t u = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= (size_t)tiles * KB * 256) return;
    const uint32_t lane = (uint32_t)(u & 63), t = (uint32_t)((u >> 6) & 3);
    const size_t blk = u >> 8;
    const uint32_t kb = (uint32_t)(blk % KB), tile = (uint32_t)(blk / KB);
    const uint32_t m = lane & 31, hh = lane >> 5;
    const uint32_t k0 = kMfmaStep * kb + 32 * hh + 8 * t;
    const uint16_t *Hb = reinterpret_cast<const uint16_t *>(H);
    uint16_t v[8];
    if (!transposed) {
        const uint32_t row = 32 * tile + m;
        for (int j = 0; j < 8; j++) v[j] = row < nb ? Hb[(size_t)row * ld + k0 + j] : 0;
    } else {
        const uint32_t col = 32 * tile + m;
        for (int j = 0; j < 8; j++) v[j] = k0 + j < nb ? Hb[(size_t)(k0 + j) * ld + col] : 0;
    }
    uint4 o;
    o.x = v[0] | ((uint32_t)v[1] << 16);
    o.y = v[2] | ((uint32_t)v[3] << 16);
    o.z = v[4] | ((uint32_t)v[5] << 16);
    o.w = v[6] | ((uint32_t)v[7] << 16);
    F[u] = o;
}

// part[split][r][i] = sum_{k in split} A[i][k] Xh[r][k] with A in fragment
// order (the operator, or its transpose); rows i < nrows are written
template <int NT>
__global__ __launch_bounds__(256) void k_gemm16(const uint4 *__restrict__ F, uint32_t tiles,
                                                uint32_t nrows, uint32_t ld,
                                                const _Float16 *__restrict__ Xh, uint32_t kbchunk,
                                                float *__restrict__ part)
{
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t m = lane & 31, hh = lane >> 5;
    const uint32_t tile = blockIdx.x * 4 + wave;
    if (tile >= tiles) return;
    const uint32_t KB = ld / kMfmaStep;
    const uint32_t kb0 = blockIdx.y * kbchunk, kb1 = min(kb0 + kbchunk, KB);
    const uint4 *fa = F + ((size_t)tile * KB) * 256 + lane;
    const uint4 *fb = reinterpret_cast<const uint4 *>(Xh) + lane;
    f32x16_t acc[NT];
    for (int nt = 0; nt < NT; nt++)
        for (int v = 0; v < 16; v++) acc[nt][v] = 0.0f;
    for (uint32_t kb = kb0; kb < kb1; kb++) {
        uint4 a[4];
        for (int t = 0; t < 4; t++) a[t] = fa[(size_t)kb * 256 + 64 * t];
        for (int nt = 0; nt < NT; nt++) {
            uint4 b[4];
            for (int t = 0; t < 4; t++) b[t] = fb[((size_t)nt * KB + kb) * 256 + 64 * t];
            for (int t = 0; t < 4; t++)
                acc[nt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_half8(a[t]), as_half8(b[t]),
                                                                 acc[nt], 0, 0, 0);
        }
    }
    // output tile: lane holds column m of each vector tile, rows
    // 8 (v / 4) + 4 hh + v % 4
    float *o = part + (size_t)blockIdx.y * NT * 32 * ld;
    const uint32_t row0 = 32 * tile;
    for (int nt = 0; nt < NT; nt++)
        for (int v = 0; v < 16; v++) {
            uint32_t i = row0 + 8 * (v >> 2) + 4 * hh + (v & 3);
            if (i < nrows) o[(size_t)(nt * 32 + m) * ld + i] = acc[nt][v];
        }
}

// Y[r][i] = ys_i xs[r] sum_split part + cd_i X[r][i] for r < R
__global__ void k_gemm_reduce(const float *__restrict__ part, uint32_t nsplit, uint32_t rpad,
                              uint32_t R, uint32_t n, uint32_t ld, const float *__restrict__ ys,
                              const float *__restrict__ xs, const double *__restrict__ cd,
                              const double *__restrict__ X, double *__restrict__ Y)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)R * ld) return;
    uint32_t r = (uint32_t)(t / ld), i = (uint32_t)(t % ld);
    if (i >= n) {
        Y[t] = 0.0;
        return;
    }
    double a = 0.0;
    for (uint32_t sp = 0; sp < nsplit; sp++) a += (double)part[((size_t)sp * rpad + r) * ld + i];
    a *= (double)xs[r] * (ys ? (double)ys[i] : 1.0);
    if (cd) a += cd[i] * X[t];
    Y[t] = a;
}

// ---- many right-hand sides on the FP32 operator: FP64 MFMA GEMM -----------
// Y = T X with v_mfma_f64_16x16x4_f64: the FP32 entries are widened exactly,
// products and sums are FP64, so the result has the accuracy of the FP64-
// accumulated single-vector product (the refinement residual of the FP16
// form runs here).  A wave owns 64 rows (four 16-row tiles) and up to 32
// vectors (two 16-vector tiles); per 16 k every lane loads one float4 of each
// of its rows and one double4 of each vector and feeds them to four MFMAs
// (k slot g of MFMA t holds k = 4 g + t in both operands).
typedef double f64x4_t __attribute__((ext_vector_type(4)));
constexpr uint32_t kM64 = 4;  // 16-row tiles per wave

template <int NT>
__global__ __launch_bounds__(256) void k_gemm64(const float *__restrict__ T, uint32_t nb,
                                                uint32_t ld, const double *__restrict__ X,
                                                uint32_t kchunk, double *__restrict__ part)
{
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t m = lane & 15, g = lane >> 4;
    const uint32_t row0 = (blockIdx.x * 4 + wave) * 16 * kM64;
    if (row0 >= nb) return;
    const uint32_t k0 = blockIdx.y * kchunk, k1 = min(k0 + kchunk, ld);
    const float *arow[kM64];
    for (uint32_t a = 0; a < kM64; a++)
        arow[a] = T + (size_t)min(row0 + 16 * a + m, nb - 1) * ld + 4 * g;
    const double *brow[NT];
    for (int b = 0; b < NT; b++) brow[b] = X + (size_t)(16 * b + m) * ld + 4 * g;
    f64x4_t acc[kM64][NT];
    for (uint32_t a = 0; a < kM64; a++)
        for (int b = 0; b < NT; b++) acc[a][b] = f64x4_t{0.0, 0.0, 0.0, 0.0};
    for (uint32_t k = k0; k < k1; k += 16) {
        float4 av[kM64];
        for (uint32_t a = 0; a < kM64; a++) av[a] = *reinterpret_cast<const float4 *>(arow[a] + k);
        double4 bv[NT];
        for (int b = 0; b < NT; b++) bv[b] = *reinterpret_cast<const double4 *>(brow[b] + k);
        for (uint32_t a = 0; a < kM64; a++)
            for (int b = 0; b < NT; b++) {
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[a].x, bv[b].x,
                                                                 acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[a].y, bv[b].y,
                                                                 acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[a].z, bv[b].z,
                                                                 acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[a].w, bv[b].w,
                                                                 acc[a][b], 0, 0, 0);
            }
    }
    // output tile: lane holds vector m of each tile, rows 4 v + g
    double *o = part + (size_t)blockIdx.y * NT * 16 * ld;
    for (uint32_t a = 0; a < kM64; a++)
        for (int b = 0; b < NT; b++)
            for (int v = 0; v < 4; v++) {
                uint32_t i = row0 + 16 * a + 4 * v + g;
                if (i < nb) o[(size_t)(16 * b + m) * ld + i] = acc[a][b][v];
            }
}

// transposed: part[split][r][c] = sum_i T[i][c] X[r][i]; the wave owns 64
// columns, each lane reads its column down four rows per 16 k
template <int NT>
__global__ __launch_bounds__(256) void k_gemm64_t(const float *__restrict__ T, uint32_t nb,
                                                  uint32_t ld, const double *__restrict__ X,
                                                  uint32_t kchunk, double *__restrict__ part)
{
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t m = lane & 15, g = lane >> 4;
    const uint32_t col0 = (blockIdx.x * 4 + wave) * 16 * kM64;
    if (col0 >= ld) return;
    const uint32_t k0 = blockIdx.y * kchunk, k1 = min(k0 + kchunk, ld);
    const double *brow[NT];
    for (int b = 0; b < NT; b++) brow[b] = X + (size_t)(16 * b + m) * ld + 4 * g;
    f64x4_t acc[kM64][NT];
    for (uint32_t a = 0; a < kM64; a++)
        for (int b = 0; b < NT; b++) acc[a][b] = f64x4_t{0.0, 0.0, 0.0, 0.0};
    for (uint32_t k = k0; k < k1; k += 16) {
        double4 bv[NT];
        for (int b = 0; b < NT; b++) bv[b] = *reinterpret_cast<const double4 *>(brow[b] + k);
        for (uint32_t a = 0; a < kM64; a++) {
            const uint32_t c = col0 + 16 * a + m;
            float av[4];
            for (int t = 0; t < 4; t++) {
                // rows beyond nb meet zero padding in X; clamp the read
                const uint32_t i = min(k + 4 * g + t, nb - 1);
                av[t] = T[(size_t)i * ld + c];
            }
            for (int b = 0; b < NT; b++) {
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[0], bv[b].x,
                                                                 acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[1], bv[b].y,
                                                                 acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[2], bv[b].z,
                                                                 acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[3], bv[b].w,
                                                                 acc[a][b], 0, 0, 0);
            }
        }
    }
    double *o = part + (size_t)blockIdx.y * NT * 16 * ld;
    for (uint32_t a = 0; a < kM64; a++)
        for (int b = 0; b < NT; b++)
            for (int v = 0; v < 4; v++) {
                uint32_t c = col0 + 16 * a + 4 * v + g;
                o[(size_t)(16 * b + m) * ld + c] = acc[a][b][v];
            }
}

// Y[r][i] = sum_split part (fixed order), zero past n
__global__ void k_gemm64_reduce(const double *__restrict__ part, uint32_t nsplit, uint32_t rpad,
                                uint32_t R, uint32_t n, uint32_t ld, double *__restrict__ Y)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)R * ld) return;
    uint32_t r = (uint32_t)(t / ld), i = (uint32_t)(t % ld);
    double a = 0.0;
    if (i < n)
        for (uint32_t sp = 0; sp < nsplit; sp++) a += part[((size_t)sp * rpad + r) * ld + i];
    Y[t] = a;
}

// y += cd * x
__global__ void k_add_diag(uint32_t n, const double *__restrict__ cd, const double *__restrict__ x,
                           double *__restrict__ y)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += cd[i] * x[i];
}

}  // namespace

// back to the dense FP32 form; the other forms are released and the sparse
// layout dropped, its buffers kept for the next sparse build (a new
// operator, or new geometry)
void op_reset(Scene &s)
{
    s.op_kind = IRTT_OP_DENSE;
    s.sp.invalidate();
    s.f16_valid = false;
    s.d_h16.release();
    s.d_h16s.release();
    s.d_h16f.release();
    s.d_h16ft.release();
    s.aca.reset();
    s.hm.reset();
}

void op_select(Scene &s, uint32_t kind)
{
    if (!s.has_transport) throw Error(IRTT_ERR_STATE, "transport operator not built");
    const bool sparse_kind = kind == IRTT_OP_SPARSE || kind == IRTT_OP_HMATRIX;
    if (sparse_kind != (s.layout == IRTT_LAYOUT_SPARSE))
        throw Error(IRTT_ERR_STATE,
                    s.layout == IRTT_LAYOUT_SPARSE
                        ? "the sparse transport layout offers IRTT_OP_SPARSE and IRTT_OP_HMATRIX only"
                        : "IRTT_OP_SPARSE and IRTT_OP_HMATRIX need the sparse transport layout");
    if (kind == IRTT_OP_SPARSE) {
        // the compressed rows and columns of the build
    } else if (kind == IRTT_OP_HMATRIX) {
        if (!s.hm.valid)
            throw Error(IRTT_ERR_STATE, "no H-matrix operator (irtt_compress_hierarchical)");
    } else if (kind == IRTT_OP_ACA) {
        if (!s.aca.valid) throw Error(IRTT_ERR_STATE, "no ACA operator (irtt_compress_transport)");
    } else if (kind == IRTT_OP_DENSE_F16) {
        if (!s.f16_valid) {
            const uint32_t nb = s.nb, ld = s.ld;
            hipStream_t stm = s.stream;
            s.d_h16.alloc((size_t)nb * ld);
            s.d_h16s.alloc(nb);
            k_row_scale16<<<nb, 256, 0, stm>>>(s.d_T.p, nb, ld, s.d_h16s.p);
            k_quant16<<<cdiv((size_t)nb * ld, 256), 256, 0, stm>>>(
                s.d_T.p, nb, ld, s.d_h16s.p, reinterpret_cast<__half *>(s.d_h16.p));
            IRTT_HIP(hipGetLastError());
            IRTT_HIP(hipStreamSynchronize(stm));
            s.f16_valid = true;
        }
    } else if (kind != IRTT_OP_DENSE) {
        throw Error(IRTT_ERR_ARG, "unknown operator kind");
    }
    s.op_kind = kind;
}

// diagonal kept outside the product P of op_apply (nb doubles, empty = none)
const std::vector<double> &op_diag_outside(const Scene &s, bool refine)
{
    static const std::vector<double> none;
    if (s.op_kind == IRTT_OP_ACA) return s.aca.cdiag;
    if (s.op_kind == IRTT_OP_DENSE_F16 && !refine) return s.Tdiag;
    return none;
}

// full diagonal of the operator the Krylov products use
const std::vector<double> &op_diag_total(const Scene &s)
{
    return s.op_kind == IRTT_OP_ACA ? s.aca.opdiag : s.Tdiag;
}

// the allocations of op_apply, made ahead of a graph capture of products
void op_prepare(Scene &s, bool transposed)
{
    const uint32_t nchunk = cdiv(s.nb, kTRows);
    if (transposed) s.d_tpart.alloc((size_t)nchunk * s.ld);
    if (s.op_kind == IRTT_OP_DENSE_F16) s.d_u32.alloc(s.ld);
}

// y = P w (or P^T w).  refine: the product of the residual ("refinement")
// side, which uses the FP32 operator whenever the Krylov side is FP16.
void op_apply(Scene &s, const double *w, double *y, bool transposed, bool refine)
{
    const uint32_t nb = s.nb, ld = s.ld;
    hipStream_t st = s.stream;
    uint32_t kind = s.op_kind;
    if (kind == IRTT_OP_DENSE_F16 && refine) kind = IRTT_OP_DENSE;
    if (kind == IRTT_OP_HMATRIX && refine) kind = IRTT_OP_SPARSE;
    if (kind == IRTT_OP_HMATRIX) {
        hm_apply(s, w, y, transposed);
        return;
    }
    if (kind == IRTT_OP_ACA) {
        aca_apply(s, w, y, transposed);
        return;
    }
    if (kind == IRTT_OP_SPARSE) {
        sp_apply(s, w, y, transposed);
        return;
    }
    const uint32_t nchunk = cdiv(nb, kTRows);
    if (kind == IRTT_OP_DENSE) {
        if (!transposed) {
            k_gemv_rows<<<cdiv(nb, 4 * kGemvRows), 256, 0, st>>>(s.d_T.p, nb, ld, w, y);
        } else {
            s.d_tpart.alloc((size_t)nchunk * ld);
            k_gemv_t_partial<<<dim3(cdiv(ld / 4, 256), nchunk), 256, 0, st>>>(s.d_T.p, nb, ld, w,
                                                                           s.d_tpart.p);
            k_chunk_reduce<<<cdiv(ld, 256), 256, 0, st>>>(s.d_tpart.p, nchunk, ld, y);
        }
        return;
    }
    const __half *H = reinterpret_cast<const __half *>(s.d_h16.p);
    if (!transposed) {
        s.d_u32.alloc(ld);
        k_vec_f32<<<cdiv(ld, 256), 256, 0, st>>>(w, ld, s.d_u32.p);
        k_gemv16_rows<<<cdiv(nb, 4 * kGemvRows), 256, 0, st>>>(H, s.d_h16s.p, nb, ld, s.d_u32.p, y);
    } else {
        s.d_tpart.alloc((size_t)nchunk * ld);
        k_gemv16_t_partial<<<dim3(cdiv(ld / 8, 256), nchunk), 256, 0, st>>>(H, s.d_h16s.p, nb, ld,
                                                                            w, s.d_tpart.p);
        k_chunk_reduce<<<cdiv(ld, 256), 256, 0, st>>>(s.d_tpart.p, nchunk, ld, y);
    }
}

// Y = Op X (or Op^T X) for R vectors: X, Y are R x ld device arrays (FP64,
// padded with zeros), Op the full operator of the Krylov side (P plus the
// diagonal kept outside it).  The FP16 form runs one MFMA GEMM per 64
// vectors; the other forms one product per vector.
void op_apply_batch(Scene &s, uint32_t R, const double *X, double *Y, bool transposed,
                    bool refine)
{
    const uint32_t nb = s.nb, ld = s.ld;
    hipStream_t st = s.stream;
    if (s.op_kind == IRTT_OP_SPARSE || (refine && s.op_kind == IRTT_OP_HMATRIX)) {
        sp_apply_batch(s, R, X, Y, transposed);
        return;
    }
    if (s.op_kind == IRTT_OP_DENSE || (refine && s.op_kind == IRTT_OP_DENSE_F16)) {
        // the FP32 operator (diagonal inside) with FP64 MFMA, 32 vectors per pass
        const uint32_t tiles = transposed ? ld / (16 * kM64) : cdiv(nb, 16 * kM64);
        const uint32_t ksteps = ld / 16;
        uint32_t ksplit = std::min<uint32_t>(std::max<uint32_t>(1, cdiv(4096, tiles)), ksteps);
        const uint32_t kchunk = cdiv(ksteps, ksplit) * 16;
        ksplit = cdiv(ld, kchunk);
        s.d_x64.alloc((size_t)32 * ld);
        s.d_gpart64.alloc((size_t)ksplit * 32 * ld);
        for (uint32_t r0 = 0; r0 < R; r0 += 32) {
            const uint32_t rc = std::min<uint32_t>(32, R - r0), nt = cdiv(rc, 16), rpad = 16 * nt;
            IRTT_HIP(hipMemcpyAsync(s.d_x64.p, X + (size_t)r0 * ld, (size_t)rc * ld * 8,
                                    hipMemcpyDeviceToDevice, st));
            if (rpad > rc)
                IRTT_HIP(hipMemsetAsync(s.d_x64.p + (size_t)rc * ld, 0, (size_t)(rpad - rc) * ld * 8,
                                        st));
            dim3 grid(cdiv(tiles, 4), ksplit);
            if (!transposed) {
                if (nt == 1) k_gemm64<1><<<grid, 256, 0, st>>>(s.d_T.p, nb, ld, s.d_x64.p, kchunk,
                                                               s.d_gpart64.p);
                else k_gemm64<2><<<grid, 256, 0, st>>>(s.d_T.p, nb, ld, s.d_x64.p, kchunk,
                                                        s.d_gpart64.p);
            } else {
                if (nt == 1) k_gemm64_t<1><<<grid, 256, 0, st>>>(s.d_T.p, nb, ld, s.d_x64.p, kchunk,
                                                                 s.d_gpart64.p);
                else k_gemm64_t<2><<<grid, 256, 0, st>>>(s.d_T.p, nb, ld, s.d_x64.p, kchunk,
                                                          s.d_gpart64.p);
            }
            k_gemm64_reduce<<<cdiv((size_t)rc * ld, 256), 256, 0, st>>>(
                s.d_gpart64.p, ksplit, rpad, rc, nb, ld, Y + (size_t)r0 * ld);
        }
        return;
    }
    const std::vector<double> &cdv = op_diag_outside(s, false);
    const double *cd = nullptr;
    if (!cdv.empty()) {
        s.d_cd.alloc(ld);
        IRTT_HIP(hipMemsetAsync(s.d_cd.p, 0, s.d_cd.bytes(), st));
        IRTT_HIP(hipMemcpyAsync(s.d_cd.p, cdv.data(), nb * sizeof(double), hipMemcpyHostToDevice,
                                st));
        cd = s.d_cd.p;
    }
    if (s.op_kind != IRTT_OP_DENSE_F16) {
        for (uint32_t r = 0; r < R; r++) {
            op_apply(s, X + (size_t)r * ld, Y + (size_t)r * ld, transposed, false);
            if (cd) k_add_diag<<<cdiv(nb, 256), 256, 0, st>>>(nb, cd, X + (size_t)r * ld,
                                                              Y + (size_t)r * ld);
        }
        return;
    }
    const uint32_t tiles = transposed ? ld / 32 : cdiv(nb, 32);  // 32-wide wave tiles
    const uint32_t nrows = transposed ? ld : nb;
    const uint32_t KB = ld / kMfmaStep;
    DBuf<uint16_t> &fr = transposed ? s.d_h16ft : s.d_h16f;
    if (!fr.p) {
        // fragment-order copy of the FP16 operator, built on first use
        const __half *H = reinterpret_cast<const __half *>(s.d_h16.p);
        fr.alloc((size_t)tiles * KB * 2048);
        k_frag_pack<<<cdiv((size_t)tiles * KB * 256, 256), 256, 0, st>>>(
            H, nb, ld, tiles, transposed ? 1 : 0, reinterpret_cast<uint4 *>(fr.p));
    }
    const uint4 *F = reinterpret_cast<const uint4 *>(fr.p);
    uint32_t ksplit = std::min<uint32_t>(std::max<uint32_t>(1, cdiv(4096, tiles)), KB);
    const uint32_t kbchunk = cdiv(KB, ksplit);
    ksplit = cdiv(KB, kbchunk);
    s.d_xs.alloc(64);
    s.d_xsm.alloc(64 * kRhsSlices);
    s.d_xh.alloc((size_t)64 * ld);
    s.d_gpart.alloc((size_t)ksplit * 64 * ld);
    _Float16 *Xh = reinterpret_cast<_Float16 *>(s.d_xh.p);
    const float *rs = s.d_h16s.p;
    for (uint32_t r0 = 0; r0 < R; r0 += 64) {
        const uint32_t rc = std::min<uint32_t>(64, R - r0), nt = cdiv(rc, 32), rpad = 32 * nt;
        const double *Xc = X + (size_t)r0 * ld;
        k_rhs_scale<<<dim3(rc, kRhsSlices), 256, 0, st>>>(Xc, nb, ld, transposed ? rs : nullptr,
                                                           s.d_xsm.p);
        k_rhs_scale_fin<<<1, 64, 0, st>>>(s.d_xsm.p, rc, s.d_xs.p);
        k_rhs_quant<<<cdiv((size_t)rpad * ld, 256), 256, 0, st>>>(
            Xc, nb, ld, rc, rpad, transposed ? rs : nullptr, s.d_xs.p, Xh);
        dim3 grid(cdiv(tiles, 4), ksplit);
        if (nt == 1)
            k_gemm16<1><<<grid, 256, 0, st>>>(F, tiles, nrows, ld, Xh, kbchunk, s.d_gpart.p);
        else k_gemm16<2><<<grid, 256, 0, st>>>(F, tiles, nrows, ld, Xh, kbchunk, s.d_gpart.p);
        k_gemm_reduce<<<cdiv((size_t)rc * ld, 256), 256, 0, st>>>(
            s.d_gpart.p, ksplit, rpad, rc, nb, ld, transposed ? nullptr : rs, s.d_xs.p, cd, Xc,
            Y + (size_t)r0 * ld);
    }
}

}  // namespace irtt
