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
constexpr uint32_t kGemvRows = 4;

// y_i = sum_j T_ij u_j: four rows per wave, float4 loads, FP64 accumulation.
__global__ __launch_bounds__(256) void k_gemv_rows(const float *__restrict__ T, uint32_t nb,
                                                   uint32_t ld, const double *__restrict__ u,
                                                   double *__restrict__ y)
{
    const uint32_t row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * kGemvRows;
    const uint32_t lane = threadIdx.x & 63;
    if (row0 >= nb) return;
    const float4 *r[kGemvRows];
    for (uint32_t q = 0; q < kGemvRows; q++)
        r[q] = reinterpret_cast<const float4 *>(T + (size_t)min(row0 + q, nb - 1) * ld);
    double acc[kGemvRows] = {};
    const uint32_t n4 = ld >> 2;
    for (uint32_t k = lane; k < n4; k += 64) {
        const double4 uu = reinterpret_cast<const double4 *>(u)[k];
        for (uint32_t q = 0; q < kGemvRows; q++) {
            float4 v = r[q][k];
            // the padded tail of each row is zero and u is padded with zeros
            acc[q] += (double)v.x * uu.x + (double)v.y * uu.y + (double)v.z * uu.z +
                      (double)v.w * uu.w;
        }
    }
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
    for (uint32_t i = r0; i < r1; i++) {
        const float4 v = reinterpret_cast<const float4 *>(T + (size_t)i * ld)[j4];
        const double w = u[i];
        a0 += (double)v.x * w;
        a1 += (double)v.y * w;
        a2 += (double)v.z * w;
        a3 += (double)v.w * w;
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
    for (uint32_t q = 0; q < kGemvRows; q++)
        r[q] = reinterpret_cast<const uint4 *>(H + (size_t)min(row0 + q, nb - 1) * ld);
    const float4 *u4 = reinterpret_cast<const float4 *>(u);
    float acc[kGemvRows] = {};
    const uint32_t n8 = ld >> 3;
    for (uint32_t k = lane; k < n8; k += 64) {
        const float4 ua = u4[2 * k], ub = u4[2 * k + 1];
        for (uint32_t q = 0; q < kGemvRows; q++) {
            uint4 h = r[q][k];
            const __half2 *hp = reinterpret_cast<const __half2 *>(&h);
            float2 f0 = __half22float2(hp[0]), f1 = __half22float2(hp[1]);
            float2 f2 = __half22float2(hp[2]), f3 = __half22float2(hp[3]);
            acc[q] += f0.x * ua.x + f0.y * ua.y + f1.x * ua.z + f1.y * ua.w + f2.x * ub.x +
                      f2.y * ub.y + f3.x * ub.z + f3.y * ub.w;
        }
    }
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
    for (uint32_t i = r0; i < r1; i++) {
        float w = (float)((double)scale[i] * u[i]);
        uint4 h = reinterpret_cast<const uint4 *>(H + (size_t)i * ld)[c8];
        const __half2 *hp = reinterpret_cast<const __half2 *>(&h);
        for (int q = 0; q < 4; q++) {
            float2 f = __half22float2(hp[q]);
            acc[2 * q] += f.x * w;
            acc[2 * q + 1] += f.y * w;
        }
    }
    double *o = part + (size_t)blockIdx.y * ld + 8 * c8;
    for (int q = 0; q < 8; q++) o[q] = acc[q];
}

}  // namespace

// back to the dense FP32 form; the other forms are released (a new
// operator, or new geometry)
void op_reset(Scene &s)
{
    s.op_kind = IRTT_OP_DENSE;
    s.f16_valid = false;
    s.d_h16.release();
    s.d_h16s.release();
    s.aca.reset();
}

void op_select(Scene &s, uint32_t kind)
{
    if (!s.has_transport) throw Error(IRTT_ERR_STATE, "transport operator not built");
    if (kind == IRTT_OP_ACA) {
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

// y = P w (or P^T w).  refine: the product of the residual ("refinement")
// side, which uses the FP32 operator whenever the Krylov side is FP16.
void op_apply(Scene &s, const double *w, double *y, bool transposed, bool refine)
{
    const uint32_t nb = s.nb, ld = s.ld;
    hipStream_t st = s.stream;
    uint32_t kind = s.op_kind;
    if (kind == IRTT_OP_DENSE_F16 && refine) kind = IRTT_OP_DENSE;
    if (kind == IRTT_OP_ACA) {
        aca_apply(s, w, y, transposed);
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

}  // namespace irtt
