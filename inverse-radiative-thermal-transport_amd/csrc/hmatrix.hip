// hmatrix.hip -- hierarchical block compression (H-matrix) of the transport
// operator and its products.
//
// The reference compresses T with one global rank-k cross approximation
// (ThermalTransport::compressTransportMatrix, thermal_transport.cpp:281-597:
// partial pivoting from row 0, findArgmaxNotInList pivots :463-474, residual
// rows / columns :476-500) and leaves its early-termination test commented
// out (:439-451).  The 400k-triangle facade asks for hierarchical blocks at a
// relative tolerance with a rank cap instead.  Built here from the sparse
// layout's entries:
//
//   bin order     30-bit Morton codes of the bin positions (host)
//   cluster tree  index ranges halved down to a leaf size, FP64 boxes (host)
//   block tree    far (admissible: min diam <= eta dist) / near pairs (host)
//   operator      relabelled into that order: compressed rows and columns
//                 (radix sort of (row, column) keys on the device)
//   far blocks    one workgroup per block runs partial-pivot ACA in FP64:
//                 the reference's residual sums (l = 0 .. k-1 in order) and
//                 pivot rule, a first row holding the block's largest entry,
//                 vanishing residual rows skipped, stop at
//                 |u_k|^2 |v_k|^2 <= (tol/4)^2 |S_k|_F^2, rank <= 32
//   acceptance    exact error |A - U V|_F from the Gram matrices of U and V
//                 and the block's entries; a block stays in low rank when
//                 the error is within tol |A|_F and (by default) its factors
//                 take fewer bytes than its entries
//   storage       U (m x k) and V^T (n x k) row-major, FP16 with one FP32
//                 scale per row (max|row| 2^-14), or FP32; every entry of a
//                 block not kept in low rank stays in the FP32 near field
//                 (compressed rows and columns in the Morton order)
//
// A product y = H x gathers x into the Morton order, runs the near field
// (one wave per row, FP64 sums), the far blocks in two stages (t_b = V_b x
// per block, then per leaf cluster y_i += U_b(i,:) t_b over the blocks that
// cover the leaf, in a fixed order) and scatters y back.  Every sum has a
// fixed order, so products are reproducible.  The transposed product swaps
// the roles of U and V.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "irtt_internal.hpp"

namespace irtt {
namespace {

constexpr int kT = 256;
constexpr uint32_t kR = kHMaxRank;          // rank cap and factor row pitch bound
constexpr double kAcaMargin = 0.25;         // ACA stops at tol / 4, blocks accepted at tol
constexpr uint32_t kItem = 2048;            // indices per first-stage item of a product
constexpr uint32_t kNone = 0xffffffffu;

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *__restrict__ a, uint32_t b,
                                                    uint32_t e, uint32_t v)
{
    while (b < e) {
        const uint32_t m = b + (e - b) / 2;
        if (a[m] < v) b = m + 1;
        else e = m;
    }
    return b;
}

// halving-tree sum of one partial per thread; every thread receives it
__device__ double tree_sum(double p, double *red)
{
    red[threadIdx.x] = p;
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + off];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// eight such sums at once (red: 8 x kT)
__device__ void tree_sum8(const double *p, double *red, double *out)
{
    for (int q = 0; q < 8; q++) red[q * kT + threadIdx.x] = p[q];
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int q = 0; q < 8; q++)
                red[q * kT + threadIdx.x] = red[q * kT + threadIdx.x] + red[q * kT + threadIdx.x + off];
        __syncthreads();
    }
    for (int q = 0; q < 8; q++) out[q] = red[q * kT];
    __syncthreads();
}

// largest strictly positive value, ties to the smallest index
// (findArgmaxNotInList, thermal_transport.cpp:463-474); index kNone if none
struct Best {
    double v;
    uint32_t i;
};

__device__ __forceinline__ void best_merge(Best &a, double v, uint32_t i)
{
    if (v > a.v || (v == a.v && v > 0.0 && i < a.i)) {
        a.v = v;
        a.i = i;
    }
}

__device__ Best block_best(Best b, double *rv, uint32_t *ri)
{
    rv[threadIdx.x] = b.v;
    ri[threadIdx.x] = b.i;
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            Best a{rv[threadIdx.x], ri[threadIdx.x]};
            best_merge(a, rv[threadIdx.x + off], ri[threadIdx.x + off]);
            rv[threadIdx.x] = a.v;
            ri[threadIdx.x] = a.i;
        }
        __syncthreads();
    }
    Best r{rv[0], ri[0]};
    __syncthreads();
    if (!(r.v > 0.0)) r.i = kNone;
    return r;
}

__device__ uint32_t block_min_u32(uint32_t v, uint32_t *ri)
{
    ri[threadIdx.x] = v;
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) ri[threadIdx.x] = min(ri[threadIdx.x], ri[threadIdx.x + off]);
        __syncthreads();
    }
    const uint32_t r = ri[0];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------- relabel
// one wave per row of the operator: Morton-order key of each entry
// (row' << bits | column', or column' << bits | row' for the column form)
__global__ __launch_bounds__(256) void k_hm_keys(const uint32_t *__restrict__ rowptr,
                                                 const uint32_t *__restrict__ col,
                                                 const float *__restrict__ val, uint32_t nb,
                                                 const uint32_t *__restrict__ iperm, uint32_t bits,
                                                 uint32_t by_col,
                                                 unsigned long long *__restrict__ keys,
                                                 uint32_t *__restrict__ vals)
{
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (r >= nb) return;
    const unsigned long long pr = iperm[r];
    for (uint32_t e = rowptr[r] + lane; e < rowptr[r + 1]; e += 64) {
        const unsigned long long pc = iperm[col[e]];
        keys[e] = by_col ? ((pc << bits) | pr) : ((pr << bits) | pc);
        vals[e] = __float_as_uint(val[e]);
    }
}

__global__ void k_hm_count(const unsigned long long *__restrict__ keys, uint64_t nnz, uint32_t bits,
                           uint32_t *__restrict__ cnt)
{
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz) atomicAdd(&cnt[(uint32_t)(keys[e] >> bits)], 1u);
}

// exclusive scan of n counts into n + 1 offsets; one block of 1024 threads
__global__ __launch_bounds__(1024) void k_hm_scan_u32(const uint32_t *__restrict__ c, uint32_t n,
                                                      uint32_t *__restrict__ off)
{
    __shared__ uint32_t sums[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t b = min(tid * per, n), e = min(b + per, n);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; i++) s += c[i];
    sums[tid] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint32_t v = tid >= o ? sums[tid - o] : 0u;
        __syncthreads();
        sums[tid] += v;
        __syncthreads();
    }
    uint32_t run = sums[tid] - s;
    for (uint32_t i = b; i < e; i++) {
        off[i] = run;
        run += c[i];
    }
    if (tid == 1023) off[n] = sums[1023];
}

__global__ void k_hm_fill(const unsigned long long *__restrict__ keys,
                          const uint32_t *__restrict__ vals, uint64_t nnz, uint32_t bits,
                          uint32_t *__restrict__ idx, float *__restrict__ v)
{
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    idx[e] = (uint32_t)(keys[e] & ((1ull << bits) - 1ull));
    v[e] = __uint_as_float(vals[e]);
}

// ------------------------------------------------------- far-block scan
// entries, squared Frobenius norm and the row of the largest entry of each
// far block (r0, r1, c0, c1)
__global__ __launch_bounds__(kT) void k_hm_blockscan(const uint4 *__restrict__ fb,
                                                     const uint32_t *__restrict__ prp,
                                                     const uint32_t *__restrict__ pcol,
                                                     const float *__restrict__ pval,
                                                     uint32_t *__restrict__ nnz_out,
                                                     double *__restrict__ fro2_out,
                                                     uint32_t *__restrict__ i0_out)
{
    __shared__ double rv[kT];
    __shared__ uint32_t ri[kT];
    const uint4 B = fb[blockIdx.x];
    const uint32_t m = B.y - B.x;
    uint32_t cnt = 0;
    double f = 0.0;
    Best best{0.0, kNone};
    for (uint32_t i = threadIdx.x; i < m; i += kT) {
        const uint32_t R = B.x + i;
        const uint32_t lo = lower_bound_u32(pcol, prp[R], prp[R + 1], B.z);
        const uint32_t hi = lower_bound_u32(pcol, lo, prp[R + 1], B.w);
        cnt += hi - lo;
        for (uint32_t e = lo; e < hi; e++) {
            const double v = (double)pval[e];
            f = f + v * v;
            best_merge(best, fabs(v), i);
        }
    }
    const double fs = tree_sum(f, rv);
    const double cs = tree_sum((double)cnt, rv);
    const Best b = block_best(best, rv, ri);
    if (threadIdx.x == 0) {
        nnz_out[blockIdx.x] = (uint32_t)cs;
        fro2_out[blockIdx.x] = fs;
        i0_out[blockIdx.x] = b.i == kNone ? 0u : b.i;
    }
}

// next unused row after i0 (cyclic) with an entry in columns [c0, c1); -1: none
__device__ int next_row(uint32_t i0, uint32_t m, const uint8_t *used, const uint4 &B,
                        const uint32_t *__restrict__ prp, const uint32_t *__restrict__ pcol,
                        uint32_t *ri)
{
    for (uint32_t base = 1; base <= m; base += kT) {
        const uint32_t d = base + threadIdx.x;
        uint32_t cand = kNone;
        if (d <= m) {
            const uint32_t r = (i0 + d) % m;
            if (!used[r]) {
                const uint32_t R = B.x + r;
                const uint32_t lo = lower_bound_u32(pcol, prp[R], prp[R + 1], B.z);
                if (lo < prp[R + 1] && pcol[lo] < B.w) cand = d;
            }
        }
        const uint32_t dm = block_min_u32(cand, ri);
        if (dm != kNone) return (int)((i0 + dm) % m);
    }
    return -1;
}

// --------------------------------------------------------------- block ACA
// One workgroup per far block with entries.  U (rank-major: u_l at U + l m)
// and V (v_l at V + l n) in FP64; used: one byte per row.
__global__ __launch_bounds__(kT) void k_hm_aca(
    const uint4 *__restrict__ fb, const uint32_t *__restrict__ list,
    const uint32_t *__restrict__ i0s, const uint32_t *__restrict__ prp,
    const uint32_t *__restrict__ pcol, const float *__restrict__ pval,
    const uint32_t *__restrict__ pcp, const uint32_t *__restrict__ prow,
    const float *__restrict__ pcv, const unsigned long long *__restrict__ uoff,
    const unsigned long long *__restrict__ voff, const unsigned long long *__restrict__ foff,
    double tol2, uint32_t kmax, double *__restrict__ Ub, double *__restrict__ Vb,
    uint8_t *__restrict__ usedb, uint32_t *__restrict__ rank_out, uint32_t *__restrict__ conv_out,
    int32_t *__restrict__ piv_out)
{
    __shared__ double red[8 * kT];
    __shared__ uint32_t ri[kT];
    __shared__ double cu[kR];
    __shared__ uint32_t pcl[kR], prl[kR];
    __shared__ uint32_t s_lo, s_hi;
    const uint32_t q = blockIdx.x, b = list[q], t = threadIdx.x;
    const uint4 B = fb[b];
    const uint32_t m = B.y - B.x, n = B.w - B.z;
    double *U = Ub + uoff[q];
    double *V = Vb + voff[q];
    uint8_t *used = usedb + foff[q];
    for (uint32_t i = t; i < m; i += kT) used[i] = 0;
    uint32_t i = i0s[b];
    uint32_t k = 0;
    double s2 = 0.0;
    uint32_t conv = 0;
    __syncthreads();
    while (k < kmax) {
        // residual row i into v_k
        double *r = V + (size_t)k * n;
        if (t < k) cu[t] = U[(size_t)t * m + i];
        for (uint32_t j = t; j < n; j += kT) r[j] = 0.0;
        if (t == 0) {
            const uint32_t R = B.x + i;
            const uint32_t lo = lower_bound_u32(pcol, prp[R], prp[R + 1], B.z);
            s_lo = lo;
            s_hi = lower_bound_u32(pcol, lo, prp[R + 1], B.w);
            used[i] = 1;
        }
        __syncthreads();
        for (uint32_t e = s_lo + t; e < s_hi; e += kT) r[pcol[e] - B.z] = (double)pval[e];
        __syncthreads();
        Best best{0.0, kNone};
        for (uint32_t j = t; j < n; j += kT) {
            double a = r[j];
            for (uint32_t l = 0; l < k; l++) a = a - cu[l] * V[(size_t)l * n + j];
            r[j] = a;
            bool ex = false;
            for (uint32_t l = 0; l < k; l++) ex = ex || pcl[l] == j;
            best_merge(best, ex ? 0.0 : fabs(a), j);
        }
        best = block_best(best, red, ri);
        if (best.i == kNone) {
            // the residual row vanishes: the row is used, no term is added
            const int nr = next_row(i, m, used, B, prp, pcol, ri);
            if (nr < 0) {
                conv = 1;
                break;
            }
            i = (uint32_t)nr;
            continue;
        }
        const uint32_t j = best.i;
        const double p = r[j];
        __syncthreads();
        for (uint32_t jj = t; jj < n; jj += kT) r[jj] = r[jj] / p;
        if (t == 0) {
            pcl[k] = j;
            prl[k] = i;
        }
        // residual column j into u_k
        double *c = U + (size_t)k * m;
        if (t < k) cu[t] = V[(size_t)t * n + j];
        for (uint32_t ii = t; ii < m; ii += kT) c[ii] = 0.0;
        if (t == 0) {
            const uint32_t C = B.z + j;
            const uint32_t lo = lower_bound_u32(prow, pcp[C], pcp[C + 1], B.x);
            s_lo = lo;
            s_hi = lower_bound_u32(prow, lo, pcp[C + 1], B.y);
        }
        __syncthreads();
        for (uint32_t e = s_lo + t; e < s_hi; e += kT) c[prow[e] - B.x] = (double)pcv[e];
        __syncthreads();
        for (uint32_t ii = t; ii < m; ii += kT) {
            double a = c[ii];
            for (uint32_t l = 0; l < k; l++) a = a - cu[l] * U[(size_t)l * m + ii];
            c[ii] = a;
        }
        __syncthreads();
        // |u_k|^2 |v_k|^2 and the cross terms of the Frobenius update
        double part[8], out[8];
        for (int w = 0; w < 8; w++) part[w] = 0.0;
        for (uint32_t ii = t; ii < m; ii += kT) part[0] = part[0] + c[ii] * c[ii];
        for (uint32_t jj = t; jj < n; jj += kT) part[1] = part[1] + r[jj] * r[jj];
        tree_sum8(part, red, out);
        const double nu2 = out[0] * out[1];
        double cross = 0.0;
        for (uint32_t l0 = 0; l0 < k; l0 += 4) {
            for (int w = 0; w < 8; w++) part[w] = 0.0;
            for (uint32_t w = 0; w < 4 && l0 + w < k; w++) {
                const double *ul = U + (size_t)(l0 + w) * m;
                const double *vl = V + (size_t)(l0 + w) * n;
                for (uint32_t ii = t; ii < m; ii += kT) part[w] = part[w] + c[ii] * ul[ii];
                for (uint32_t jj = t; jj < n; jj += kT) part[4 + w] = part[4 + w] + r[jj] * vl[jj];
            }
            tree_sum8(part, red, out);
            for (uint32_t w = 0; w < 4 && l0 + w < k; w++) cross = cross + out[w] * out[4 + w];
        }
        s2 = s2 + nu2 + 2.0 * cross;
        k++;
        if (nu2 <= tol2 * s2) {
            conv = 1;
            break;
        }
        // next row: the largest |u_k| among the unused rows
        Best bi{0.0, kNone};
        for (uint32_t ii = t; ii < m; ii += kT) best_merge(bi, used[ii] ? 0.0 : fabs(c[ii]), ii);
        bi = block_best(bi, red, ri);
        if (bi.i != kNone) {
            i = bi.i;
        } else {
            const int nr = next_row(i, m, used, B, prp, pcol, ri);
            if (nr < 0) {
                conv = 1;
                break;
            }
            i = (uint32_t)nr;
        }
    }
    __syncthreads();
    if (t == 0) {
        rank_out[q] = k;
        conv_out[q] = conv;
    }
    if (t < kR) {
        piv_out[(size_t)q * 2 * kR + t] = t < k ? (int32_t)prl[t] : -1;
        piv_out[(size_t)q * 2 * kR + kR + t] = t < k ? (int32_t)pcl[t] : -1;
    }
}

// ------------------------------------------------------------ exact error
// Factor access: FP64 as built (rank-major), or as stored (row-major FP16
// with a scale per row, or FP32).
struct FacBuilt {
    const double *p;
    uint32_t len;
    __device__ double operator()(uint32_t l, uint32_t i) const { return p[(size_t)l * len + i]; }
};
struct FacStored {
    const __half *h;
    const float *f;
    const float *s;
    uint32_t k;
    __device__ double operator()(uint32_t l, uint32_t i) const
    {
        if (h) return (double)__half2float(h[(size_t)i * k + l]) * (double)s[i];
        return (double)f[(size_t)i * k + l];
    }
};

// Gram matrix entries of the pairs (l, l') of one factor over its length,
// through 256-index tiles in LDS; pair p of the thread (p = t, t + 256,
// t + 512 below k (k + 1) / 2) accumulates in g[]
template <typename Fac>
__device__ void gram(const Fac &F, uint32_t len, uint32_t k, double (*tile)[kT],
                     const uint8_t *pl, const uint8_t *pm, double *g)
{
    const uint32_t np = k * (k + 1) / 2;
    for (int w = 0; w < 3; w++) g[w] = 0.0;
    for (uint32_t base = 0; base < len; base += kT) {
        const uint32_t idx = base + threadIdx.x;
        for (uint32_t l = 0; l < k; l++) tile[l][threadIdx.x] = idx < len ? F(l, idx) : 0.0;
        __syncthreads();
        for (int w = 0; w < 3; w++) {
            const uint32_t p = threadIdx.x + (uint32_t)w * kT;
            if (p >= np) break;
            const double *a = tile[pl[p]], *b = tile[pm[p]];
            double s = 0.0;
            for (int x = 0; x < kT; x++) s = s + a[x] * b[x];
            g[w] = g[w] + s;
        }
        __syncthreads();
    }
}

// |A - U V|_F^2 of each listed block: <U^T U, V V^T> - 2 sum_nz A (U V) + |A|_F^2
template <bool Stored>
__global__ __launch_bounds__(kT) void k_hm_err(
    const uint4 *__restrict__ fb, const uint32_t *__restrict__ list, const uint32_t *__restrict__ rk,
    const unsigned long long *__restrict__ uoff, const unsigned long long *__restrict__ voff,
    const unsigned long long *__restrict__ suoff, const unsigned long long *__restrict__ svoff,
    const double *__restrict__ Ub, const double *__restrict__ Vb, const __half *__restrict__ hu,
    const __half *__restrict__ hv, const float *__restrict__ fu, const float *__restrict__ fv,
    const float *__restrict__ su, const float *__restrict__ sv, const double *__restrict__ fro2,
    const uint32_t *__restrict__ prp, const uint32_t *__restrict__ pcol,
    const float *__restrict__ pval, double *__restrict__ err2)
{
    __shared__ double tile[kR][kT];
    __shared__ double gu[kR * (kR + 1) / 2];
    __shared__ double red[kT];
    __shared__ uint8_t pl[kR * (kR + 1) / 2], pm[kR * (kR + 1) / 2];
    const uint32_t q = blockIdx.x, b = list[q], t = threadIdx.x;
    const uint4 B = fb[b];
    const uint32_t m = B.y - B.x, n = B.w - B.z, k = rk[q];
    const uint32_t np = k * (k + 1) / 2;
    if (t == 0) {
        uint32_t p = 0;
        for (uint32_t l = 0; l < k; l++)
            for (uint32_t l2 = l; l2 < k; l2++) {
                pl[p] = (uint8_t)l;
                pm[p] = (uint8_t)l2;
                p++;
            }
    }
    __syncthreads();
    double g[3];
    double e = 0.0;
    if (Stored) {
        const FacStored FU{hu ? hu + uoff[q] : nullptr, fu ? fu + uoff[q] : nullptr,
                           su ? su + suoff[q] : nullptr, k};
        const FacStored FV{hv ? hv + voff[q] : nullptr, fv ? fv + voff[q] : nullptr,
                           sv ? sv + svoff[q] : nullptr, k};
        gram(FU, m, k, tile, pl, pm, g);
        for (int w = 0; w < 3; w++)
            if (t + w * kT < np) gu[t + w * kT] = g[w];
        __syncthreads();
        gram(FV, n, k, tile, pl, pm, g);
        double x = 0.0;
        for (int w = 0; w < 3; w++) {
            const uint32_t p = t + (uint32_t)w * kT;
            if (p < np) x = x + (pl[p] == pm[p] ? 1.0 : 2.0) * gu[p] * g[w];
        }
        double cr = 0.0;
        for (uint32_t i = t; i < m; i += kT) {
            const uint32_t R = B.x + i;
            const uint32_t lo = lower_bound_u32(pcol, prp[R], prp[R + 1], B.z);
            const uint32_t hi = lower_bound_u32(pcol, lo, prp[R + 1], B.w);
            for (uint32_t ee = lo; ee < hi; ee++) {
                const uint32_t j = pcol[ee] - B.z;
                double d = 0.0;
                for (uint32_t l = 0; l < k; l++) d = d + FU(l, i) * FV(l, j);
                cr = cr + (double)pval[ee] * d;
            }
        }
        e = x - 2.0 * cr;
    } else {
        const FacBuilt FU{Ub + uoff[q], m}, FV{Vb + voff[q], n};
        gram(FU, m, k, tile, pl, pm, g);
        for (int w = 0; w < 3; w++)
            if (t + w * kT < np) gu[t + w * kT] = g[w];
        __syncthreads();
        gram(FV, n, k, tile, pl, pm, g);
        double x = 0.0;
        for (int w = 0; w < 3; w++) {
            const uint32_t p = t + (uint32_t)w * kT;
            if (p < np) x = x + (pl[p] == pm[p] ? 1.0 : 2.0) * gu[p] * g[w];
        }
        double cr = 0.0;
        for (uint32_t i = t; i < m; i += kT) {
            const uint32_t R = B.x + i;
            const uint32_t lo = lower_bound_u32(pcol, prp[R], prp[R + 1], B.z);
            const uint32_t hi = lower_bound_u32(pcol, lo, prp[R + 1], B.w);
            for (uint32_t ee = lo; ee < hi; ee++) {
                const uint32_t j = pcol[ee] - B.z;
                double d = 0.0;
                for (uint32_t l = 0; l < k; l++) d = d + FU(l, i) * FV(l, j);
                cr = cr + (double)pval[ee] * d;
            }
        }
        e = x - 2.0 * cr;
    }
    const double tot = tree_sum(e, red);
    if (t == 0) err2[q] = fmax(0.0, tot + fro2[b]);
}

// ------------------------------------------------------------------- pack
// accepted block q: u_l (rank-major FP64) -> U row-major (FP16 + row scale,
// or FP32); the same for V -> V^T
__global__ __launch_bounds__(kT) void k_hm_pack(
    const uint32_t *__restrict__ lpos, const uint32_t *__restrict__ lenv,
    const uint32_t *__restrict__ rk, const unsigned long long *__restrict__ boff,
    const unsigned long long *__restrict__ poff, const unsigned long long *__restrict__ soff,
    const double *__restrict__ Fb, uint32_t f16, __half *__restrict__ H, float *__restrict__ Fo,
    float *__restrict__ S)
{
    const uint32_t q = blockIdx.y;
    const uint32_t len = lenv[q], k = rk[q];
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i >= len) return;
    const double *src = Fb + boff[lpos[q]];
    if (!f16) {
        for (uint32_t l = 0; l < k; l++) Fo[poff[q] + (size_t)i * k + l] = (float)src[(size_t)l * len + i];
        return;
    }
    double mx = 0.0;
    for (uint32_t l = 0; l < k; l++) mx = fmax(mx, fabs(src[(size_t)l * len + i]));
    const float s = (float)mx * 0x1p-14f;
    S[soff[q] + i] = s;
    for (uint32_t l = 0; l < k; l++) {
        const double v = src[(size_t)l * len + i];
        H[poff[q] + (size_t)i * k + l] =
            s > 0.0f ? __float2half_rn((float)(v / (double)s)) : __float2half_rn(0.0f);
    }
}

// ------------------------------------------------------------ near field
// is column c inside one of the sorted, disjoint intervals iv[b, e)?
__device__ __forceinline__ bool in_intervals(const uint2 *__restrict__ iv, uint32_t b, uint32_t e,
                                             uint32_t c)
{
    // last interval with start <= c
    uint32_t lo = b, hi = e;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (iv[mid].x <= c) lo = mid + 1;
        else hi = mid;
    }
    return lo > b && c < iv[lo - 1].y;
}

// entries of each row (of the Morton-ordered operator) outside the low-rank
// blocks covering its leaf; one wave per row
__global__ __launch_bounds__(256) void k_hm_near_count(const uint32_t *__restrict__ ptr,
                                                       const uint32_t *__restrict__ idx, uint32_t n,
                                                       const uint32_t *__restrict__ leaf_of,
                                                       const uint32_t *__restrict__ lptr,
                                                       const uint2 *__restrict__ liv,
                                                       uint32_t *__restrict__ cnt)
{
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (r >= n) return;
    const uint32_t L = leaf_of[r], ib = lptr[L], ie = lptr[L + 1];
    uint32_t c = 0;
    for (uint32_t e = ptr[r] + lane; e < ptr[r + 1]; e += 64)
        if (!in_intervals(liv, ib, ie, idx[e])) c++;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if (lane == 0) cnt[r] = c;
}

__global__ __launch_bounds__(256) void k_hm_near_fill(
    const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ idx,
    const float *__restrict__ val, uint32_t n, const uint32_t *__restrict__ leaf_of,
    const uint32_t *__restrict__ lptr, const uint2 *__restrict__ liv,
    const uint32_t *__restrict__ nptr, uint32_t *__restrict__ nidx, float *__restrict__ nval)
{
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (r >= n) return;
    const uint32_t L = leaf_of[r], ib = lptr[L], ie = lptr[L + 1];
    uint32_t out = nptr[r];
    for (uint32_t e0 = ptr[r]; e0 < ptr[r + 1]; e0 += 64) {
        const uint32_t e = e0 + lane;
        const bool keep = e < ptr[r + 1] && !in_intervals(liv, ib, ie, idx[e]);
        const unsigned long long mask = __ballot(keep);
        if (keep) {
            const uint32_t o = out + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
            nidx[o] = idx[e];
            nval[o] = val[e];
        }
        out += (uint32_t)__popcll(mask);
    }
}

// ---------------------------------------------------------------- products
__global__ void k_hm_gather(const double *__restrict__ x, const uint32_t *__restrict__ perm,
                            uint32_t n, double *__restrict__ xp)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) xp[i] = x[perm[i]];
}

__global__ void k_hm_scatter(const double *__restrict__ yp, const uint32_t *__restrict__ perm,
                             uint32_t n, double *__restrict__ y)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[perm[i]] = yp[i];
}

// near field, one wave per row, FP64 sums in a fixed order
__global__ __launch_bounds__(256) void k_hm_spmv(const uint32_t *__restrict__ ptr,
                                                 const uint32_t *__restrict__ idx,
                                                 const float *__restrict__ val, uint32_t n,
                                                 const double *__restrict__ u,
                                                 double *__restrict__ y)
{
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (r >= n) return;
    // four entries per lane in flight, four partial sums in a fixed order
    const uint32_t e = ptr[r + 1];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    uint32_t k = ptr[r] + lane;
    for (; k + 192 < e; k += 256) {
        const uint32_t i0 = idx[k], i1 = idx[k + 64], i2 = idx[k + 128], i3 = idx[k + 192];
        const float v0 = val[k], v1 = val[k + 64], v2 = val[k + 128], v3 = val[k + 192];
        a0 += (double)v0 * u[i0];
        a1 += (double)v1 * u[i1];
        a2 += (double)v2 * u[i2];
        a3 += (double)v3 * u[i3];
    }
    for (; k < e; k += 64) a0 += (double)val[k] * u[idx[k]];
    double a = (a0 + a1) + (a2 + a3);
    for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
    if (lane == 0) y[r] = a;
}

// first stage: item (block q, local begin, local end, global base) adds
// sum_j F_q(j, :) x[base + j] into tpart[item][0..k)
template <bool F16>
__global__ __launch_bounds__(kT) void k_hm_stage1(const uint4 *__restrict__ items,
                                                  const uint32_t *__restrict__ rk,
                                                  const unsigned long long *__restrict__ foff,
                                                  const unsigned long long *__restrict__ soff,
                                                  const __half *__restrict__ H,
                                                  const float *__restrict__ F,
                                                  const float *__restrict__ S,
                                                  const double *__restrict__ xp,
                                                  double *__restrict__ tpart)
{
    __shared__ double wsum[4][kR];
    const uint4 it = items[blockIdx.x];
    const uint32_t q = it.x, k = rk[q];
    double acc[kR];
#pragma unroll
    for (uint32_t l = 0; l < kR; l++) acc[l] = 0.0;
    for (uint32_t j = it.y + threadIdx.x; j < it.z; j += kT) {
        double w = xp[it.w + j];
        if (F16) {
            w = w * (double)S[soff[q] + j];
            const __half *row = H + foff[q] + (size_t)j * k;
#pragma unroll
            for (uint32_t l = 0; l < kR; l++)
                if (l < k) acc[l] = acc[l] + (double)__half2float(row[l]) * w;
        } else {
            const float *row = F + foff[q] + (size_t)j * k;
#pragma unroll
            for (uint32_t l = 0; l < kR; l++)
                if (l < k) acc[l] = acc[l] + (double)row[l] * w;
        }
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (uint32_t l = 0; l < kR; l++) {
        double a = acc[l];
        for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
        if (lane == 0) wsum[wave][l] = a;
    }
    __syncthreads();
    // every slot written (zeros past the rank): the per-block sums read all kR
    if (threadIdx.x < kR)
        tpart[(size_t)blockIdx.x * kR + threadIdx.x] =
            ((wsum[0][threadIdx.x] + wsum[1][threadIdx.x]) + wsum[2][threadIdx.x]) +
            wsum[3][threadIdx.x];
}

// t[q][l] = sum over block q's items, in order
__global__ void k_hm_stage1_sum(const uint32_t *__restrict__ iptr, uint32_t nq,
                                const double *__restrict__ tpart, double *__restrict__ tq)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t q = g / kR, l = g % kR;
    if (q >= nq) return;
    double s = 0.0;
    for (uint32_t it = iptr[q]; it < iptr[q + 1]; it++) s = s + tpart[(size_t)it * kR + l];
    tq[(size_t)q * kR + l] = s;
}

// second stage: one wave per leaf; rows i of the leaf add sum_l F_q(i - base_q, l) t_q[l]
// over the blocks q listed for the leaf, in list order
template <bool F16>
__global__ __launch_bounds__(256) void k_hm_stage2(
    const uint32_t *__restrict__ leaf_beg, uint32_t nleaf, const uint32_t *__restrict__ lptr,
    const uint32_t *__restrict__ lblk, const uint32_t *__restrict__ base,
    const uint32_t *__restrict__ rk, const unsigned long long *__restrict__ foff,
    const unsigned long long *__restrict__ soff, const __half *__restrict__ H,
    const float *__restrict__ F, const float *__restrict__ S, const double *__restrict__ tq,
    double *__restrict__ yp)
{
    const uint32_t L = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (L >= nleaf) return;
    for (uint32_t r = leaf_beg[L] + lane; r < leaf_beg[L + 1]; r += 64) {
        double acc = yp[r];
        for (uint32_t x = lptr[L]; x < lptr[L + 1]; x++) {
            const uint32_t q = lblk[x], k = rk[q];
            const uint32_t loc = r - base[q];
            const double *t = tq + (size_t)q * kR;
            double d = 0.0;
            if (F16) {
                const __half *row = H + foff[q] + (size_t)loc * k;
                for (uint32_t l = 0; l < k; l++) d = d + (double)__half2float(row[l]) * t[l];
                d = d * (double)S[soff[q] + loc];
            } else {
                const float *row = F + foff[q] + (size_t)loc * k;
                for (uint32_t l = 0; l < k; l++) d = d + (double)row[l] * t[l];
            }
            acc = acc + d;
        }
        yp[r] = acc;
    }
}

template <typename K, typename V>
size_t sort_temp_bytes(uint64_t n, uint32_t end_bit)
{
    size_t b = 0;
    rocprim::double_buffer<K> kb(nullptr, nullptr);
    rocprim::double_buffer<V> vb(nullptr, nullptr);
    IRTT_HIP(rocprim::radix_sort_pairs(nullptr, b, kb, vb, (size_t)n, 0, end_bit));
    return b;
}

// ---------------------------------------------------------------- host trees
struct Cl {
    uint32_t b, e;
    double lo[3], hi[3];
    int32_t kid[2];
    uint32_t leaf0, leaf1;  // leaves under the cluster: [leaf0, leaf1)
};

double diam(const Cl &c)
{
    const double d0 = c.hi[0] - c.lo[0], d1 = c.hi[1] - c.lo[1], d2 = c.hi[2] - c.lo[2];
    return std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

double dist(const Cl &t, const Cl &s)
{
    double g[3];
    for (int a = 0; a < 3; a++) g[a] = std::max(0.0, std::max(s.lo[a] - t.hi[a], t.lo[a] - s.hi[a]));
    return std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
}

// positions of the bins in FP64: vertices, or triangle centroids
std::vector<double> bin_positions(const Scene &s)
{
    std::vector<double> pos((size_t)s.nb * 3);
    if (s.bin_mode == IRTT_BIN_VERTEX) {
        for (size_t i = 0; i < (size_t)s.nb * 3; i++) pos[i] = (double)s.xyz[i];
    } else {
        for (uint32_t t = 0; t < s.nt; t++)
            for (int a = 0; a < 3; a++) {
                const double v0 = s.xyz[(size_t)s.idx[3 * t] * 3 + a];
                const double v1 = s.xyz[(size_t)s.idx[3 * t + 1] * 3 + a];
                const double v2 = s.xyz[(size_t)s.idx[3 * t + 2] * 3 + a];
                pos[(size_t)t * 3 + a] = (v0 + v1 + v2) / 3.0;
            }
    }
    return pos;
}

std::vector<uint32_t> morton_order(const std::vector<double> &pos, uint32_t n)
{
    double lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = hi[a] = n ? pos[a] : 0.0;
        for (uint32_t i = 0; i < n; i++) {
            lo[a] = std::min(lo[a], pos[(size_t)i * 3 + a]);
            hi[a] = std::max(hi[a], pos[(size_t)i * 3 + a]);
        }
    }
    std::vector<std::pair<uint64_t, uint32_t>> key(n);
    for (uint32_t i = 0; i < n; i++) {
        uint64_t q[3] = {0, 0, 0};
        for (int a = 0; a < 3; a++) {
            const double ext = hi[a] - lo[a];
            if (ext > 0.0) {
                const double v = std::floor((pos[(size_t)i * 3 + a] - lo[a]) / ext * 1024.0);
                q[a] = (uint64_t)std::min(1023.0, std::max(0.0, v));
            }
        }
        uint64_t code = 0;
        for (int b = 0; b < 10; b++)
            for (int a = 0; a < 3; a++) code |= ((q[a] >> b) & 1ull) << (3 * b + (2 - a));
        key[i] = {code, i};
    }
    std::sort(key.begin(), key.end());
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; i++) perm[i] = key[i].second;
    return perm;
}

}  // namespace

// ------------------------------------------------------------------ build
void hm_compress(Scene &s, const irtt_hmatrix_params &pin, irtt_hmatrix_stats &st)
{
    auto t_start = std::chrono::steady_clock::now();
    if (!s.has_transport || s.layout != IRTT_LAYOUT_SPARSE || !s.sp.valid)
        throw Error(IRTT_ERR_STATE, "irtt_compress_hierarchical needs a sparse-layout transport build");
    irtt_hmatrix_params p = pin;
    if (!(p.tol > 0.0)) p.tol = 1e-4;
    if (!(p.eta > 0.0)) p.eta = 1.0;
    if (p.max_rank == 0) p.max_rank = kR;
    if (p.leaf_size == 0) p.leaf_size = 64;
    if (p.max_rank > kR) throw Error(IRTT_ERR_ARG, "max_rank above 32");
    if (p.leaf_size < 8 || p.leaf_size > 4096) throw Error(IRTT_ERR_ARG, "leaf_size outside 8..4096");
    if (p.precision > IRTT_STORE_F16) throw Error(IRTT_ERR_ARG, "unknown precision");
    Hmat &h = s.hm;
    h.reset();
    const uint32_t n = s.nb;
    hipStream_t stm = s.stream;
    st = irtt_hmatrix_stats{};

    // bin order, clusters, blocks (host)
    const std::vector<double> pos = bin_positions(s);
    h.perm = morton_order(pos, n);
    std::vector<uint32_t> iperm(n);
    for (uint32_t i = 0; i < n; i++) iperm[h.perm[i]] = i;
    std::vector<Cl> cl;
    std::vector<uint32_t> leaf_beg;
    {
        struct Frame {
            uint32_t b, e;
            int32_t parent;
            int slot;
        };
        // preorder, children in order: leaves numbered left to right
        std::vector<Frame> stack{{0, n, -1, 0}};
        while (!stack.empty()) {
            Frame f = stack.back();
            stack.pop_back();
            Cl c{};
            c.b = f.b;
            c.e = f.e;
            for (int a = 0; a < 3; a++) {
                c.lo[a] = pos[(size_t)h.perm[f.b] * 3 + a];
                c.hi[a] = c.lo[a];
            }
            for (uint32_t i = f.b; i < f.e; i++)
                for (int a = 0; a < 3; a++) {
                    const double v = pos[(size_t)h.perm[i] * 3 + a];
                    c.lo[a] = std::min(c.lo[a], v);
                    c.hi[a] = std::max(c.hi[a], v);
                }
            c.kid[0] = c.kid[1] = -1;
            const int32_t id = (int32_t)cl.size();
            cl.push_back(c);
            if (f.parent >= 0) cl[f.parent].kid[f.slot] = id;
            if (f.e - f.b > p.leaf_size) {
                const uint32_t mid = f.b + (f.e - f.b) / 2;
                stack.push_back({mid, f.e, id, 1});
                stack.push_back({f.b, mid, id, 0});
            }
        }
        // leaf numbering and each cluster's leaf range (children follow parents)
        for (size_t i = 0; i < cl.size(); i++)
            if (cl[i].kid[0] < 0) {
                cl[i].leaf0 = (uint32_t)leaf_beg.size();
                cl[i].leaf1 = cl[i].leaf0 + 1;
                leaf_beg.push_back(cl[i].b);
            }
        leaf_beg.push_back(n);
        for (size_t i = cl.size(); i-- > 0;)
            if (cl[i].kid[0] >= 0) {
                cl[i].leaf0 = cl[cl[i].kid[0]].leaf0;
                cl[i].leaf1 = cl[cl[i].kid[1]].leaf1;
            }
    }
    // the cluster of each leaf in leaf order is increasing in b: leaf_beg sorted
    std::vector<uint4> far;
    std::vector<std::pair<int32_t, int32_t>> farcl;
    uint32_t n_near = 0;
    {
        std::vector<std::pair<int32_t, int32_t>> stack{{0, 0}};
        while (!stack.empty()) {
            auto [ti, si] = stack.back();
            stack.pop_back();
            const Cl &t = cl[ti], &sc = cl[si];
            const double d = dist(t, sc);
            if (d > 0.0 && std::min(diam(t), diam(sc)) <= p.eta * d) {
                far.push_back(make_uint4(t.b, t.e, sc.b, sc.e));
                farcl.push_back({ti, si});
            } else if (t.kid[0] < 0 || sc.kid[0] < 0) {
                n_near++;
            } else {
                // children pairs in order (last pushed runs first)
                for (int a = 1; a >= 0; a--)
                    for (int c = 1; c >= 0; c--) stack.push_back({t.kid[a], sc.kid[c]});
            }
        }
    }
    const uint32_t nfar = (uint32_t)far.size();
    const uint32_t nleaf = (uint32_t)leaf_beg.size() - 1;
    st.n_clusters = (uint32_t)cl.size();
    st.n_leaves = nleaf;
    st.n_far = nfar;
    st.n_near = n_near;
    h.far = far;
    h.nleaf = nleaf;

    // Morton-ordered operator: compressed rows and columns
    const uint64_t nnz = s.sp.nnz;
    uint32_t bits = 0;
    while ((1ull << bits) < n) bits++;
    const uint32_t end_bit = std::max<uint32_t>(1, 2 * bits);
    h.d_perm.alloc(n);
    IRTT_HIP(hipMemcpyAsync(h.d_perm.p, h.perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, stm));
    DBuf<uint32_t> d_iperm, cnt;
    d_iperm.alloc(n);
    cnt.alloc(n);
    IRTT_HIP(hipMemcpyAsync(d_iperm.p, iperm.data(), (size_t)n * 4, hipMemcpyHostToDevice, stm));
    {
        DBuf<unsigned long long> ka, kb;
        DBuf<uint32_t> va, vb;
        DBuf<uint8_t> tmp;
        ka.alloc(nnz);
        kb.alloc(nnz);
        va.alloc(nnz);
        vb.alloc(nnz);
        const size_t tb0 = sort_temp_bytes<unsigned long long, uint32_t>(nnz, end_bit);
        tmp.alloc(std::max<size_t>(tb0, 1));
        for (int form = 0; form < 2; form++) {
            const bool bycol = form == 1;
            k_hm_keys<<<cdiv(n, 4), 256, 0, stm>>>(s.sp.rowptr.p, s.sp.col.p, s.sp.val.p, n,
                                                   d_iperm.p, bits, bycol ? 1u : 0u, ka.p, va.p);
            IRTT_HIP(hipGetLastError());
            size_t tb = tb0;
            rocprim::double_buffer<unsigned long long> kd(ka.p, kb.p);
            rocprim::double_buffer<uint32_t> vd(va.p, vb.p);
            IRTT_HIP(rocprim::radix_sort_pairs(tmp.p, tb, kd, vd, (size_t)nnz, 0, end_bit, stm));
            DBuf<uint32_t> &ptr = bycol ? h.pcp : h.prp;
            DBuf<uint32_t> &idx = bycol ? h.prow : h.pcol;
            DBuf<float> &val = bycol ? h.pcv : h.pval;
            ptr.alloc((size_t)n + 1);
            idx.alloc(nnz);
            val.alloc(nnz);
            IRTT_HIP(hipMemsetAsync(cnt.p, 0, (size_t)n * 4, stm));
            k_hm_count<<<cdiv(nnz, 256), 256, 0, stm>>>(kd.current(), nnz, bits, cnt.p);
            k_hm_scan_u32<<<1, 1024, 0, stm>>>(cnt.p, n, ptr.p);
            k_hm_fill<<<cdiv(nnz, 256), 256, 0, stm>>>(kd.current(), vd.current(), nnz, bits,
                                                       idx.p, val.p);
            IRTT_HIP(hipGetLastError());
        }
    }

    // far blocks: entries, norms, first rows
    DBuf<uint4> d_fb;
    DBuf<uint32_t> d_fnnz, d_i0;
    DBuf<double> d_fro2;
    std::vector<uint32_t> fnnz(nfar);
    std::vector<double> ffro2(nfar);
    if (nfar) {
        d_fb.alloc(nfar);
        d_fnnz.alloc(nfar);
        d_i0.alloc(nfar);
        d_fro2.alloc(nfar);
        IRTT_HIP(hipMemcpyAsync(d_fb.p, far.data(), (size_t)nfar * sizeof(uint4),
                                hipMemcpyHostToDevice, stm));
        k_hm_blockscan<<<nfar, kT, 0, stm>>>(d_fb.p, h.prp.p, h.pcol.p, h.pval.p, d_fnnz.p,
                                            d_fro2.p, d_i0.p);
        IRTT_HIP(hipGetLastError());
        IRTT_HIP(hipMemcpyAsync(fnnz.data(), d_fnnz.p, (size_t)nfar * 4, hipMemcpyDeviceToHost, stm));
        IRTT_HIP(hipMemcpyAsync(ffro2.data(), d_fro2.p, (size_t)nfar * 8, hipMemcpyDeviceToHost,
                                stm));
        IRTT_HIP(hipStreamSynchronize(stm));
    }
    // total |T|_F^2 (host, from the entries)
    double fro_all = 0.0, fro_far = 0.0;
    {
        std::vector<float> v(nnz);
        IRTT_HIP(hipMemcpyAsync(v.data(), s.sp.val.p, nnz * 4, hipMemcpyDeviceToHost, stm));
        IRTT_HIP(hipStreamSynchronize(stm));
        for (float x : v) fro_all += (double)x * (double)x;
    }
    for (uint32_t b = 0; b < nfar; b++) fro_far += ffro2[b];

    // ACA over the far blocks with entries (largest first)
    std::vector<uint32_t> list;
    for (uint32_t b = 0; b < nfar; b++)
        if (fnnz[b]) list.push_back(b);
    std::stable_sort(list.begin(), list.end(), [&](uint32_t a, uint32_t b) {
        return (uint64_t)(far[a].y - far[a].x) + (far[a].w - far[a].z) >
               (uint64_t)(far[b].y - far[b].x) + (far[b].w - far[b].z);
    });
    const uint32_t nl = (uint32_t)list.size();
    const uint32_t kmax = p.max_rank;
    // the FP64 work space of the cross approximations (U, V: (m + n) kmax
    // doubles per block, plus a byte per row) is taken in batches of the
    // sorted list that fit a budget (all at once when the factors are kept);
    // offsets are relative to the batch
    std::vector<unsigned long long> uoff(nl), voff(nl), foff(nl);
    std::vector<uint32_t> batch_beg{0};
    unsigned long long peak_ws = 0;
    {
        const unsigned long long budget =
            p.keep_factors ? ~0ull
            : p.workspace_mib
                ? (unsigned long long)p.workspace_mib << 20
                : std::max<unsigned long long>(
                      1ull << 30, std::min<unsigned long long>(8ull << 30, dev_available() / 4));
        unsigned long long ut = 0, vt = 0, ft = 0;
        for (uint32_t q = 0; q < nl; q++) {
            const uint4 B = far[list[q]];
            const unsigned long long m = B.y - B.x, nn = B.w - B.z;
            const unsigned long long need = (m + nn) * kmax * 8 + m;
            if (q > batch_beg.back() && (ut + vt) * 8 + ft + need > budget) {
                batch_beg.push_back(q);
                ut = vt = ft = 0;
            }
            uoff[q] = ut;
            voff[q] = vt;
            foff[q] = ft;
            ut += m * kmax;
            vt += nn * kmax;
            ft += m;
            peak_ws = std::max(peak_ws, (ut + vt) * 8 + ft);
        }
        batch_beg.push_back(nl);
    }
    st.build_bytes = peak_ws;
    std::vector<uint32_t> rank(nl), conv(nl);
    h.piv.assign((size_t)nl * 2 * kR, -1);
    std::vector<double> err2(nl, 0.0);
    h.status.assign(nfar, IRTT_HBLOCK_ZERO);
    h.rank.assign(nfar, 0);
    h.err_rel.assign(nfar, 0.0);
    h.list_pos.assign(nfar, -1);
    DBuf<uint32_t> d_list, d_rank, d_conv;
    DBuf<unsigned long long> d_uoff, d_voff, d_foff;
    DBuf<int32_t> d_piv;
    DBuf<uint8_t> d_used;
    const double tol_aca = kAcaMargin * p.tol;
    // ACA of the list entries [q0, q1) into the batch work space
    auto run_aca = [&](uint32_t q0, uint32_t q1) {
        const uint4 L = far[list[q1 - 1]];
        unsigned long long ut = uoff[q1 - 1] + (unsigned long long)(L.y - L.x) * kmax;
        unsigned long long vt = voff[q1 - 1] + (unsigned long long)(L.w - L.z) * kmax;
        unsigned long long ft = foff[q1 - 1] + (L.y - L.x);
        h.Ub.alloc(ut);
        h.Vb.alloc(vt);
        d_used.alloc(ft);
        k_hm_aca<<<q1 - q0, kT, 0, stm>>>(d_fb.p, d_list.p + q0, d_i0.p, h.prp.p, h.pcol.p,
                                          h.pval.p, h.pcp.p, h.prow.p, h.pcv.p, d_uoff.p + q0,
                                          d_voff.p + q0, d_foff.p + q0, tol_aca * tol_aca, kmax,
                                          h.Ub.p, h.Vb.p, d_used.p, d_rank.p + q0, d_conv.p + q0,
                                          d_piv.p + (size_t)q0 * 2 * kR);
        IRTT_HIP(hipGetLastError());
    };
    if (nl) {
        d_list.upload(list, stm);
        d_uoff.upload(uoff, stm);
        d_voff.upload(voff, stm);
        d_foff.upload(foff, stm);
        d_rank.alloc(nl);
        d_conv.alloc(nl);
        d_piv.alloc((size_t)nl * 2 * kR);
        h.d_err.alloc(nl);
        for (size_t bt = 0; bt + 1 < batch_beg.size(); bt++) {
            const uint32_t q0 = batch_beg[bt], q1 = batch_beg[bt + 1];
            run_aca(q0, q1);
            // exact error of every block of the batch (blocks that did not
            // stop get one too; they stay sparse)
            k_hm_err<false><<<q1 - q0, kT, 0, stm>>>(
                d_fb.p, d_list.p + q0, d_rank.p + q0, d_uoff.p + q0, d_voff.p + q0, nullptr,
                nullptr, h.Ub.p, h.Vb.p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                d_fro2.p, h.prp.p, h.pcol.p, h.pval.p, h.d_err.p + q0);
            IRTT_HIP(hipGetLastError());
        }
        IRTT_HIP(hipMemcpyAsync(rank.data(), d_rank.p, (size_t)nl * 4, hipMemcpyDeviceToHost, stm));
        IRTT_HIP(hipMemcpyAsync(conv.data(), d_conv.p, (size_t)nl * 4, hipMemcpyDeviceToHost, stm));
        IRTT_HIP(hipMemcpyAsync(h.piv.data(), d_piv.p, (size_t)nl * 2 * kR * 4,
                                hipMemcpyDeviceToHost, stm));
        IRTT_HIP(hipMemcpyAsync(err2.data(), h.d_err.p, (size_t)nl * 8, hipMemcpyDeviceToHost, stm));
        IRTT_HIP(hipStreamSynchronize(stm));
    }

    // decisions
    const size_t fb_bytes = p.precision == IRTT_STORE_F16 ? 2 : 4;
    std::vector<uint32_t> acc;  // list positions kept in low rank
    uint64_t nnz_lr = 0, bytes_lr = 0, rank_sum = 0;
    for (uint32_t q = 0; q < nl; q++) {
        const uint32_t b = list[q];
        const uint4 B = far[b];
        const uint64_t mn = (uint64_t)(B.y - B.x) + (B.w - B.z);
        h.rank[b] = rank[q];
        h.list_pos[b] = (int32_t)q;
        h.err_rel[b] = ffro2[b] > 0.0 ? std::sqrt(err2[q] / ffro2[b]) : 0.0;
        const uint64_t lrb = mn * rank[q] * fb_bytes + (p.precision == IRTT_STORE_F16 ? mn * 4 : 0);
        if (!conv[q]) {
            h.status[b] = IRTT_HBLOCK_NOCONV;
            st.n_noconv++;
        } else if (h.err_rel[b] > p.tol || (!p.allow_larger && lrb >= 8ull * fnnz[b])) {
            h.status[b] = IRTT_HBLOCK_REJECTED;
            st.n_reject++;
        } else {
            h.status[b] = IRTT_HBLOCK_LOWRANK;
            acc.push_back(q);
            nnz_lr += fnnz[b];
            bytes_lr += lrb;
            rank_sum += rank[q];
            st.max_rank_used = std::max(st.max_rank_used, rank[q]);
        }
    }
    st.n_zero = nfar - nl;
    st.n_lowrank = (uint32_t)acc.size();
    st.mean_rank = acc.empty() ? 0.0 : (double)rank_sum / (double)acc.size();

    // stored factors of the accepted blocks
    const uint32_t na = (uint32_t)acc.size();
    h.nlr = na;
    h.precision = p.precision;
    std::vector<uint32_t> a_rk(na), a_r0(na), a_c0(na), a_m(na), a_n(na), a_lpos(na);
    std::vector<unsigned long long> a_uo(na), a_vo(na), a_suo(na), a_svo(na);
    unsigned long long fu = 0, fv = 0, su = 0, sv = 0;
    for (uint32_t a = 0; a < na; a++) {
        const uint32_t q = acc[a];
        const uint4 B = far[list[q]];
        a_rk[a] = rank[q];
        a_r0[a] = B.x;
        a_c0[a] = B.z;
        a_m[a] = B.y - B.x;
        a_n[a] = B.w - B.z;
        a_lpos[a] = q;
        a_uo[a] = fu;
        a_vo[a] = fv;
        a_suo[a] = su;
        a_svo[a] = sv;
        fu += (unsigned long long)a_m[a] * a_rk[a];
        fv += (unsigned long long)a_n[a] * a_rk[a];
        su += a_m[a];
        sv += a_n[a];
    }
    double err2_stored_far = 0.0, err2_far = 0.0;
    for (uint32_t a = 0; a < na; a++) err2_far += err2[acc[a]];
    if (na) {
        h.d_rk.upload(a_rk, stm);
        h.d_r0.upload(a_r0, stm);
        h.d_c0.upload(a_c0, stm);
        h.d_uo.upload(a_uo, stm);
        h.d_vo.upload(a_vo, stm);
        h.d_suo.upload(a_suo, stm);
        h.d_svo.upload(a_svo, stm);
        DBuf<uint32_t> d_lpos, d_m, d_n;
        d_lpos.upload(a_lpos, stm);
        d_m.upload(a_m, stm);
        d_n.upload(a_n, stm);
        const bool f16 = p.precision == IRTT_STORE_F16;
        if (f16) {
            h.uh.alloc(fu);
            h.vh.alloc(fv);
            h.su.alloc(su);
            h.sv.alloc(sv);
        } else {
            h.uf.alloc(fu);
            h.vf.alloc(fv);
        }
        // pack batch by batch: a batch's factors are rebuilt (the same
        // deterministic steps) unless it is the one still in the work space
        const size_t nbt = batch_beg.size() - 1;
        size_t in_ws = nbt - 1;  // the batch whose factors the work space holds
        size_t a0 = 0;
        for (size_t bt = 0; bt < nbt; bt++) {
            const uint32_t q0 = batch_beg[bt], q1 = batch_beg[bt + 1];
            size_t a1 = a0;
            while (a1 < na && acc[a1] < q1) a1++;
            if (a1 == a0) continue;
            if (bt != in_ws) {
                run_aca(q0, q1);
                in_ws = bt;
            }
            uint32_t mmax = 0, nmax = 0;
            for (size_t a = a0; a < a1; a++) {
                mmax = std::max(mmax, a_m[a]);
                nmax = std::max(nmax, a_n[a]);
            }
            const uint32_t cnt = (uint32_t)(a1 - a0);
            k_hm_pack<<<dim3(cdiv(mmax, kT), cnt), kT, 0, stm>>>(
                d_lpos.p + a0, d_m.p + a0, h.d_rk.p + a0, d_uoff.p, h.d_uo.p + a0,
                h.d_suo.p + a0, h.Ub.p, f16 ? 1u : 0u, reinterpret_cast<__half *>(h.uh.p), h.uf.p,
                h.su.p);
            k_hm_pack<<<dim3(cdiv(nmax, kT), cnt), kT, 0, stm>>>(
                d_lpos.p + a0, d_n.p + a0, h.d_rk.p + a0, d_voff.p, h.d_vo.p + a0,
                h.d_svo.p + a0, h.Vb.p, f16 ? 1u : 0u, reinterpret_cast<__half *>(h.vh.p), h.vf.p,
                h.sv.p);
            IRTT_HIP(hipGetLastError());
            a0 = a1;
        }
        // exact error with the stored factors
        std::vector<uint32_t> alist(na);
        for (uint32_t a = 0; a < na; a++) alist[a] = list[acc[a]];
        DBuf<uint32_t> d_alist;
        d_alist.upload(alist, stm);
        DBuf<double> d_e2;
        d_e2.alloc(na);
        k_hm_err<true><<<na, kT, 0, stm>>>(
            d_fb.p, d_alist.p, h.d_rk.p, h.d_uo.p, h.d_vo.p, h.d_suo.p, h.d_svo.p, nullptr, nullptr,
            f16 ? reinterpret_cast<const __half *>(h.uh.p) : nullptr,
            f16 ? reinterpret_cast<const __half *>(h.vh.p) : nullptr, f16 ? nullptr : h.uf.p,
            f16 ? nullptr : h.vf.p, f16 ? h.su.p : nullptr, f16 ? h.sv.p : nullptr, d_fro2.p,
            h.prp.p, h.pcol.p, h.pval.p, d_e2.p);
        IRTT_HIP(hipGetLastError());
        std::vector<double> e2s(na);
        IRTT_HIP(hipMemcpyAsync(e2s.data(), d_e2.p, (size_t)na * 8, hipMemcpyDeviceToHost, stm));
        IRTT_HIP(hipStreamSynchronize(stm));
        for (double x : e2s) err2_stored_far += x;
    }

    // per-leaf lists of the accepted blocks: by row cluster (sorted by c0)
    // and by column cluster (sorted by r0)
    {
        std::vector<std::vector<uint32_t>> rl(nleaf), cll(nleaf);
        for (uint32_t a = 0; a < na; a++) {
            const auto [ti, si] = farcl[list[acc[a]]];
            for (uint32_t L = cl[ti].leaf0; L < cl[ti].leaf1; L++) rl[L].push_back(a);
            for (uint32_t L = cl[si].leaf0; L < cl[si].leaf1; L++) cll[L].push_back(a);
        }
        std::vector<uint32_t> rptr(nleaf + 1, 0), cptr(nleaf + 1, 0), rblk, cblk;
        std::vector<uint2> riv, civ;
        for (uint32_t L = 0; L < nleaf; L++) {
            std::sort(rl[L].begin(), rl[L].end(),
                      [&](uint32_t x, uint32_t y) { return a_c0[x] < a_c0[y]; });
            std::sort(cll[L].begin(), cll[L].end(),
                      [&](uint32_t x, uint32_t y) { return a_r0[x] < a_r0[y]; });
            for (uint32_t a : rl[L]) {
                rblk.push_back(a);
                riv.push_back(make_uint2(a_c0[a], a_c0[a] + a_n[a]));
            }
            for (uint32_t a : cll[L]) {
                cblk.push_back(a);
                civ.push_back(make_uint2(a_r0[a], a_r0[a] + a_m[a]));
            }
            rptr[L + 1] = (uint32_t)rblk.size();
            cptr[L + 1] = (uint32_t)cblk.size();
        }
        std::vector<uint32_t> leaf_of(n);
        for (uint32_t L = 0; L < nleaf; L++)
            for (uint32_t i = leaf_beg[L]; i < leaf_beg[L + 1]; i++) leaf_of[i] = L;
        h.d_leaf_beg.upload(leaf_beg, stm);
        h.d_rptr.upload(rptr, stm);
        h.d_cptr.upload(cptr, stm);
        if (rblk.empty()) rblk.push_back(0), riv.push_back(make_uint2(0, 0));
        if (cblk.empty()) cblk.push_back(0), civ.push_back(make_uint2(0, 0));
        h.d_rblk.upload(rblk, stm);
        h.d_cblk.upload(cblk, stm);
        DBuf<uint2> d_riv, d_civ;
        DBuf<uint32_t> d_leaf_of;
        d_riv.upload(riv, stm);
        d_civ.upload(civ, stm);
        d_leaf_of.upload(leaf_of, stm);
        // near field in rows and in columns
        for (int form = 0; form < 2; form++) {
            const bool bycol = form == 1;
            const uint32_t *ptr = bycol ? h.pcp.p : h.prp.p;
            const uint32_t *idx = bycol ? h.prow.p : h.pcol.p;
            const float *val = bycol ? h.pcv.p : h.pval.p;
            const uint32_t *lp = bycol ? h.d_cptr.p : h.d_rptr.p;
            const uint2 *iv = bycol ? d_civ.p : d_riv.p;
            k_hm_near_count<<<cdiv(n, 4), 256, 0, stm>>>(ptr, idx, n, d_leaf_of.p, lp, iv, cnt.p);
            DBuf<uint32_t> &nptr = bycol ? h.ncp : h.nrp;
            nptr.alloc((size_t)n + 1);
            k_hm_scan_u32<<<1, 1024, 0, stm>>>(cnt.p, n, nptr.p);
            uint32_t nn = 0;
            IRTT_HIP(hipMemcpyAsync(&nn, nptr.p + n, 4, hipMemcpyDeviceToHost, stm));
            IRTT_HIP(hipStreamSynchronize(stm));
            DBuf<uint32_t> &nidx = bycol ? h.nrow : h.ncol;
            DBuf<float> &nval = bycol ? h.ncv : h.nval;
            nidx.alloc(std::max<uint32_t>(nn, 1));
            nval.alloc(std::max<uint32_t>(nn, 1));
            k_hm_near_fill<<<cdiv(n, 4), 256, 0, stm>>>(ptr, idx, val, n, d_leaf_of.p, lp, iv,
                                                        nptr.p, nidx.p, nval.p);
            IRTT_HIP(hipGetLastError());
            if (!bycol) h.nnz_near = nn;
        }
    }
    // first-stage items of both product directions (over V^T rows for
    // H x, over U rows for H^T x)
    {
        std::vector<uint4> itn, itt;
        std::vector<uint32_t> ipn(na + 1, 0), ipt(na + 1, 0);
        for (uint32_t a = 0; a < na; a++) {
            for (uint32_t j = 0; j < a_n[a]; j += kItem)
                itn.push_back(make_uint4(a, j, std::min(a_n[a], j + kItem), a_c0[a]));
            for (uint32_t i = 0; i < a_m[a]; i += kItem)
                itt.push_back(make_uint4(a, i, std::min(a_m[a], i + kItem), a_r0[a]));
            ipn[a + 1] = (uint32_t)itn.size();
            ipt[a + 1] = (uint32_t)itt.size();
        }
        h.n_itn = (uint32_t)itn.size();
        h.n_itt = (uint32_t)itt.size();
        if (na) {
            h.d_itn.upload(itn, stm);
            h.d_itt.upload(itt, stm);
            h.d_ipn.upload(ipn, stm);
            h.d_ipt.upload(ipt, stm);
            h.tpart.alloc((size_t)std::max(h.n_itn, h.n_itt) * kR);
            h.tq.alloc((size_t)na * kR);
        }
    }
    h.xp.alloc(n);
    h.yp.alloc(n);
    IRTT_HIP(hipStreamSynchronize(stm));
    if (!p.keep_factors) {
        h.Ub.release();
        h.Vb.release();
        h.piv.clear();
    }
    h.keep = p.keep_factors != 0;
    h.list_uoff = uoff;
    h.list_voff = voff;
    h.valid = true;

    st.nnz = nnz;
    st.nnz_near = h.nnz_near;
    st.nnz_lowrank = nnz_lr;
    st.bytes_sparse = nnz * 8 + ((uint64_t)n + 1) * 4;
    st.bytes_lowrank = bytes_lr;
    st.bytes_hmatrix = h.nnz_near * 8 + ((uint64_t)n + 1) * 4 + bytes_lr;
    st.err_frob_rel = fro_all > 0.0 ? std::sqrt(err2_far / fro_all) : 0.0;
    st.err_frob_rel_stored = fro_all > 0.0 ? std::sqrt(err2_stored_far / fro_all) : 0.0;
    st.fro_far_rel = fro_all > 0.0 ? std::sqrt(fro_far / fro_all) : 0.0;
    st.build_ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                                   t_start).count();
    h.bytes = st.bytes_hmatrix;
}

// y = H w (or H^T w)
void hm_apply(Scene &s, const double *w, double *y, bool transposed)
{
    Hmat &h = s.hm;
    if (!h.valid) throw Error(IRTT_ERR_STATE, "no H-matrix operator (irtt_compress_hierarchical)");
    const uint32_t n = s.nb;
    hipStream_t st = s.stream;
    const uint32_t g = cdiv(n, 256);
    k_hm_gather<<<g, 256, 0, st>>>(w, h.d_perm.p, n, h.xp.p);
    k_hm_spmv<<<cdiv(n, 4), 256, 0, st>>>(transposed ? h.ncp.p : h.nrp.p,
                                          transposed ? h.nrow.p : h.ncol.p,
                                          transposed ? h.ncv.p : h.nval.p, n, h.xp.p, h.yp.p);
    if (h.nlr) {
        const bool f16 = h.precision == IRTT_STORE_F16;
        const __half *HU = reinterpret_cast<const __half *>(h.uh.p);
        const __half *HV = reinterpret_cast<const __half *>(h.vh.p);
        // stage 1 runs over the factor whose rows meet x: V^T for H x, U for H^T x
        const uint32_t nit = transposed ? h.n_itt : h.n_itn;
        const uint4 *items = transposed ? h.d_itt.p : h.d_itn.p;
        const unsigned long long *fo1 = transposed ? h.d_uo.p : h.d_vo.p;
        const unsigned long long *so1 = transposed ? h.d_suo.p : h.d_svo.p;
        const __half *H1 = transposed ? HU : HV;
        const float *F1 = transposed ? h.uf.p : h.vf.p;
        const float *S1 = transposed ? h.su.p : h.sv.p;
        if (f16)
            k_hm_stage1<true><<<nit, kT, 0, st>>>(items, h.d_rk.p, fo1, so1, H1, F1, S1, h.xp.p,
                                                   h.tpart.p);
        else
            k_hm_stage1<false><<<nit, kT, 0, st>>>(items, h.d_rk.p, fo1, so1, H1, F1, S1, h.xp.p,
                                                    h.tpart.p);
        k_hm_stage1_sum<<<cdiv((uint64_t)h.nlr * kR, 256), 256, 0, st>>>(
            transposed ? h.d_ipt.p : h.d_ipn.p, h.nlr, h.tpart.p, h.tq.p);
        const unsigned long long *fo2 = transposed ? h.d_vo.p : h.d_uo.p;
        const unsigned long long *so2 = transposed ? h.d_svo.p : h.d_suo.p;
        const __half *H2 = transposed ? HV : HU;
        const float *F2 = transposed ? h.vf.p : h.uf.p;
        const float *S2 = transposed ? h.sv.p : h.su.p;
        const uint32_t *lp = transposed ? h.d_cptr.p : h.d_rptr.p;
        const uint32_t *lb = transposed ? h.d_cblk.p : h.d_rblk.p;
        const uint32_t *base = transposed ? h.d_c0.p : h.d_r0.p;
        if (f16)
            k_hm_stage2<true><<<cdiv(h.nleaf, 4), 256, 0, st>>>(h.d_leaf_beg.p, h.nleaf, lp, lb,
                                                                 base, h.d_rk.p, fo2, so2, H2, F2,
                                                                 S2, h.tq.p, h.yp.p);
        else
            k_hm_stage2<false><<<cdiv(h.nleaf, 4), 256, 0, st>>>(h.d_leaf_beg.p, h.nleaf, lp, lb,
                                                                  base, h.d_rk.p, fo2, so2, H2, F2,
                                                                  S2, h.tq.p, h.yp.p);
    }
    k_hm_scatter<<<g, 256, 0, st>>>(h.yp.p, h.d_perm.p, n, y);
    IRTT_HIP(hipGetLastError());
}

void hm_get(Scene &s, uint32_t *perm, uint32_t *n_far, uint32_t *blocks, double *err_rel,
            int32_t *pivots)
{
    const Hmat &h = s.hm;
    if (!h.valid) throw Error(IRTT_ERR_STATE, "no H-matrix operator (irtt_compress_hierarchical)");
    if (n_far) *n_far = (uint32_t)h.far.size();
    if (perm) std::copy(h.perm.begin(), h.perm.end(), perm);
    for (size_t b = 0; b < h.far.size(); b++) {
        if (blocks) {
            uint32_t *o = blocks + b * 6;
            o[0] = h.far[b].x;
            o[1] = h.far[b].y;
            o[2] = h.far[b].z;
            o[3] = h.far[b].w;
            o[4] = h.rank[b];
            o[5] = h.status[b];
        }
        if (err_rel) err_rel[b] = h.err_rel[b];
        if (pivots) {
            const int32_t q = h.list_pos[b];
            for (uint32_t l = 0; l < 2 * kR; l++)
                pivots[b * 2 * kR + l] = (q >= 0 && !h.piv.empty()) ? h.piv[(size_t)q * 2 * kR + l] : -1;
        }
    }
}

void hm_get_factors(Scene &s, uint32_t b, double *U, double *V)
{
    const Hmat &h = s.hm;
    if (!h.valid) throw Error(IRTT_ERR_STATE, "no H-matrix operator (irtt_compress_hierarchical)");
    if (!h.keep) throw Error(IRTT_ERR_STATE, "factors not kept (irtt_hmatrix_params.keep_factors)");
    if (b >= h.far.size()) throw Error(IRTT_ERR_ARG, "block index out of range");
    const int32_t q = h.list_pos[b];
    const uint32_t k = h.rank[b];
    if (q < 0 || k == 0) return;
    const uint4 B = h.far[b];
    const uint32_t m = B.y - B.x, n = B.w - B.z;
    for (uint32_t l = 0; l < k; l++) {
        if (U)
            IRTT_HIP(hipMemcpyAsync(U + (size_t)l * m, h.Ub.p + h.list_uoff[q] + (size_t)l * m,
                                    (size_t)m * 8, hipMemcpyDeviceToHost, s.stream));
        if (V)
            IRTT_HIP(hipMemcpyAsync(V + (size_t)l * n, h.Vb.p + h.list_voff[q] + (size_t)l * n,
                                    (size_t)n * 8, hipMemcpyDeviceToHost, s.stream));
    }
    IRTT_HIP(hipStreamSynchronize(s.stream));
}

}  // namespace irtt
