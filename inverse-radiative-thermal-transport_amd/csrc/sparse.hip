// sparse.hip -- the transport operator in compressed rows and columns.
//
// The reference stores T densely (thermal_transport.cpp:612-660) and caps it
// at 46 340 bins (:143).  A Monte-Carlo operator reaches only the (absorber,
// emitter) pairs some photon connected: with ~1300 photons per triangle over
// 36k bins the city block's operator is a few percent full, and the 400k-
// triangle scenes cannot be held densely at all.  Here the photon kernel adds
// each absorbed photon's fixed-point weight into a device hash table keyed by
// (absorbing bin, emitting bin) (sparse_table.hpp); the occupied slots are
// then compacted, radix-sorted by key and converted to FP32 exactly as the
// dense finalize converts (T_ij = delta_ij sigma eps_i A_i - acc_ij Q), so
// every stored entry equals the dense layout's bit for bit.  A second sort by
// (column, row) gives the compressed columns used by T^T products.  Every
// step is integer arithmetic or a per-entry conversion, so the stored
// operator does not depend on the order in which photons landed.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <vector>

#include "irtt_internal.hpp"
#include "sparse_table.hpp"

namespace irtt {
namespace {

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

__global__ void k_tab_clear(TabSlot *tab, uint64_t n)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        tab[i].key = kTabEmpty;
        tab[i].val = 0ull;
    }
}

// every bin's diagonal entry exists (sigma eps A even when no photon returned)
__global__ void k_tab_diag(TabSlot *tab, uint32_t lg, uint32_t bits, uint32_t nb,
                           uint32_t *overflow)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) tab_add(tab, lg, tab_key(i, i, bits), 0ull, overflow);
}

// The table passes give each 256-thread block one chunk of kChunk slots
// (thread t reads slots t, t + 256, ...): the occupied slots of a chunk are
// counted, the chunk counts scanned, and each chunk written at its offset,
// so no global atomics are needed.
constexpr uint32_t kChunkIters = 16;
constexpr uint32_t kChunk = 256 * kChunkIters;

__global__ __launch_bounds__(256) void k_tab_count(const TabSlot *__restrict__ tab, uint64_t n,
                                                   uint32_t *__restrict__ ccount)
{
    __shared__ uint32_t ws[4];
    const uint64_t base = (uint64_t)blockIdx.x * kChunk + threadIdx.x;
    uint32_t c = 0;
    for (uint32_t j = 0; j < kChunkIters; j++) {
        const uint64_t i = base + (uint64_t)j * 256;
        if (i < n && tab[i].key != kTabEmpty) c++;
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) ccount[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// exclusive scan of m chunk counts into 64-bit offsets, total in off[m]; one
// block of 1024 threads
__global__ __launch_bounds__(1024) void k_scan_chunks(const uint32_t *__restrict__ c, uint32_t m,
                                                      unsigned long long *__restrict__ off)
{
    __shared__ unsigned long long sums[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (m + 1023u) / 1024u;
    const uint32_t b = min(tid * per, m), e = min(b + per, m);
    unsigned long long s = 0;
    for (uint32_t i = b; i < e; i++) s += c[i];
    sums[tid] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const unsigned long long v = tid >= o ? sums[tid - o] : 0ull;
        __syncthreads();
        sums[tid] += v;
        __syncthreads();
    }
    unsigned long long run = sums[tid] - s;
    for (uint32_t i = b; i < e; i++) {
        off[i] = run;
        run += c[i];
    }
    if (tid == 1023) off[m] = sums[1023];
}

// occupied slots of each chunk -> (key, value) at the chunk's offset, in
// slot order (the slot of a key depends on the insertion race, so the
// entries are sorted next)
__global__ __launch_bounds__(256) void k_tab_compact(const TabSlot *__restrict__ tab, uint64_t n,
                                                     const unsigned long long *__restrict__ coff,
                                                     unsigned long long *__restrict__ keys,
                                                     unsigned long long *__restrict__ vals)
{
    __shared__ uint32_t wc[4];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * kChunk + threadIdx.x;
    unsigned long long run = coff[blockIdx.x];
    for (uint32_t j = 0; j < kChunkIters; j++) {
        const uint64_t i = base + (uint64_t)j * 256;
        TabSlot t{kTabEmpty, 0ull};
        if (i < n) t = tab[i];
        const bool occ = t.key != kTabEmpty;
        const unsigned long long m = __ballot(occ);
        if (lane == 0) wc[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t w = 0; w < 4; w++) {
            const uint32_t x = wc[w];
            before += w < wave ? x : 0u;
            total += x;
        }
        if (occ) {
            const unsigned long long o = run + before + __popcll(m & ((1ull << lane) - 1ull));
            keys[o] = t.key;
            vals[o] = t.val;
        }
        run += total;
        __syncthreads();
    }
}

// sorted (row, column) keys -> compressed rows; FP32 entry as the dense
// finalize computes it
__global__ void k_csr_fill(const unsigned long long *__restrict__ keys,
                           const unsigned long long *__restrict__ vals, uint64_t nnz,
                           uint32_t bits, uint32_t nb, const double *__restrict__ diag, double Q,
                           uint32_t *__restrict__ col, float *__restrict__ val,
                           uint32_t *__restrict__ rowptr, float *__restrict__ tdiag)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    const unsigned long long key = keys[i];
    const uint32_t r = (uint32_t)(key >> bits), c = (uint32_t)(key & ((1ull << bits) - 1ull));
    const double v = (r == c ? diag[r] : 0.0) - (double)vals[i] * Q;
    const float f = (float)v;
    col[i] = c;
    val[i] = f;
    if (r == c) tdiag[r] = f;
    if (i == 0 || (uint32_t)(keys[i - 1] >> bits) != r) rowptr[r] = (uint32_t)i;
    if (i == nnz - 1) rowptr[nb] = (uint32_t)nnz;
}

__global__ void k_csc_keys(const unsigned long long *__restrict__ keys,
                           const unsigned long long *__restrict__ acc, uint64_t nnz,
                           uint32_t bits, unsigned long long *__restrict__ k2,
                           unsigned long long *__restrict__ a2)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    const unsigned long long key = keys[i];
    const unsigned long long cmask = (1ull << bits) - 1ull;
    k2[i] = ((key & cmask) << bits) | (key >> bits);
    a2[i] = acc[i];
}

// sorted (column, row) keys -> compressed columns, entries converted as in
// k_csr_fill (hence equal to the row form's)
__global__ void k_csc_fill(const unsigned long long *__restrict__ k2,
                           const unsigned long long *__restrict__ a2, uint64_t nnz, uint32_t bits,
                           uint32_t nb, const double *__restrict__ diag, double Q,
                           uint32_t *__restrict__ row, float *__restrict__ cval,
                           uint32_t *__restrict__ colptr)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    const unsigned long long key = k2[i];
    const uint32_t c = (uint32_t)(key >> bits), r = (uint32_t)(key & ((1ull << bits) - 1ull));
    row[i] = r;
    cval[i] = (float)((r == c ? diag[c] : 0.0) - (double)a2[i] * Q);
    if (i == 0 || (uint32_t)(k2[i - 1] >> bits) != c) colptr[c] = (uint32_t)i;
    if (i == nnz - 1) colptr[nb] = (uint32_t)nnz;
}

// integer column sums of the accumulators, one wave per column
__global__ __launch_bounds__(256) void k_col_sums(const uint32_t *__restrict__ colptr,
                                                  const unsigned long long *__restrict__ a2,
                                                  uint32_t nb, unsigned long long *__restrict__ colacc)
{
    const uint32_t c = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (c >= nb) return;
    unsigned long long s = 0;
    for (uint32_t k = colptr[c] + lane; k < colptr[c + 1]; k += 64) s += a2[k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) colacc[c] = s;
}

// y_i = sum_k val_k u_{idx_k} over the entries of row i (or column i of the
// compressed columns): one wave per row, FP64 accumulation, fixed-order
// wave reduction, so the result is reproducible.
__global__ __launch_bounds__(256) void k_spmv(const uint32_t *__restrict__ ptr,
                                              const uint32_t *__restrict__ idx,
                                              const float *__restrict__ val, uint32_t n,
                                              const double *__restrict__ u, double *__restrict__ y)
{
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (r >= n) return;
    const uint32_t b = ptr[r], e = ptr[r + 1];
    // four entries per lane in flight (their index loads, then their
    // gathers), summed in four partial sums in a fixed order
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    uint32_t k = b + lane;
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

// X (rc vectors of pitch ld) -> Xt (n x RP, vector index fastest; zeros past rc)
__global__ void k_interleave(const double *__restrict__ X, uint32_t ld, uint32_t n, uint32_t rc,
                             uint32_t RP, double *__restrict__ Xt)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * RP) return;
    const uint32_t c = (uint32_t)(i / RP), r = (uint32_t)(i - (uint64_t)c * RP);
    Xt[i] = r < rc ? X[(size_t)r * ld + c] : 0.0;
}

// Y[r][i] = sum_k val_k Xt[idx_k][r] for RP vectors: one wave per row, lane
// (s, r) = (lane / RP, lane % RP) takes entries s, s + 64/RP, ... so one
// gather of Xt reads RP consecutive doubles.
template <int RP>
__global__ __launch_bounds__(256) void k_spmm(const uint32_t *__restrict__ ptr,
                                              const uint32_t *__restrict__ idx,
                                              const float *__restrict__ val, uint32_t n,
                                              const double *__restrict__ Xt, uint32_t rc,
                                              uint32_t ld, double *__restrict__ Y)
{
    constexpr uint32_t S = 64 / RP;
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (r >= n) return;
    const uint32_t v = lane % RP, slot = lane / RP;
    const uint32_t b = ptr[r], e = ptr[r + 1];
    // four entries per lane in flight, four partial sums in a fixed order
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    uint32_t k = b + slot;
    for (; k + 3 * S < e; k += 4 * S) {
        const uint32_t i0 = idx[k], i1 = idx[k + S], i2 = idx[k + 2 * S], i3 = idx[k + 3 * S];
        const float v0 = val[k], v1 = val[k + S], v2 = val[k + 2 * S], v3 = val[k + 3 * S];
        a0 += (double)v0 * Xt[(size_t)i0 * RP + v];
        a1 += (double)v1 * Xt[(size_t)i1 * RP + v];
        a2 += (double)v2 * Xt[(size_t)i2 * RP + v];
        a3 += (double)v3 * Xt[(size_t)i3 * RP + v];
    }
    for (; k < e; k += S) a0 += (double)val[k] * Xt[(size_t)idx[k] * RP + v];
    double a = (a0 + a1) + (a2 + a3);
    for (int off = 32; off >= RP; off >>= 1) a += __shfl_down(a, off, 64);
    if (slot == 0 && v < rc) Y[(size_t)v * ld + r] = a;
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

}  // namespace

// Hash table for the photons of one build: sized to 5/4 of (photons + bins)
// rounded up to a power of two, an upper bound of the distinct keys, so the
// load factor stays below 0.8 even if every photon reached a new pair.
void sparse_table_begin(Scene &s, uint64_t photons, PhotonSink &sink)
{
    const uint32_t nb = s.nb;
    uint32_t bits = 0;
    while ((1ull << bits) < nb) bits++;
    const uint64_t keys = photons + nb;
    const uint64_t want = keys + keys / 4;
    uint32_t lg = 10;
    while (lg < 63 && (1ull << lg) < want) lg++;
    if (lg > 36)
        throw Error(IRTT_ERR_LIMIT, "sparse transport: " + std::to_string(photons) +
                                        " photons need a hash table above 2^36 slots");
    hipStream_t stm = s.stream;
    s.sp.invalidate();
    const uint64_t cap = 1ull << lg;
    s.sp.tab.alloc(cap);
    s.sp.flag.alloc(1);
    IRTT_HIP(hipMemsetAsync(s.sp.flag.p, 0, sizeof(uint32_t), stm));
    k_tab_clear<<<(uint32_t)std::min<uint64_t>(cdiv(cap, 256), 65536), 256, 0, stm>>>(s.sp.tab.p,
                                                                                     cap);
    k_tab_diag<<<cdiv(nb, 256), 256, 0, stm>>>(s.sp.tab.p, lg, bits, nb, s.sp.flag.p);
    IRTT_HIP(hipGetLastError());
    sink.acc = nullptr;
    sink.tab = s.sp.tab.p;
    sink.lg = lg;
    sink.key_bits = bits;
    sink.overflow = s.sp.flag.p;
}

// Table -> compressed rows and columns.  Needs s.d_diag and s.quantum;
// fills s.d_tdiag and s.d_colacc (exact integer column sums).
void sparse_table_finish(Scene &s, const PhotonSink &sink)
{
    const uint32_t nb = s.nb, bits = sink.key_bits;
    const uint64_t cap = 1ull << sink.lg;
    hipStream_t stm = s.stream;
    const uint64_t nchunk64 = (cap + kChunk - 1) / kChunk;
    if (nchunk64 > 0x7fffffffull) throw Error(IRTT_ERR_LIMIT, "sparse transport: table too large");
    const uint32_t nchunk = (uint32_t)nchunk64;
    Sparse &p = s.sp;
    DBuf<uint32_t> &ccount = p.ccount;
    DBuf<unsigned long long> &coff = p.coff;
    ccount.alloc(nchunk);
    coff.alloc((size_t)nchunk + 1);
    k_tab_count<<<nchunk, 256, 0, stm>>>(s.sp.tab.p, cap, ccount.p);
    k_scan_chunks<<<1, 1024, 0, stm>>>(ccount.p, nchunk, coff.p);
    IRTT_HIP(hipGetLastError());
    unsigned long long nnz = 0;
    uint32_t overflow = 0;
    IRTT_HIP(hipMemcpyAsync(&nnz, coff.p + nchunk, sizeof(nnz), hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipMemcpyAsync(&overflow, s.sp.flag.p, sizeof(overflow), hipMemcpyDeviceToHost, stm));
    IRTT_HIP(hipStreamSynchronize(stm));
    if (overflow) throw Error(IRTT_ERR_LIMIT, "sparse transport: hash table overflow");
    if (nnz >= 0xffffffffull)
        throw Error(IRTT_ERR_LIMIT, "sparse transport: " + std::to_string(nnz) +
                                        " entries exceed 32-bit offsets");
    const uint32_t end_bit = std::max<uint32_t>(1, 2 * bits);

    // ka/va: compacted, kb/vb: the sort's other halves; the column sort
    // reuses all four
    DBuf<unsigned long long> &ka = p.ka, &kb = p.kb, &va = p.va, &vb = p.vb;
    ka.alloc(nnz);
    kb.alloc(nnz);
    va.alloc(nnz);
    vb.alloc(nnz);
    k_tab_compact<<<nchunk, 256, 0, stm>>>(s.sp.tab.p, cap, coff.p, ka.p, va.p);
    IRTT_HIP(hipGetLastError());

    const size_t t1 = sort_temp_bytes<unsigned long long, unsigned long long>(nnz, end_bit);
    DBuf<uint8_t> &tmp = p.sort_tmp;
    tmp.alloc(std::max<size_t>(t1, 1));
    size_t tb = t1;
    rocprim::double_buffer<unsigned long long> kd(ka.p, kb.p);
    rocprim::double_buffer<unsigned long long> vd(va.p, vb.p);
    IRTT_HIP(rocprim::radix_sort_pairs(tmp.p, tb, kd, vd, (size_t)nnz, 0, end_bit, stm));

    p.rowptr.alloc((size_t)nb + 1);
    p.col.alloc(nnz);
    p.val.alloc(nnz);
    s.d_tdiag.alloc(nb);
    const uint32_t ge = cdiv(nnz, 256);
    k_csr_fill<<<ge, 256, 0, stm>>>(kd.current(), vd.current(), nnz, bits, nb, s.d_diag.p,
                                    s.quantum, p.col.p, p.val.p, p.rowptr.p, s.d_tdiag.p);
    IRTT_HIP(hipGetLastError());

    // compressed columns: the same entries keyed (column, row), written into
    // the halves the row sort left free; the row-sorted halves are free once
    // k_csc_keys has read them and serve as the second sort's other halves
    unsigned long long *k2a = kd.current() == ka.p ? kb.p : ka.p;
    unsigned long long *a2a = vd.current() == va.p ? vb.p : va.p;
    k_csc_keys<<<ge, 256, 0, stm>>>(kd.current(), vd.current(), nnz, bits, k2a, a2a);
    IRTT_HIP(hipGetLastError());
    rocprim::double_buffer<unsigned long long> kd2(k2a, kd.current());
    rocprim::double_buffer<unsigned long long> ad2(a2a, vd.current());
    tb = t1;
    IRTT_HIP(rocprim::radix_sort_pairs(tmp.p, tb, kd2, ad2, (size_t)nnz, 0, end_bit, stm));
    p.colptr.alloc((size_t)nb + 1);
    p.row.alloc(nnz);
    p.cval.alloc(nnz);
    k_csc_fill<<<ge, 256, 0, stm>>>(kd2.current(), ad2.current(), nnz, bits, nb, s.d_diag.p,
                                    s.quantum, p.row.p, p.cval.p, p.colptr.p);
    k_col_sums<<<cdiv(nb, 4), 256, 0, stm>>>(p.colptr.p, ad2.current(), nb, s.d_colacc.p);
    IRTT_HIP(hipGetLastError());
    p.flag.release();
    p.nnz = nnz;
    p.valid = true;
}

void sp_apply(Scene &s, const double *w, double *y, bool transposed)
{
    const Sparse &p = s.sp;
    if (!p.valid) throw Error(IRTT_ERR_STATE, "no sparse operator");
    k_spmv<<<cdiv(s.nb, 4), 256, 0, s.stream>>>(transposed ? p.colptr.p : p.rowptr.p,
                                                 transposed ? p.row.p : p.col.p,
                                                 transposed ? p.cval.p : p.val.p, s.nb, w, y);
    IRTT_HIP(hipGetLastError());
}

// Y = T X (or T^T X) for R vectors of pitch ld, up to 64 per pass.
void sp_apply_batch(Scene &s, uint32_t R, const double *X, double *Y, bool transposed)
{
    Sparse &p = s.sp;
    if (!p.valid) throw Error(IRTT_ERR_STATE, "no sparse operator");
    const uint32_t nb = s.nb, ld = s.ld;
    hipStream_t st = s.stream;
    const uint32_t *ptr = transposed ? p.colptr.p : p.rowptr.p;
    const uint32_t *idx = transposed ? p.row.p : p.col.p;
    const float *val = transposed ? p.cval.p : p.val.p;
    p.d_xt.alloc((size_t)64 * nb);
    for (uint32_t r0 = 0; r0 < R; r0 += 64) {
        const uint32_t rc = std::min<uint32_t>(64, R - r0);
        uint32_t RP = 1;
        while (RP < rc) RP <<= 1;
        const double *Xc = X + (size_t)r0 * ld;
        double *Yc = Y + (size_t)r0 * ld;
        k_interleave<<<cdiv((uint64_t)nb * RP, 256), 256, 0, st>>>(Xc, ld, nb, rc, RP, p.d_xt.p);
        const uint32_t g = cdiv(nb, 4);
        switch (RP) {
        case 1: k_spmm<1><<<g, 256, 0, st>>>(ptr, idx, val, nb, p.d_xt.p, rc, ld, Yc); break;
        case 2: k_spmm<2><<<g, 256, 0, st>>>(ptr, idx, val, nb, p.d_xt.p, rc, ld, Yc); break;
        case 4: k_spmm<4><<<g, 256, 0, st>>>(ptr, idx, val, nb, p.d_xt.p, rc, ld, Yc); break;
        case 8: k_spmm<8><<<g, 256, 0, st>>>(ptr, idx, val, nb, p.d_xt.p, rc, ld, Yc); break;
        case 16: k_spmm<16><<<g, 256, 0, st>>>(ptr, idx, val, nb, p.d_xt.p, rc, ld, Yc); break;
        case 32: k_spmm<32><<<g, 256, 0, st>>>(ptr, idx, val, nb, p.d_xt.p, rc, ld, Yc); break;
        default: k_spmm<64><<<g, 256, 0, st>>>(ptr, idx, val, nb, p.d_xt.p, rc, ld, Yc); break;
        }
        IRTT_HIP(hipGetLastError());
    }
}

void sp_download(Scene &s, uint32_t *rowptr, uint32_t *col, float *val)
{
    const Sparse &p = s.sp;
    hipStream_t st = s.stream;
    IRTT_HIP(hipMemcpyAsync(rowptr, p.rowptr.p, ((size_t)s.nb + 1) * 4, hipMemcpyDeviceToHost, st));
    IRTT_HIP(hipMemcpyAsync(col, p.col.p, p.nnz * 4, hipMemcpyDeviceToHost, st));
    IRTT_HIP(hipMemcpyAsync(val, p.val.p, p.nnz * 4, hipMemcpyDeviceToHost, st));
    IRTT_HIP(hipStreamSynchronize(st));
}

}  // namespace irtt
