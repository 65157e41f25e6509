// path_records.hpp -- path records of photons, written by the adjoint trace's
// path replay (adjoint.hip) or by a transport build that keeps its paths
// (trace.hip, IRTT_TRANSPORT_KEEP_PATHS), read by the adjoint trace's terms.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace irtt {

// corner of the smallest barycentric weight (ties go to the earlier corner)
__device__ __forceinline__ int min_corner(float b0, float b1, float b2)
{
    int idx = 0;
    float v = b0;
    if (b1 < v) {
        idx = 1;
        v = b1;
    }
    if (b2 < v) idx = 2;
    return idx;
}

// Edge record: one hit near an edge, its occlusion ray still to trace.  The
// path writer stores the hit as it found it (raw form); k_edge_prep derives
// the occlusion ray and the edge's terms from it (ray form).
//   raw  r0: ray direction, distance from the segment's start to the hit
//        r1: barycentric weights of corners 1 and 2 at the hit, hit triangle,
//            source triangle
//        r2: start triangle, its corner nearest the start, source bin,
//            absorbing bin (~0: none, or not yet known)
//        r3: -, -, -, photon index
//   ray  r0: occlusion ray origin, distance from the segment's start to the hit
//        r1: ray direction, hit triangle
//        r2: edge-irradiance derivative dE, edge vertex
//        r3: origin vertex, source bin, absorbing bin, photon index
struct EdgeRec {
    float4 r0, r1, r2;
    uint4 r3;
};

// Segment record: one accepted path segment of a photon, its terms still to
// compute.
//   r0: start triangle (the emitter on the first segment), hit triangle,
//       source bin, absorbing bin (~0: none, or not yet known)
//   r1: barycentric weights of corners 1 and 2 at the start and at the hit
//   r2: photon index, source triangle, first segment (1) or not (0)
struct SegRec {
    uint4 r0;
    float4 r1;
    uint4 r2;
};

// The edge ratio of the edge records a transport build keeps (the adjoint
// trace's default): a hit whose smallest barycentric weight is at most this
// is near an edge.
constexpr float kKeptEdgeRatio = 0.05f;

// Records are written in chunks of kChunk consecutive slots, each chunk
// owned by one wave of the writing kernel (one atomic per chunk, not per
// write); fills[c] is the number of records chunk c holds.
constexpr uint32_t kChunk = 256;
constexpr uint32_t kNoChunk = 0xffffffffu;

struct RecBuf {
    void *p;                       // SegRec or EdgeRec slots
    uint32_t n_chunks;             // capacity in chunks
    uint32_t *fills;               // per chunk
    unsigned long long *next;      // chunks taken (may run past n_chunks: dropped records)
};

// The wave's next slots for the lanes with `want`, called by all 64 lanes
// together: chunk and fill are the wave's current chunk (the same in every
// lane).  The current chunk is filled to the end before a new one is taken.
// Returns the slot; ok is false when the buffer is full.
__device__ __forceinline__ uint64_t chunk_slot(bool want, uint32_t &chunk, uint32_t &fill,
                                               const RecBuf &b, bool &ok)
{
    const unsigned long long m = __ballot(want);
    const uint32_t k = (uint32_t)__popcll(m);
    ok = false;
    if (k == 0) return 0;
    const uint32_t prefix = (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
    const uint32_t rem = chunk == kNoChunk ? 0u : kChunk - fill;
    if (k <= rem) {
        ok = chunk < b.n_chunks;
        const uint64_t slot = (uint64_t)chunk * kChunk + fill + prefix;
        fill += k;
        return slot;
    }
    unsigned long long nc = 0;
    if (__lane_id() == 0) nc = atomicAdd(b.next, 1ull);
    nc = __shfl(nc, 0, 64);
    const uint32_t ncu = nc < (unsigned long long)(kNoChunk - 1) ? (uint32_t)nc : kNoChunk - 1;
    uint64_t slot;
    if (prefix < rem) {
        ok = chunk < b.n_chunks;
        slot = (uint64_t)chunk * kChunk + fill + prefix;
    } else {
        ok = ncu < b.n_chunks;
        slot = (uint64_t)ncu * kChunk + (prefix - rem);
    }
    if (chunk != kNoChunk && chunk < b.n_chunks && __lane_id() == 0) b.fills[chunk] = kChunk;
    chunk = ncu;
    fill = k - rem;
    return slot;
}

__device__ __forceinline__ void chunk_close(uint32_t chunk, uint32_t fill, const RecBuf &b)
{
    if (chunk != kNoChunk && chunk < b.n_chunks && __lane_id() == 0) b.fills[chunk] = fill;
}

}  // namespace irtt
