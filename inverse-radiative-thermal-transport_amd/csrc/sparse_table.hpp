// sparse_table.hpp -- device side of the hash table that combines the
// photon hits of the sparse operator layout (sparse.hip builds the rows and
// columns from it).
//
// Open addressing with linear probing over 2^lg slots of 16 bytes.  A key is
// claimed once (empty -> key by compare-and-swap) and never changes, so a
// stale read of a slot can only show it empty, and the compare-and-swap then
// returns the key that won.  Values are u64 fixed-point sums: integer adds
// commute, so a slot's final value does not depend on the order of the hits.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "irtt_internal.hpp"

namespace irtt {

constexpr unsigned long long kTabEmpty = ~0ull;

__device__ __forceinline__ unsigned long long tab_key(uint32_t row, uint32_t col, uint32_t bits)
{
    return ((unsigned long long)row << bits) | col;
}

// Fibonacci hashing: the top lg bits of key * 2^64 / golden ratio.
__device__ __forceinline__ unsigned long long tab_home(unsigned long long key, uint32_t lg)
{
    return (key * 0x9E3779B97F4A7C15ull) >> (64 - lg);
}

// Adds q to the slot of key.  A table with every slot taken raises *overflow
// (the host sizes the table above the number of distinct keys, so this does
// not happen; the probe loop is bounded regardless).
__device__ __forceinline__ void tab_add(TabSlot *tab, uint32_t lg, unsigned long long key,
                                        unsigned long long q, uint32_t *overflow)
{
    const unsigned long long mask = (1ull << lg) - 1ull;
    unsigned long long h = tab_home(key, lg);
    for (unsigned long long probe = 0; probe <= mask; probe++) {
        unsigned long long k =
            __hip_atomic_load(&tab[h].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == kTabEmpty) k = atomicCAS(&tab[h].key, kTabEmpty, key);
        if (k == kTabEmpty || k == key) {
            if (q) atomicAdd(&tab[h].val, q);
            return;
        }
        h = (h + 1) & mask;
    }
    atomicOr(overflow, 1u);
}

}  // namespace irtt
