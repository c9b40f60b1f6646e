// dev_count.h — per-lane work counters of the counting kernel variants
// (COUNT = true) and their flush into the context's counter slots.
#pragma once
#include "context.h"
#include "dev_math.h"

namespace ltd {

// Rays traced and BVH8 work of one lane: nodes / tris of closest-hit
// queries, snodes / stris of occlusion queries.
struct Ctr {
    uint32_t closest, shadow, nodes, tris, snodes, stris;
};

// Sums a wave's counters and adds them to counts (kCountSlots u64 slots:
// closest rays, shadow rays, nodes and tris of all queries, ..., nodes and
// tris of the occlusion queries at 7 and 8).  Every lane of the wave must
// call it.
LT_DEV void flush_counts(const Ctr& c, unsigned long long* counts) {
    unsigned long long v[6] = {c.closest, c.shadow, (unsigned long long)c.nodes + c.snodes,
                               (unsigned long long)c.tris + c.stris, c.snodes, c.stris};
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int k = 0; k < 6; k++) v[k] += __shfl_xor(v[k], o);
    if ((threadIdx.x & 63u) == 0u) {
        atomicAdd(&counts[0], v[0]);
        atomicAdd(&counts[1], v[1]);
        atomicAdd(&counts[2], v[2]);
        atomicAdd(&counts[3], v[3]);
        atomicAdd(&counts[7], v[4]);
        atomicAdd(&counts[8], v[5]);
    }
}

}  // namespace ltd
