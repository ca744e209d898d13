// romein.hpp -- software-pipelined record loop shared by the Romein kernels.
//
// Records are consumed in batches of U: `fetch(j, batch)` reads records
// [j, j + U) from LDS, computes each lane's footprint cell and issues the U
// kernel-tap loads; `update(batch)` applies them in stream order.  Two
// batches alternate so the loads of batch i + 1 are in flight while batch i
// is applied.
#pragma once

#include "binning.hpp"

namespace rg {

template <typename T, int U> struct RecBatch {
    int cell[U];
    typename Vec2<T>::type k[U];
    typename Vec2<T>::type v[U];
};

template <int U, typename B, typename Fetch, typename Update>
__device__ inline void pipelined_records(uint32_t n, Fetch &&fetch, Update &&update)
{
    B a, b;
    if (n == 0)
        return;
    fetch(0u, a);
    for (uint32_t j = 0; j < n; j += 2 * U) {
        if (j + U < n)
            fetch(j + U, b);
        update(a);
        if (j + U >= n)
            break;
        if (j + 2 * U < n)
            fetch(j + 2 * U, a);
        update(b);
    }
}

} // namespace rg
