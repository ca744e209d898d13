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

// Kernel taps through a buffer resource: offsets beyond the table read zeros
// (TAP_NONE for a lane whose row the footprint misses, unpadded table).
__device__ inline float2 tap_load(__amdgpu_buffer_rsrc_t r, uint32_t off, float2 *)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ inline double2 tap_load(__amdgpu_buffer_rsrc_t r, uint32_t off, double2 *)
{
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
constexpr uint32_t TAP_NONE = 0x80000000u;
// word 3 of a raw (unswizzled, stride 0) buffer resource on gfx9
constexpr int BUF_RSRC_W3 = 0x00020000;

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
