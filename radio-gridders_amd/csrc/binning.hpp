// binning.hpp -- device-side counting sort of visibilities into work items.
//
// A "keyer" maps visibility i to (key, packed tile position, kernel-tap index).
// Keys are dense in [0, nkeys).  The sort is:
//   count   : wave-aggregated atomics on count[key]
//   scan    : three-phase exclusive scan of (count, ceil(count / K))
//   items   : one work item per K visibilities of a key
//   scatter : wave-aggregated atomics on cursor[key], record written in place
// Consecutive visibilities of a wavefront that share a key keep their stream
// order, so baseline-track runs stay contiguous inside a key (the locality
// the Romein accumulation relies on).
#pragma once

#include "rg_internal.hpp"

namespace rg {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_PER_THREAD = 16;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_PER_THREAD; // keys per scan block

struct Item {
    uint32_t key;
    uint32_t start; // first record
    uint32_t count; // records
};

template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

template <typename T> struct Rec {
    uint32_t pos;  // tile-local footprint top-left: lx | ly << 16
    uint32_t kidx; // (plane * R + yf) * R + xf
    typename Vec2<T>::type vis;
};

enum VisStatus { VIS_OK = 0, VIS_DROPPED = 1, VIS_BAD_PLANE = 2 };

// Returns, for the calling lane, old counter value + rank among the lanes
// of the wavefront that have the same key.  Must be called by the whole
// wavefront in uniform control flow.
__device__ inline uint32_t wave_agg_add(uint32_t *counters, uint32_t key, bool active)
{
    const int lane = __lane_id();
    unsigned long long todo = __ballot(active);
    uint32_t result = 0;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const uint32_t lkey = __shfl(key, leader);
        const unsigned long long same = __ballot(active && key == lkey) & todo;
        uint32_t base = 0;
        if (lane == leader)
            base = atomicAdd(&counters[lkey], (uint32_t)__popcll(same));
        base = __shfl(base, leader);
        if ((same >> lane) & 1ull)
            result = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        todo &= ~same;
    }
    return result;
}

template <typename Keyer>
__global__ __launch_bounds__(256) void k_bin_count(Keyer kf, int64_t n, uint32_t *count,
                                                   unsigned long long *stats)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t key = 0, pos = 0, kidx = 0;
    int st = VIS_DROPPED;
    if (i < n)
        st = kf.classify(i, key, pos, kidx);
    wave_agg_add(count, key, i < n && st == VIS_OK);
    // per-wave status counts
    const unsigned long long dropped = __ballot(i < n && st == VIS_DROPPED);
    const unsigned long long badp = __ballot(i < n && st == VIS_BAD_PLANE);
    if (__lane_id() == 0) {
        if (dropped)
            atomicAdd(&stats[CNT_DROPPED], (unsigned long long)__popcll(dropped));
        if (badp)
            atomicAdd(&stats[CNT_BAD_PLANE], (unsigned long long)__popcll(badp));
    }
}

template <typename Keyer, typename T>
__global__ __launch_bounds__(256) void k_bin_scatter(Keyer kf, int64_t n, uint32_t *cursor,
                                                     Rec<T> *recs)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t key = 0, pos = 0, kidx = 0;
    int st = VIS_DROPPED;
    if (i < n)
        st = kf.classify(i, key, pos, kidx);
    const bool ok = i < n && st == VIS_OK;
    const uint32_t dst = wave_agg_add(cursor, key, ok);
    if (ok) {
        Rec<T> r;
        r.pos = pos;
        r.kidx = kidx;
        r.vis = kf.template vis_value<T>(i);
        recs[dst] = r;
    }
}

// Phase 1: per-block totals of (count, items).
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const uint32_t *count, int64_t nkeys,
                                                            uint32_t per_item, uint2 *block_sums);
// Phase 2: exclusive scan of block totals (single block), totals to stats.
__global__ __launch_bounds__(1024) void k_scan_blocks(uint2 *block_sums, int nblocks,
                                                      unsigned long long *stats);
// Phase 3: per-key starts, work items and scatter cursors.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down(const uint32_t *count, int64_t nkeys,
                                                          uint32_t per_item,
                                                          const uint2 *block_sums,
                                                          uint32_t *vis_start,
                                                          uint32_t *item_start, uint32_t *cursor,
                                                          Item *items);

// Host driver for count + scan + items + scatter.  Scratch layout is
// managed by the caller through the context's scratch slots 0..5.
struct BinResult {
    uint32_t *vis_start;  // [nkeys + 1] (last entry written by the scan)
    uint32_t *item_start; // [nkeys + 1]
    Item *items;          // [max_items]
    void *recs;           // Rec<T>[n]
    int64_t max_items;
};

template <typename Keyer, typename T>
int bin_visibilities(rg_context *ctx, const Keyer &kf, int64_t n, int64_t nkeys, uint32_t per_item,
                     hipStream_t s, BinResult &out);

} // namespace rg
