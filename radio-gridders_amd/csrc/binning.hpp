// binning.hpp -- device-side counting sort of visibilities into work items.
//
// A "keyer" maps visibility i to (key, packed tile position, kernel-tap
// index).  Keys are dense in [0, nkeys).  The sort is:
//   count   : wave-aggregated atomics on count[rep][key]; each visibility
//             keeps the counter value it got (its rank inside the replica's
//             share of the key) with its key, position and tap index
//   scan    : three-phase exclusive scan of (count, ceil(count / K)), where
//             count[key] is the sum over the replicas, then the start of
//             every replica's share of every key
//   items   : one work item per K records of a key
//   scatter : record i written at start[rep][key] + rank (no atomics)
// Block b uses counter replica rep = b mod NREP.  Device-scope atomics on one
// address serialise, and a few keys (the short-baseline tiles at the uv
// centre) take thousands of wavefront updates, so the counters of a key are
// spread over NREP separate arrays; replica rep's records of a key follow
// those of replicas 0..rep-1.
// Consecutive visibilities of a wavefront that share a key keep their stream
// order, so baseline-track runs stay contiguous inside a key (the locality
// the W-projection Romein accumulation relies on).
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

// VIS_OVERFLOW: outside a capacity the call was sized for (counted as dropped
// and as an overrun of the call)
enum VisStatus { VIS_OK = 0, VIS_DROPPED = 1, VIS_BAD_PLANE = 2, VIS_OVERFLOW = 3 };

// Returns, for the calling lane, old counter value + rank inside its run (when
// `want` is set).  A run is a maximal stretch of consecutive active lanes with
// the same key; its first lane adds the run length to the key's counter with
// one atomic.  All runs of the wavefront issue their atomics in one
// instruction and the results are gathered once, so a wavefront costs one
// atomic round trip whatever its keys (visibility streams are baseline-track
// ordered, so neighbouring lanes mostly share keys).  Must be called by the
// whole wavefront in uniform control flow.
__device__ inline uint32_t wave_agg_add(uint32_t *counters, uint32_t key, bool active, bool want)
{
    const int lane = __lane_id();
    const unsigned long long act = __ballot(active);
    const uint32_t prev = (uint32_t)__shfl_up((int)key, 1);
    const bool head = active && (lane == 0 || !((act >> (lane - 1)) & 1ull) || prev != key);
    const unsigned long long heads = __ballot(head);
    const unsigned long long ends = heads | ~act; // a run stops at the next head or idle lane
    uint32_t base = 0;
    if (head) {
        const unsigned long long nxt = lane == 63 ? 0ull : (ends >> (lane + 1));
        const uint32_t len = nxt ? (uint32_t)__ffsll((long long)nxt) : (uint32_t)(64 - lane);
        base = atomicAdd(&counters[key], len);
    }
    if (!want)
        return 0;
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long hm = heads & upto;
    const int my_head = hm ? 63 - __clzll((long long)hm) : 0;
    return (uint32_t)__shfl((int)base, my_head) + (uint32_t)(lane - my_head);
}

// Counter replicas: NREP arrays of `stride` entries (stride >= nkeys + 1).
struct Replicas {
    uint32_t mask;   // NREP - 1, NREP a power of two
    int64_t stride;
};

// Per visibility after the count pass: key (KEY_NONE if not gridded), rank,
// packed position, tap index.
constexpr uint32_t KEY_NONE = 0xffffffffu;

template <typename Keyer>
__global__ __launch_bounds__(256) void k_bin_count(Keyer kf, int64_t n, uint32_t *count,
                                                   Replicas rp, uint4 *krp,
                                                   unsigned long long *stats)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t key = 0, pos = 0, kidx = 0;
    int st = VIS_DROPPED;
    if (i < n)
        st = kf.classify(i, key, pos, kidx);
    const bool ok = i < n && st == VIS_OK;
    const uint32_t rank =
        wave_agg_add(count + (blockIdx.x & rp.mask) * rp.stride, key, ok, true);
    if (i < n)
        krp[i] = make_uint4(ok ? key : KEY_NONE, rank, pos, kidx);
    // per-wave status counts
    const unsigned long long dropped =
        __ballot(i < n && (st == VIS_DROPPED || st == VIS_OVERFLOW));
    const unsigned long long badp = __ballot(i < n && st == VIS_BAD_PLANE);
    const unsigned long long over = __ballot(i < n && st == VIS_OVERFLOW);
    if (__lane_id() == 0) {
        if (dropped)
            atomicAdd(&stats[CNT_DROPPED], (unsigned long long)__popcll(dropped));
        if (badp)
            atomicAdd(&stats[CNT_BAD_PLANE], (unsigned long long)__popcll(badp));
        if (over)
            atomicAdd(&stats[CNT_OVERFLOW], (unsigned long long)__popcll(over));
    }
}

// Same grid as k_bin_count (block b, replica b mod NREP).
template <typename Keyer, typename T>
__global__ __launch_bounds__(256) void k_bin_scatter(Keyer kf, int64_t n, const uint32_t *start,
                                                     Replicas rp, const uint4 *__restrict__ krp,
                                                     Rec<T> *recs)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint4 e = krp[i];
    if (e.x == KEY_NONE)
        return;
    Rec<T> r;
    r.pos = e.z;
    r.kidx = e.w;
    r.vis = kf.template vis_value<T>(i);
    recs[start[(blockIdx.x & rp.mask) * rp.stride + e.x] + e.y] = r;
}

// Phase 1: per-block totals of (count, items).
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const uint32_t *count, Replicas rp,
                                                            int64_t nkeys, uint32_t per_item,
                                                            uint2 *block_sums);
// Phase 2: exclusive scan of block totals (single block), totals to stats.
__global__ __launch_bounds__(1024) void k_scan_blocks(uint2 *block_sums, int nblocks,
                                                      unsigned long long *stats);
// Phase 3: per-key starts, work items and the scatter cursors of every replica.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down(const uint32_t *count, Replicas rp,
                                                          int64_t nkeys, uint32_t per_item,
                                                          const uint2 *block_sums,
                                                          uint32_t *vis_start,
                                                          uint32_t *item_start, uint32_t *cursor,
                                                          Item *items);

// Host driver for count + scan + items + scatter.
struct BinResult {
    uint32_t *vis_start;  // [nkeys + 1] (last entry written by the scan)
    uint32_t *item_start; // [nkeys + 1]
    Item *items;          // [max_items]
    void *recs;           // Rec<T>[n]
    int64_t max_items;
};

// Device arrays of one sort and their sizes (bin_plan).
enum BinArray {
    BIN_COUNT = 0,      // per-replica key counters
    BIN_VIS_START = 1,  // first record of every key
    BIN_ITEM_START = 2, // first work item of every key
    BIN_CURSOR = 3,     // per-replica scatter starts
    BIN_ITEMS = 4,      // work items
    BIN_RECS = 5,       // sorted records
    BIN_BSUMS = 6,      // scan block totals
    BIN_KRP = 7,        // per visibility: key, rank, position, tap index
    BIN__N = 8
};

struct BinPlan {
    Replicas rp;
    int64_t nblocks_scan, max_items;
    size_t bytes[BIN__N];
};

struct BinBufs {
    uint32_t *count, *vis_start, *item_start, *cursor;
    Item *items;
    void *recs;
    uint2 *bsums;
    uint4 *krp;
    unsigned long long *stats; // CNT__N call counters
};

// Validates the sort's sizes and fills the plan (no device work).
int bin_plan(int64_t n, int64_t nkeys, uint32_t per_item, int64_t max_counters, size_t rec_bytes,
             BinPlan &pl);
// The sort into caller-provided arrays of the plan's sizes; timer_ctx (may be
// NULL) receives the RG_TIMER_BINNING events.
template <typename Keyer, typename T>
int bin_run(const Keyer &kf, int64_t n, int64_t nkeys, uint32_t per_item, const BinPlan &pl,
            const BinBufs &b, rg_context *timer_ctx, hipStream_t s, BinResult &out);
// bin_run into the context's scratch slots 0-6 and 15.
template <typename Keyer, typename T>
int bin_visibilities(rg_context *ctx, const Keyer &kf, int64_t n, int64_t nkeys, uint32_t per_item,
                     hipStream_t s, BinResult &out, int64_t max_counters = (int64_t)1 << 22);

// Work items over per-key lengths len[0..nkeys): prefix[k] = exclusive sum of
// len, item_start[k] = first item of key k (both nkeys + 1 entries), items of
// at most per_item entries with start = prefix[key] + q * per_item (callers
// shift them to their own offsets).  Writes the item count to the context's
// counters; uses scratch slot 6.
int items_from_lengths(rg_context *ctx, const uint32_t *len, int64_t nkeys, uint32_t per_item,
                       uint32_t *prefix, uint32_t *item_start, Item *items, hipStream_t s);

} // namespace rg
