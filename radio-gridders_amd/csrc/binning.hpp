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
// Chunk c of BIN_CHUNK visibilities uses counter replica rep = c mod NREP.
// Device-scope atomics on one address serialise, and a few keys (the
// short-baseline tiles at the uv centre) take thousands of wavefront updates,
// so the counters of a key are spread over NREP separate arrays; replica
// rep's records of a key follow those of replicas 0..rep-1.
// Consecutive visibilities of a wavefront that share a key keep their stream
// order, so baseline-track runs stay contiguous inside a key (the locality
// the W-projection Romein accumulation relies on).
#pragma once

#include "rg_internal.hpp"

namespace rg {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_PER_THREAD = 4; // 1024 keys per block: W-towers key spaces of ~1 M keys
                                   // give ~1000 blocks, enough to fill the CUs
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

// A run is a maximal stretch of consecutive active lanes with the same key;
// its first lane (the head) adds the run length to the key's counter with one
// atomic.  All runs of the wavefront issue their atomics in one instruction
// and the results are gathered once, so a wavefront costs one atomic round
// trip whatever its keys (visibility streams are baseline-track ordered, so
// neighbouring lanes often share keys).  wave_agg_issue returns the atomic's
// old value (meaningful on head lanes) and the lane's head; wave_agg_rank
// turns them into the lane's rank.  Splitting the two lets a thread keep
// several atomics in flight.  Must be called by the whole wavefront in
// uniform control flow.
__device__ inline uint32_t wave_agg_issue(uint32_t *counters, uint32_t key, bool active,
                                          int &my_head)
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
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long hm = heads & upto;
    my_head = hm ? 63 - __clzll((long long)hm) : 0;
    return base;
}

__device__ inline uint32_t wave_agg_rank(uint32_t base, int my_head)
{
    return (uint32_t)__shfl((int)base, my_head) + (uint32_t)(__lane_id() - my_head);
}

// Both steps; the rank only when `want` (uniform) is set.
__device__ inline uint32_t wave_agg_add(uint32_t *counters, uint32_t key, bool active, bool want)
{
    int my_head;
    const uint32_t base = wave_agg_issue(counters, key, active, my_head);
    return want ? wave_agg_rank(base, my_head) : 0u;
}

// Counter replicas: NREP arrays of `stride` entries (stride >= nkeys + 1).
struct Replicas {
    uint32_t mask;   // NREP - 1, NREP a power of two
    int64_t stride;
};

// Per visibility after the count pass: key (KEY_NONE if not gridded), rank,
// packed position, tap index.
constexpr uint32_t KEY_NONE = 0xffffffffu;

// Count and scatter workgroups take BIN_CHUNK consecutive visibilities,
// BIN_VPT per thread (visibility chunk + q BIN_TB + thread, so the lanes of a
// wavefront hold consecutive visibilities); chunk c uses counter replica
// c mod NREP.
constexpr int BIN_TB = 256;
constexpr int BIN_VPT = 4;
constexpr int BIN_CHUNK = BIN_TB * BIN_VPT;

template <typename Keyer>
__global__ __launch_bounds__(BIN_TB) void k_bin_count(Keyer kf, int64_t n, uint32_t *count,
                                                      Replicas rp, uint4 *krp,
                                                      unsigned long long *stats)
{
    const int64_t i0 = (int64_t)blockIdx.x * BIN_CHUNK + threadIdx.x;
    uint32_t *cnt = count + (blockIdx.x & rp.mask) * rp.stride;
    uint32_t key[BIN_VPT], pos[BIN_VPT], kidx[BIN_VPT], base[BIN_VPT];
    int st[BIN_VPT], head[BIN_VPT];
#pragma unroll
    for (int q = 0; q < BIN_VPT; ++q) {
        const int64_t i = i0 + q * BIN_TB;
        key[q] = pos[q] = kidx[q] = 0;
        st[q] = i < n ? kf.classify(i, key[q], pos[q], kidx[q]) : VIS_DROPPED;
    }
    // every atomic in flight before the first result is used
#pragma unroll
    for (int q = 0; q < BIN_VPT; ++q)
        base[q] = wave_agg_issue(cnt, key[q], i0 + q * BIN_TB < n && st[q] == VIS_OK, head[q]);
    uint32_t dropped = 0, badp = 0, over = 0;
#pragma unroll
    for (int q = 0; q < BIN_VPT; ++q) {
        const int64_t i = i0 + q * BIN_TB;
        const bool in = i < n, ok = in && st[q] == VIS_OK;
        const uint32_t rank = wave_agg_rank(base[q], head[q]);
        if (in)
            krp[i] = make_uint4(ok ? key[q] : KEY_NONE, rank, pos[q], kidx[q]);
        // per-wave status counts
        dropped += (uint32_t)__popcll(__ballot(in && (st[q] == VIS_DROPPED || st[q] == VIS_OVERFLOW)));
        badp += (uint32_t)__popcll(__ballot(in && st[q] == VIS_BAD_PLANE));
        over += (uint32_t)__popcll(__ballot(in && st[q] == VIS_OVERFLOW));
    }
    if (__lane_id() == 0) {
        if (dropped)
            atomicAdd(&stats[CNT_DROPPED], (unsigned long long)dropped);
        if (badp)
            atomicAdd(&stats[CNT_BAD_PLANE], (unsigned long long)badp);
        if (over)
            atomicAdd(&stats[CNT_OVERFLOW], (unsigned long long)over);
    }
}

// Same grid and chunks as k_bin_count.
template <typename Keyer, typename T>
__global__ __launch_bounds__(BIN_TB) void k_bin_scatter(Keyer kf, int64_t n,
                                                        const uint32_t *start, Replicas rp,
                                                        const uint4 *__restrict__ krp,
                                                        Rec<T> *recs)
{
    const int64_t i0 = (int64_t)blockIdx.x * BIN_CHUNK + threadIdx.x;
    const uint32_t *st = start + (blockIdx.x & rp.mask) * rp.stride;
    uint4 e[BIN_VPT];
    typename Vec2<T>::type v[BIN_VPT];
#pragma unroll
    for (int q = 0; q < BIN_VPT; ++q) {
        const int64_t i = i0 + q * BIN_TB;
        e[q] = make_uint4(KEY_NONE, 0, 0, 0);
        if (i < n) {
            e[q] = krp[i];
            v[q] = kf.template vis_value<T>(i);
        }
    }
#pragma unroll
    for (int q = 0; q < BIN_VPT; ++q) {
        if (e[q].x == KEY_NONE)
            continue;
        Rec<T> r;
        r.pos = e[q].z;
        r.kidx = e[q].w;
        r.vis = v[q];
        recs[st[e[q].x] + e[q].y] = r;
    }
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
