// binning.hip -- scan kernels of the device-side counting sort.
#include "binning.hpp"

namespace rg {

// Exclusive block scan of uint2 values (x and y scanned independently).
// Returns the exclusive prefix of the calling thread and the block total.
template <int NT>
__device__ inline uint2 block_exclusive_scan(uint2 v, uint2 &total, uint2 *lds /* NT */)
{
    lds[threadIdx.x] = v;
    __syncthreads();
    // Hillis-Steele inclusive scan in LDS (NT <= 1024)
    for (int off = 1; off < NT; off <<= 1) {
        uint2 add = make_uint2(0, 0);
        if ((int)threadIdx.x >= off)
            add = lds[threadIdx.x - off];
        __syncthreads();
        uint2 cur = lds[threadIdx.x];
        cur.x += add.x;
        cur.y += add.y;
        lds[threadIdx.x] = cur;
        __syncthreads();
    }
    uint2 inc = lds[threadIdx.x];
    total = lds[NT - 1];
    __syncthreads();
    return make_uint2(inc.x - v.x, inc.y - v.y);
}

__device__ inline uint32_t items_of(uint32_t c, uint32_t per_item)
{
    return (c + per_item - 1) / per_item;
}

// total count of key k over the replicas
__device__ inline uint32_t key_count(const uint32_t *count, Replicas rp, int64_t k)
{
    uint32_t c = 0;
    for (uint32_t r = 0; r <= rp.mask; ++r)
        c += count[r * rp.stride + k];
    return c;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const uint32_t *count, Replicas rp,
                                                            int64_t nkeys, uint32_t per_item,
                                                            uint2 *block_sums)
{
    __shared__ uint2 lds[SCAN_BLOCK];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    uint2 acc = make_uint2(0, 0);
    for (int j = 0; j < SCAN_PER_THREAD; ++j) {
        const int64_t k = base + (int64_t)j * SCAN_BLOCK + threadIdx.x;
        if (k < nkeys) {
            const uint32_t c = key_count(count, rp, k);
            acc.x += c;
            acc.y += items_of(c, per_item);
        }
    }
    uint2 total;
    block_exclusive_scan<SCAN_BLOCK>(acc, total, lds);
    if (threadIdx.x == 0)
        block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_blocks(uint2 *block_sums, int nblocks,
                                                      unsigned long long *stats)
{
    __shared__ uint2 lds[1024];
    const int per = (nblocks + 1023) / 1024;
    const int b0 = threadIdx.x * per;
    uint2 acc = make_uint2(0, 0);
    for (int j = 0; j < per; ++j) {
        if (b0 + j < nblocks) {
            const uint2 v = block_sums[b0 + j];
            acc.x += v.x;
            acc.y += v.y;
        }
    }
    uint2 total;
    uint2 run = block_exclusive_scan<1024>(acc, total, lds);
    for (int j = 0; j < per; ++j) {
        if (b0 + j < nblocks) {
            const uint2 v = block_sums[b0 + j];
            block_sums[b0 + j] = run;
            run.x += v.x;
            run.y += v.y;
        }
    }
    if (threadIdx.x == 0) {
        block_sums[nblocks] = total;
        stats[CNT_RECORDS] = total.x;
        stats[CNT_ITEMS] = total.y;
    }
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down(const uint32_t *count, Replicas rp,
                                                          int64_t nkeys, uint32_t per_item,
                                                          const uint2 *block_sums,
                                                          uint32_t *vis_start,
                                                          uint32_t *item_start, uint32_t *cursor,
                                                          Item *items)
{
    __shared__ uint32_t cnt[SCAN_TILE];
    __shared__ uint2 lds[SCAN_BLOCK];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    for (int j = 0; j < SCAN_PER_THREAD; ++j) {
        const int e = j * SCAN_BLOCK + threadIdx.x;
        cnt[e] = (base + e < nkeys) ? key_count(count, rp, base + e) : 0u;
    }
    __syncthreads();
    const int e0 = threadIdx.x * SCAN_PER_THREAD;
    uint2 acc = make_uint2(0, 0);
    for (int j = 0; j < SCAN_PER_THREAD; ++j) {
        const uint32_t c = cnt[e0 + j];
        acc.x += c;
        acc.y += items_of(c, per_item);
    }
    uint2 total;
    uint2 run = block_exclusive_scan<SCAN_BLOCK>(acc, total, lds);
    const uint2 boff = block_sums[blockIdx.x];
    run.x += boff.x;
    run.y += boff.y;
    for (int j = 0; j < SCAN_PER_THREAD; ++j) {
        const int64_t k = base + e0 + j;
        if (k >= nkeys)
            break;
        const uint32_t c = cnt[e0 + j];
        vis_start[k] = run.x;
        item_start[k] = run.y;
        uint32_t cur = run.x;
        for (uint32_t r = 0; r <= rp.mask; ++r) {
            cursor[r * rp.stride + k] = cur;
            cur += count[r * rp.stride + k];
        }
        const uint32_t ni = items_of(c, per_item);
        for (uint32_t q = 0; q < ni; ++q) {
            Item it;
            it.key = (uint32_t)k;
            it.start = run.x + q * per_item;
            it.count = min(per_item, c - q * per_item);
            items[run.y + q] = it;
        }
        run.x += c;
        run.y += ni;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        // totals live after the last block's entry of block_sums
        const uint2 tot = block_sums[gridDim.x];
        vis_start[nkeys] = tot.x;
        item_start[nkeys] = tot.y;
    }
}

int bin_plan(int64_t n, int64_t nkeys, uint32_t per_item, int64_t max_counters, size_t rec_bytes,
             BinPlan &pl)
{
    // record indices are 32-bit
    if (n >= (int64_t)0x7fffffff)
        return fail(RG_ERR_PARAM, "visibility stream too long for one call (%lld)", (long long)n);
    if (nkeys <= 0 || nkeys >= (int64_t)0x7fffffff)
        return fail(RG_ERR_PARAM, "invalid key space %lld", (long long)nkeys);
    if (per_item == 0)
        return fail(RG_ERR_PARAM, "work items need at least one record");
    pl.nblocks_scan = (nkeys + SCAN_TILE - 1) / SCAN_TILE;
    if (pl.nblocks_scan > 1024 * 256)
        return fail(RG_ERR_PARAM, "key space too large (%lld)", (long long)nkeys);
    pl.max_items = nkeys + n / per_item + 1;
    // counter replicas: up to 16, at most max_counters counters in all
    // (replicas spread the atomics of hot keys; a stream without hot keys
    // is better served by fewer counters that stay in cache)
    pl.rp.stride = (nkeys + 1 + 63) / 64 * 64;
    uint32_t nrep = 16;
    while (nrep > 1 && (int64_t)nrep * pl.rp.stride > max_counters)
        nrep >>= 1;
    pl.rp.mask = nrep - 1;
    const int64_t n1 = n > 0 ? n : 1;
    pl.bytes[BIN_COUNT] = sizeof(uint32_t) * nrep * pl.rp.stride;
    pl.bytes[BIN_VIS_START] = sizeof(uint32_t) * (nkeys + 1);
    pl.bytes[BIN_ITEM_START] = sizeof(uint32_t) * (nkeys + 1);
    pl.bytes[BIN_CURSOR] = sizeof(uint32_t) * nrep * pl.rp.stride;
    pl.bytes[BIN_ITEMS] = sizeof(Item) * pl.max_items;
    pl.bytes[BIN_RECS] = rec_bytes * n1;
    pl.bytes[BIN_BSUMS] = sizeof(uint2) * (pl.nblocks_scan + 1);
    pl.bytes[BIN_KRP] = sizeof(uint4) * n1;
    return RG_OK;
}

int items_from_lengths(rg_context *ctx, const uint32_t *len, int64_t nkeys, uint32_t per_item,
                       uint32_t *prefix, uint32_t *item_start, Item *items, hipStream_t s)
{
    const int64_t nblocks_scan = (nkeys + SCAN_TILE - 1) / SCAN_TILE;
    if (nkeys <= 0 || nblocks_scan > 1024 * 256)
        return fail(RG_ERR_PARAM, "invalid key space %lld", (long long)nkeys);
    RG_TRY(ensure_counters(ctx));
    RG_TRY(ctx->scratch[6].ensure(sizeof(uint2) * (nblocks_scan + 1)));
    Replicas one;
    one.mask = 0;
    one.stride = nkeys + 1;
    uint2 *bsums = ctx->scratch[6].as<uint2>();
    auto *stats = ctx->counters.as<unsigned long long>();
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nblocks_scan), dim3(SCAN_BLOCK), 0, s, len, one,
                       nkeys, per_item, bsums);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, bsums, (int)nblocks_scan, stats);
    // the single cursor array coincides with the prefix (same values)
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nblocks_scan), dim3(SCAN_BLOCK), 0, s, len, one,
                       nkeys, per_item, (const uint2 *)bsums, prefix, item_start, prefix, items);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

} // namespace rg
