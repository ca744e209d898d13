// binning_impl.hpp -- definitions of the sort in binning.hpp (included once
// per translation unit that instantiates bin_visibilities).
#pragma once

#include "binning.hpp"

namespace rg {

template <typename Keyer, typename T>
int bin_visibilities(rg_context *ctx, const Keyer &kf, int64_t n, int64_t nkeys, uint32_t per_item,
                     hipStream_t s, BinResult &out, int64_t max_counters)
{
    // record indices are 32-bit
    const int64_t nrec = n;
    if (nrec >= (int64_t)0x7fffffff)
        return fail(RG_ERR_PARAM, "visibility stream too long for one call (%lld)", (long long)n);
    if (nkeys <= 0 || nkeys >= (int64_t)0x7fffffff)
        return fail(RG_ERR_PARAM, "invalid key space %lld", (long long)nkeys);
    const int64_t nblocks_scan = (nkeys + SCAN_TILE - 1) / SCAN_TILE;
    if (nblocks_scan > 1024 * 64)
        return fail(RG_ERR_PARAM, "key space too large (%lld)", (long long)nkeys);
    const int64_t max_items = nkeys + nrec / per_item + 1;
    // counter replicas: up to 16, at most max_counters counters in all
    // (replicas spread the atomics of hot keys; a stream without hot keys
    // is better served by fewer counters that stay in cache)
    Replicas rp;
    rp.stride = (nkeys + 1 + 63) / 64 * 64;
    uint32_t nrep = 16;
    while (nrep > 1 && (int64_t)nrep * rp.stride > max_counters)
        nrep >>= 1;
    rp.mask = nrep - 1;

    RG_TRY(ensure_counters(ctx));
    RG_TRY(ctx->scratch[0].ensure(sizeof(uint32_t) * nrep * rp.stride)); // count
    RG_TRY(ctx->scratch[1].ensure(sizeof(uint32_t) * (nkeys + 1)));       // vis_start
    RG_TRY(ctx->scratch[2].ensure(sizeof(uint32_t) * (nkeys + 1)));       // item_start
    RG_TRY(ctx->scratch[3].ensure(sizeof(uint32_t) * nrep * rp.stride)); // cursor
    RG_TRY(ctx->scratch[4].ensure(sizeof(Item) * max_items));              // items
    RG_TRY(ctx->scratch[5].ensure(sizeof(Rec<T>) * (nrec > 0 ? nrec : 1))); // records
    RG_TRY(ctx->scratch[6].ensure(sizeof(uint2) * (nblocks_scan + 1)));   // block sums
    RG_TRY(ctx->scratch[15].ensure(sizeof(uint4) * (n > 0 ? n : 1)));     // key, rank, pos, kidx

    uint32_t *count = ctx->scratch[0].as<uint32_t>();
    KernelTimer kt(ctx, RG_TIMER_BINNING, s);
    RG_TRY(zero_fill(count, sizeof(uint32_t) * nrep * rp.stride, s));
    auto *stats = ctx->counters.as<unsigned long long>();

    const int tb = 256;
    const int64_t nb = (n + tb - 1) / tb;
    if (nb > 0)
        hipLaunchKernelGGL((k_bin_count<Keyer>), dim3((unsigned)nb), dim3(tb), 0, s, kf, n, count,
                           rp, ctx->scratch[15].as<uint4>(), stats);
    RG_HIP(hipGetLastError());
    uint2 *bsums = ctx->scratch[6].as<uint2>();
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nblocks_scan), dim3(SCAN_BLOCK), 0, s,
                       (const uint32_t *)count, rp, nkeys, per_item, bsums);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, bsums, (int)nblocks_scan, stats);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nblocks_scan), dim3(SCAN_BLOCK), 0, s,
                       (const uint32_t *)count, rp, nkeys, per_item, (const uint2 *)bsums,
                       ctx->scratch[1].as<uint32_t>(), ctx->scratch[2].as<uint32_t>(),
                       ctx->scratch[3].as<uint32_t>(), ctx->scratch[4].as<Item>());
    RG_HIP(hipGetLastError());
    if (nb > 0)
        hipLaunchKernelGGL((k_bin_scatter<Keyer, T>), dim3((unsigned)nb), dim3(tb), 0, s, kf, n,
                           (const uint32_t *)ctx->scratch[3].as<uint32_t>(), rp,
                           (const uint4 *)ctx->scratch[15].as<uint4>(),
                           ctx->scratch[5].as<Rec<T>>());
    RG_HIP(hipGetLastError());

    out.vis_start = ctx->scratch[1].as<uint32_t>();
    out.item_start = ctx->scratch[2].as<uint32_t>();
    out.items = ctx->scratch[4].as<Item>();
    out.recs = ctx->scratch[5].p;
    out.max_items = max_items;
    return RG_OK;
}

} // namespace rg
