// binning_impl.hpp -- definitions of the sort in binning.hpp (included once
// per translation unit that instantiates bin_visibilities).
#pragma once

#include "binning.hpp"

namespace rg {

template <typename Keyer, typename T>
int bin_run(const Keyer &kf, int64_t n, int64_t nkeys, uint32_t per_item, const BinPlan &pl,
            const BinBufs &b, rg_context *timer_ctx, hipStream_t s, BinResult &out)
{
    const Replicas rp = pl.rp;
    KernelTimer kt(timer_ctx, RG_TIMER_BINNING, s);
    RG_TRY(zero_fill(b.count, pl.bytes[BIN_COUNT], s));

    const int64_t nb = (n + BIN_CHUNK - 1) / BIN_CHUNK;
    if (nb > 0)
        hipLaunchKernelGGL((k_bin_count<Keyer>), dim3((unsigned)nb), dim3(BIN_TB), 0, s, kf, n,
                           b.count, rp, b.krp, b.stats);
    RG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)pl.nblocks_scan), dim3(SCAN_BLOCK), 0, s,
                       (const uint32_t *)b.count, rp, nkeys, per_item, b.bsums);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, b.bsums, (int)pl.nblocks_scan,
                       b.stats);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)pl.nblocks_scan), dim3(SCAN_BLOCK), 0, s,
                       (const uint32_t *)b.count, rp, nkeys, per_item, (const uint2 *)b.bsums,
                       b.vis_start, b.item_start, b.cursor, b.items);
    RG_HIP(hipGetLastError());
    if (nb > 0)
        hipLaunchKernelGGL((k_bin_scatter<Keyer, T>), dim3((unsigned)nb), dim3(BIN_TB), 0, s, kf, n,
                           (const uint32_t *)b.cursor, rp, (const uint4 *)b.krp,
                           (Rec<T> *)b.recs);
    RG_HIP(hipGetLastError());

    out.vis_start = b.vis_start;
    out.item_start = b.item_start;
    out.items = b.items;
    out.recs = b.recs;
    out.max_items = pl.max_items;
    return RG_OK;
}

template <typename Keyer, typename T>
int bin_visibilities(rg_context *ctx, const Keyer &kf, int64_t n, int64_t nkeys, uint32_t per_item,
                     hipStream_t s, BinResult &out, int64_t max_counters)
{
    BinPlan pl;
    RG_TRY(bin_plan(n, nkeys, per_item, max_counters, sizeof(Rec<T>), pl));
    RG_TRY(ensure_counters(ctx));
    // context scratch slots of the sort (rg_internal.hpp)
    static const int slot[BIN__N] = {0, 1, 2, 3, 4, 5, 6, 15};
    for (int k = 0; k < BIN__N; ++k)
        RG_TRY(ctx->scratch[slot[k]].ensure(pl.bytes[k]));
    BinBufs b;
    b.count = ctx->scratch[slot[BIN_COUNT]].as<uint32_t>();
    b.vis_start = ctx->scratch[slot[BIN_VIS_START]].as<uint32_t>();
    b.item_start = ctx->scratch[slot[BIN_ITEM_START]].as<uint32_t>();
    b.cursor = ctx->scratch[slot[BIN_CURSOR]].as<uint32_t>();
    b.items = ctx->scratch[slot[BIN_ITEMS]].as<Item>();
    b.recs = ctx->scratch[slot[BIN_RECS]].p;
    b.bsums = ctx->scratch[slot[BIN_BSUMS]].as<uint2>();
    b.krp = ctx->scratch[slot[BIN_KRP]].as<uint4>();
    b.stats = ctx->counters.as<unsigned long long>();
    return bin_run<Keyer, T>(kf, n, nkeys, per_item, pl, b, ctx, s, out);
}

} // namespace rg
