// wproj.hip -- Romein W-projection gridders.
//
// Plain path: the reference's launch structure (scatter_grid_kernel_flat,
// grid_cuda.cu:253-280, launched <<<512, 256>>> by wproj_CUDA_flat,
// grid_cuda.cu:893): contiguous bins of the flat stream, each thread owns one
// footprint position, FP64 coordinates per thread, FP64 atomics when the
// thread's grid point changes.
//
// Batched path: visibilities are counting-sorted by the grid tile of their
// footprint (binning.hpp), coordinates computed once per visibility during
// the sort (FP64, same formulas), then accumulated per tile in LDS by the
// Romein kernel of tiles.hpp.
#include <algorithm>
#include <mutex>

#include "binning_impl.hpp"
#include "tiles.hpp"

namespace rg {

// ---- zero fill (graph-capturable, no memset nodes) ---------------------------

__global__ void k_zero(uint4 *p, size_t n16, unsigned char *tail, size_t ntail)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        p[i] = make_uint4(0, 0, 0, 0);
    if (blockIdx.x == 0 && threadIdx.x < ntail)
        tail[threadIdx.x] = 0;
}

int zero_fill(void *p, size_t bytes, hipStream_t s)
{
    if (bytes == 0)
        return RG_OK;
    if (((uintptr_t)p & 15) != 0)
        return fail(RG_ERR_PARAM, "zero_fill: pointer not 16-byte aligned");
    const size_t n16 = bytes / 16, ntail = bytes % 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    if (blocks == 0)
        blocks = 1;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, (uint4 *)p, n16,
                       (unsigned char *)p + n16 * 16, ntail);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

// ---- plain FP64 Romein (reference launch structure) ---------------------------

__device__ inline void grid_add_f64(double2 *grid, int N, int gu, int gv, double re, double im)
{
    // scatter_grid_add bounds check (grid_cuda.cu:29-31)
    if (gu < 0 || gu >= N || gv < 0 || gv >= N)
        return;
    double *g = reinterpret_cast<double *>(grid + (size_t)gv * N + gu);
    atomicAdd(g, re);
    atomicAdd(g + 1, im);
}

__global__ __launch_bounds__(256) void k_wproj_plain(const double *__restrict__ u,
                                                     const double *__restrict__ v,
                                                     const double *__restrict__ w,
                                                     const double2 *__restrict__ vis, int64_t nvis,
                                                     int nbins, const double2 *__restrict__ kern,
                                                     int P, int R, int S, double w_min,
                                                     double w_step, double2 *__restrict__ grid,
                                                     int N, double theta,
                                                     unsigned long long *stats)
{
    // bin_flat_visibilities (radio.cuh:504-590): equal bins, remainder to the last.
    const int64_t per = nvis / nbins;
    const int64_t i0 = per * blockIdx.x;
    const int64_t i1 = (blockIdx.x == (unsigned)nbins - 1) ? nvis : i0 + per;
    const int SS = S * S;
    for (int pos = threadIdx.x; pos < SS; pos += blockDim.x) {
        const int myU = pos % S, myV = pos / S;
        int cu = INT_MIN, cv = INT_MIN;
        double sr = 0.0, si = 0.0;
        for (int64_t i = i0; i < i1; ++i) {
            int gx, gy, xf, yf;
            frac_coord(N, S, R, theta, u[i], v[i], 0.0, 0.0, gx, gy, xf, yf);
            const int p = w_plane(w[i], w_min, w_step);
            if (p < 0 || p >= P) {
                if (pos == 0)
                    atomicAdd(&stats[CNT_BAD_PLANE], 1ull);
                continue;
            }
            int conv_u = (myU - gx) % S;
            int conv_v = (myV - gy) % S;
            if (conv_u < 0)
                conv_u += S;
            if (conv_v < 0)
                conv_v += S;
            const int gu = gx + conv_u, gv = gy + conv_v;
            if (gu != cu || gv != cv) {
                if (cu != INT_MIN)
                    grid_add_f64(grid, N, cu, cv, sr, si);
                sr = 0.0;
                si = 0.0;
                cu = gu;
                cv = gv;
            }
            const double2 k =
                kern[((size_t)(p * R + yf) * R + xf) * SS + (size_t)conv_v * S + conv_u];
            const double2 vv = vis[i];
            // sum += conj(k) * vis
            sr += k.x * vv.x + k.y * vv.y;
            si += k.x * vv.y - k.y * vv.x;
        }
        if (cu != INT_MIN)
            grid_add_f64(grid, N, cu, cv, sr, si);
    }
}

int wproj_plain(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
                const rg_wkern *wk, void *grid, hipStream_t s)
{
    if (vis->vis_type != RG_C128 || wk->type != RG_C128)
        return fail(RG_ERR_PARAM, "plain mode is FP64: visibilities and kernels must be RG_C128");
    const int N = p->grid_size;
    RG_TRY(zero_fill(grid, sizeof(double2) * (size_t)N * N, s));
    const int nbins = 512; // grid_cuda.cu:881-893
    KernelTimer kt(ctx, RG_TIMER_WPROJ_PLAIN, s);
    if (vis->n > 0)
        hipLaunchKernelGGL(k_wproj_plain, dim3(nbins), dim3(256), 0, s, vis->u, vis->v, vis->w,
                           (const double2 *)vis->vis, vis->n, nbins, (const double2 *)wk->data,
                           wk->plane_count, wk->oversampling, wk->size, wk->w_min, wk->w_step,
                           (double2 *)grid, N, p->theta,
                           ctx->counters.as<unsigned long long>());
    RG_HIP(hipGetLastError());
    return RG_OK;
}

// ---- batched tiled Romein -------------------------------------------------------

struct WprojKeyer {
    const double *u, *v, *w;
    const void *vis;
    int vis_type;
    int N, S, R, P;
    double theta, w_min, w_step;
    int tile, ntx, ntiles;
    int zb, rtp; // bytes per kernel tap; cell row stride of a tile (tile + S - 1 rounded up to S)
    const int64_t *chan_off;
    int nchan;

    __device__ int chan_of(int64_t i) const
    {
        int lo = 0, hi = nchan - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (chan_off[mid] <= i)
                lo = mid;
            else
                hi = mid - 1;
        }
        return lo;
    }

    __device__ int classify(int64_t i, uint32_t &key, uint32_t &pos, uint32_t &kidx) const
    {
        int gx, gy, xf, yf;
        frac_coord(N, S, R, theta, u[i], v[i], 0.0, 0.0, gx, gy, xf, yf);
        if (gx <= -S || gx >= N || gy <= -S || gy >= N)
            return VIS_DROPPED;
        const int p = w_plane(w[i], w_min, w_step);
        if (p < 0 || p >= P)
            return VIS_BAD_PLANE;
        const int tx = (gx + S) / tile, ty = (gy + S) / tile;
        const int lx = gx + S - tx * tile, ly = gy + S - ty * tile;
        const int c = nchan > 1 ? chan_of(i) : 0;
        key = (uint32_t)(c * ntiles + ty * ntx + tx);
        // byte offsets as the tile kernel uses them: cell (lx, ly) of the tile's
        // LDS cells and row ly of a sub-kernel; first tap of the sub-kernel
        pos = (uint32_t)((ly * rtp + lx) * zb) | ((uint32_t)(ly * S * zb) << 16);
        kidx = (uint32_t)((p * R + yf) * R + xf) * (uint32_t)(S * S * zb);
        return VIS_OK;
    }

    template <typename T> __device__ typename Vec2<T>::type vis_value(int64_t i) const
    {
        typename Vec2<T>::type r;
        if (vis_type == RG_C128) {
            const double2 d = reinterpret_cast<const double2 *>(vis)[i];
            r.x = (T)d.x;
            r.y = (T)d.y;
        } else {
            const float2 f = reinterpret_cast<const float2 *>(vis)[i];
            r.x = (T)f.x;
            r.y = (T)f.y;
        }
        return r;
    }
};

template <typename T, int S, int TILE, bool COMP, typename GT>
static int launch_tiles(const BinResult &br, rg_context *timer_ctx, unsigned long long *stats,
                        const rg_wkern *wk, void *grid, int N, int ntx, int ntiles, hipStream_t s)
{
    TileSink sink;
    sink.base = grid;
    sink.gsize = N;
    sink.ntx = ntx;
    sink.ntiles = ntiles;
    sink.group_map = nullptr;
    sink.group_offset = 0;
    KernelTimer kt(timer_ctx, RG_TIMER_WPROJ_TILES, s);
    const uint64_t kbytes = (uint64_t)wk->plane_count * wk->oversampling * wk->oversampling * S *
                            S * sizeof(typename Vec2<T>::type);
    if (kbytes > 0x7fffffffull)
        return fail(RG_ERR_UNSUPPORTED, "batched W-projection kernel tables above 2 GiB");
    if (br.max_items > 0)
        hipLaunchKernelGGL((k_tiles<T, S, TILE, COMP, GT>), dim3((unsigned)br.max_items),
                           dim3(S * S), 0, s, (const Rec<T> *)br.recs, (const Item *)br.items, 0u,
                           0u, (const unsigned long long *)stats,
                           (const typename Vec2<T>::type *)wk->data, (uint32_t)kbytes, sink);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

template <typename T, int S, int TILE>
static int dispatch_comp_grid(const BinResult &br, rg_context *timer_ctx,
                              unsigned long long *stats, const rg_wkern *wk, void *grid,
                              int grid_type, bool comp, int N, int ntx, int ntiles, hipStream_t s)
{
    if constexpr (std::is_same<T, double>::value) {
        if (grid_type != RG_C128)
            return fail(RG_ERR_PARAM, "FP64 kernels need an RG_C128 grid");
        return launch_tiles<double, S, TILE, false, double>(br, timer_ctx, stats, wk, grid, N, ntx,
                                                            ntiles, s);
    } else {
        if (grid_type == RG_C128)
            return comp ? launch_tiles<float, S, TILE, true, double>(br, timer_ctx, stats, wk, grid,
                                                                      N, ntx, ntiles, s)
                        : launch_tiles<float, S, TILE, false, double>(br, timer_ctx, stats, wk,
                                                                       grid, N, ntx, ntiles, s);
        return comp ? launch_tiles<float, S, TILE, true, float>(br, timer_ctx, stats, wk, grid, N,
                                                                 ntx, ntiles, s)
                    : launch_tiles<float, S, TILE, false, float>(br, timer_ctx, stats, wk, grid, N,
                                                                  ntx, ntiles, s);
    }
}

template <typename T>
static int dispatch_tiles(const BinResult &br, rg_context *timer_ctx, unsigned long long *stats,
                          const rg_wkern *wk, void *grid, int grid_type, bool comp, int N,
                          int tile, int ntx, int ntiles, hipStream_t s)
{
    if (wk->size == 16)
        return tile == 32 ? dispatch_comp_grid<T, 16, 32>(br, timer_ctx, stats, wk, grid,
                                                          grid_type, comp, N, ntx, ntiles, s)
                          : dispatch_comp_grid<T, 16, 48>(br, timer_ctx, stats, wk, grid,
                                                          grid_type, comp, N, ntx, ntiles, s);
    return tile == 32 ? dispatch_comp_grid<T, 8, 32>(br, timer_ctx, stats, wk, grid, grid_type,
                                                     comp, N, ntx, ntiles, s)
                      : dispatch_comp_grid<T, 8, 48>(br, timer_ctx, stats, wk, grid, grid_type,
                                                     comp, N, ntx, ntiles, s);
}

// Tile geometry and sort parameters of batched W-projection.
struct WprojShape {
    int tile, ntx, ntiles;
    uint32_t per_item;
};

static int wproj_shape(const rg_grid_params *p, const rg_wkern *wk, WprojShape &sh)
{
    const int S = wk->size;
    sh.tile = p->tile > 0 ? p->tile : 32;
    if (sh.tile != 32 && sh.tile != 48)
        return fail(RG_ERR_UNSUPPORTED, "batched tile must be 32 or 48 (got %d)", sh.tile);
    if (S != 8 && S != 16)
        return fail(RG_ERR_UNSUPPORTED, "batched W-projection supports kernel sizes 8 and 16 (got %d)", S);
    sh.ntx = (p->grid_size + S + sh.tile - 1) / sh.tile;
    sh.ntiles = sh.ntx * sh.ntx;
    sh.per_item = p->vis_per_item > 0 ? (uint32_t)p->vis_per_item : 2048u;
    return RG_OK;
}

static WprojKeyer wproj_keyer(const rg_grid_params *p, const rg_flat_vis *vis, const rg_wkern *wk,
                              const WprojShape &sh, const int64_t *d_off, int n_chan, int zb)
{
    WprojKeyer kf;
    kf.u = vis->u;
    kf.v = vis->v;
    kf.w = vis->w;
    kf.vis = vis->vis;
    kf.vis_type = vis->vis_type;
    kf.N = p->grid_size;
    kf.S = wk->size;
    kf.R = wk->oversampling;
    kf.P = wk->plane_count;
    kf.theta = p->theta;
    kf.w_min = wk->w_min;
    kf.w_step = wk->w_step;
    kf.tile = sh.tile;
    kf.ntx = sh.ntx;
    kf.ntiles = sh.ntiles;
    kf.zb = zb;
    kf.rtp = (sh.tile + wk->size - 1 + wk->size - 1) / wk->size * wk->size;
    kf.chan_off = d_off;
    kf.nchan = n_chan;
    return kf;
}

template <typename T>
static int wproj_batched_t(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
                           const int64_t *d_off, int n_chan, const rg_wkern *wk, void *grid,
                           int grid_type, hipStream_t s)
{
    const int N = p->grid_size;
    WprojShape sh;
    RG_TRY(wproj_shape(p, wk, sh));

    const size_t gbytes = (size_t)(grid_type == RG_C128 ? 16 : 8) * N * N * n_chan;
    RG_TRY(zero_fill(grid, gbytes, s));

    const WprojKeyer kf =
        wproj_keyer(p, vis, wk, sh, d_off, n_chan, (int)sizeof(typename Vec2<T>::type));
    BinResult br;
    RG_TRY((bin_visibilities<WprojKeyer, T>(ctx, kf, vis->n, (int64_t)sh.ntiles * n_chan,
                                            sh.per_item, s, br)));
    return dispatch_tiles<T>(br, ctx, ctx->counters.as<unsigned long long>(), wk, grid, grid_type,
                             p->compensated != 0, N, sh.tile, sh.ntx, sh.ntiles, s);
}

int wproj_batched(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
                  const int64_t *d_off, int n_chan, const rg_wkern *wk, void *grid,
                  int grid_type, hipStream_t s)
{
    if (wk->type == RG_C64)
        return wproj_batched_t<float>(ctx, p, vis, d_off, n_chan, wk, grid, grid_type, s);
    return wproj_batched_t<double>(ctx, p, vis, d_off, n_chan, wk, grid, grid_type, s);
}

// ---- allocation-free batched W-projection ------------------------------------------------
// One file-scope device array serves every call on a device: the sort's
// counters, key starts, work items, records and per-visibility keys, and the
// call counters.  Host code never passes the array's name to a kernel: its
// device address on the current device comes from hipGetSymbolAddress, looked
// up once per device and kept.  Calls on one device share the array, so calls
// that overlap in time on one device (several streams or host threads) race.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];

static int noalloc_scratch(void **out)
{
    int dev = 0;
    RG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(RG_ERR_UNSUPPORTED, "device %d beyond the %d devices of the static scratch",
                    dev, NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        RG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr)
            return fail(RG_ERR_HIP, "no device address for the static scratch on device %d", dev);
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return RG_OK;
}

size_t noalloc_scratch_bytes() { return NOALLOC_SCRATCH_BYTES; }

template <typename T>
static int wproj_noalloc_t(const rg_grid_params *p, const rg_flat_vis *vis, const rg_wkern *wk,
                           void *grid, int grid_type, int accumulate, hipStream_t s)
{
    const int N = p->grid_size;
    WprojShape sh;
    RG_TRY(wproj_shape(p, wk, sh));
    BinPlan pl;
    RG_TRY(bin_plan(vis->n, sh.ntiles, sh.per_item, (int64_t)1 << 22, sizeof(Rec<T>), pl));
    // carve the arrays from the static scratch, 256-byte aligned
    size_t off[BIN__N + 1];
    size_t at = (sizeof(unsigned long long) * CNT__N + 255) / 256 * 256; // call counters first
    for (int k = 0; k < BIN__N; ++k) {
        off[k] = at;
        at += (pl.bytes[k] + 255) / 256 * 256;
    }
    off[BIN__N] = at;
    if (at > NOALLOC_SCRATCH_BYTES)
        return fail(RG_ERR_PARAM,
                    "rg_grid_noalloc: %lld visibilities on a %d^2 grid need %zu bytes of scratch; "
                    "the static scratch holds %zu: split the stream over calls with accumulate = 1",
                    (long long)vis->n, N, at, NOALLOC_SCRATCH_BYTES);
    void *base = nullptr;
    RG_TRY(noalloc_scratch(&base));
    char *b0 = static_cast<char *>(base);
    BinBufs b;
    b.stats = reinterpret_cast<unsigned long long *>(b0);
    b.count = reinterpret_cast<uint32_t *>(b0 + off[BIN_COUNT]);
    b.vis_start = reinterpret_cast<uint32_t *>(b0 + off[BIN_VIS_START]);
    b.item_start = reinterpret_cast<uint32_t *>(b0 + off[BIN_ITEM_START]);
    b.cursor = reinterpret_cast<uint32_t *>(b0 + off[BIN_CURSOR]);
    b.items = reinterpret_cast<Item *>(b0 + off[BIN_ITEMS]);
    b.recs = b0 + off[BIN_RECS];
    b.bsums = reinterpret_cast<uint2 *>(b0 + off[BIN_BSUMS]);
    b.krp = reinterpret_cast<uint4 *>(b0 + off[BIN_KRP]);
    // the previous call's counters are still there: start from zero
    RG_TRY(zero_fill(b.stats, sizeof(unsigned long long) * CNT__N, s));
    if (!accumulate)
        RG_TRY(zero_fill(grid, (size_t)(grid_type == RG_C128 ? 16 : 8) * N * N, s));
    if (vis->n == 0)
        return RG_OK;
    const WprojKeyer kf =
        wproj_keyer(p, vis, wk, sh, nullptr, 1, (int)sizeof(typename Vec2<T>::type));
    BinResult br;
    RG_TRY((bin_run<WprojKeyer, T>(kf, vis->n, sh.ntiles, sh.per_item, pl, b, nullptr, s, br)));
    return dispatch_tiles<T>(br, nullptr, b.stats, wk, grid, grid_type, p->compensated != 0, N,
                             sh.tile, sh.ntx, sh.ntiles, s);
}

int wproj_noalloc(const rg_grid_params *p, const rg_flat_vis *vis, const rg_wkern *wk, void *grid,
                  int grid_type, int accumulate, hipStream_t s)
{
    if (wk->type == RG_C64)
        return wproj_noalloc_t<float>(p, vis, wk, grid, grid_type, accumulate, s);
    return wproj_noalloc_t<double>(p, vis, wk, grid, grid_type, accumulate, s);
}

} // namespace rg

