// host.cpp -- C ABI front end of librg: errors, contexts, dispatch, the
// prolate spheroidal taper and the W-kernel table generator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "rg_internal.hpp"

namespace rg {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int DevBuf::ensure(size_t n)
{
    if (n <= bytes && p)
        return RG_OK;
    if (frozen && *frozen)
        return fail(RG_ERR_PARAM,
                    "this call needs %zu bytes of a scratch buffer that a captured graph uses "
                    "(%zu bytes): the context cannot reallocate it until rg_context_thaw",
                    n, bytes);
    release();
    size_t want = n < 256 ? 256 : n;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        bytes = 0;
        return fail(RG_ERR_NOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    bytes = want;
    return RG_OK;
}

void DevBuf::release()
{
    if (p)
        (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int PinnedBuf::ensure(size_t n)
{
    if (n <= bytes && p)
        return RG_OK;
    release();
    hipError_t e = hipHostMalloc(&p, n, hipHostMallocDefault);
    if (e != hipSuccess) {
        p = nullptr;
        bytes = 0;
        return fail(RG_ERR_NOMEM, "hipHostMalloc(%zu) failed: %s", n, hipGetErrorString(e));
    }
    bytes = n;
    return RG_OK;
}

void PinnedBuf::release()
{
    if (p)
        (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
}

int ensure_counters(rg_context *ctx)
{
    RG_TRY(ctx->counters.ensure(sizeof(unsigned long long) * CNT__N));
    RG_TRY(ctx->counters_host.ensure(sizeof(unsigned long long) * CNT__N));
    return RG_OK;
}

static hipEvent_t pool_event(rg_context *ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

KernelTimer::KernelTimer(rg_context *c, int w, hipStream_t st) : ctx(c), which(w), s(st)
{
    if (ctx && ctx->timing) {
        start = pool_event(ctx);
        if (start)
            (void)hipEventRecord(start, s);
    }
}

KernelTimer::~KernelTimer()
{
    if (!start)
        return;
    hipEvent_t stop = pool_event(ctx);
    if (!stop) {
        ctx->event_pool.push_back(start);
        return;
    }
    (void)hipEventRecord(stop, s);
    ctx->timer_events[which].emplace_back(start, stop);
}

int dft_run(rg_context *ctx, const rg_flat_vis *vis, const int32_t *px, const int32_t *py,
            int64_t npix, int N, double lambda, void *out, int mode, hipStream_t s);
int predict_run(const double *u, const double *v, const double *w, int64_t n, const double *l,
                const double *m, const double *flux, int nsrc, void *out, int type, hipStream_t s);

// ---- validation ----------------------------------------------------------------------

static int check_vis(const rg_flat_vis *vis)
{
    if (!vis)
        return fail(RG_ERR_PARAM, "visibility descriptor is NULL");
    if (vis->n < 0)
        return fail(RG_ERR_PARAM, "negative visibility count");
    if (vis->n > 0 && (!vis->u || !vis->v || !vis->w || !vis->vis))
        return fail(RG_ERR_PARAM, "visibility arrays must not be NULL");
    if (vis->vis_type != RG_C128 && vis->vis_type != RG_C64)
        return fail(RG_ERR_PARAM, "unknown visibility type %d", vis->vis_type);
    return RG_OK;
}

static int check_wkern(const rg_wkern *wk)
{
    if (!wk || !wk->data)
        return fail(RG_ERR_PARAM, "W-kernel table is NULL");
    if (wk->type != RG_C128 && wk->type != RG_C64)
        return fail(RG_ERR_PARAM, "unknown W-kernel type %d", wk->type);
    if (wk->plane_count <= 0 || wk->oversampling <= 0 || wk->size <= 0)
        return fail(RG_ERR_SHAPE, "invalid W-kernel dimensions P=%d R=%d S=%d", wk->plane_count,
                    wk->oversampling, wk->size);
    if (wk->plane_count > 1 && !(wk->w_step > 0))
        return fail(RG_ERR_PARAM, "W-kernel w_step must be positive");
    if ((int64_t)wk->plane_count * wk->oversampling * wk->oversampling > 0x7fffffffLL / 4)
        return fail(RG_ERR_SHAPE, "W-kernel table too large");
    return RG_OK;
}

static int check_params(const rg_grid_params *p, const rg_wkern *wk)
{
    if (!p)
        return fail(RG_ERR_PARAM, "parameters are NULL");
    const int N = p->grid_size;
    if (N <= 0 || (N & 1))
        return fail(RG_ERR_PARAM, "grid size must be positive and even (got %d)", N);
    if (!(p->theta > 0))
        return fail(RG_ERR_PARAM, "theta must be positive");
    if (p->mode != RG_MODE_PLAIN && p->mode != RG_MODE_BATCHED)
        return fail(RG_ERR_PARAM, "unknown mode %d", p->mode);
    if (p->subgrid_size != 0) {
        const int n = p->subgrid_size, M = p->subgrid_margin;
        // grid_cuda.cu:536 and grid.c:300-302
        if (n <= 0 || N % n != 0)
            return fail(RG_ERR_PARAM, "grid size %d is not a multiple of subgrid size %d", N, n);
        if (n & (n - 1))
            return fail(RG_ERR_UNSUPPORTED, "subgrid size must be a power of two (got %d)", n);
        if (M < wk->size)
            return fail(RG_ERR_PARAM, "subgrid margin %d smaller than kernel size %d", M, wk->size);
        if ((n % 2) || (M % 2) || M >= n)
            return fail(RG_ERR_PARAM, "subgrid size and margin must be even, margin < size");
        if (!(p->winc > 0))
            return fail(RG_ERR_PARAM, "winc must be positive");
    }
    if (p->fft_radix != 0 && p->fft_radix != 4 && p->fft_radix != 8)
        return fail(RG_ERR_PARAM, "fft_radix must be 0, 4 or 8 (got %d)", p->fft_radix);
    if (p->vis_per_item < 0 || p->floor_batch < 0)
        return fail(RG_ERR_PARAM, "vis_per_item and floor_batch must not be negative");
    return RG_OK;
}

} // namespace rg

using namespace rg;

// ---- C ABI -----------------------------------------------------------------------------

namespace rg {

int channel_offsets(rg_context *ctx, const int64_t *offsets, int n_chan, hipStream_t s,
                    const int64_t **d_off)
{
    *d_off = nullptr;
    if (n_chan <= 1)
        return RG_OK;
    // resident copy of the offsets: a call with the offsets of the previous
    // multi-channel call issues no copy (and so can be graph-captured)
    const bool same = ctx->chan_off.size() == (size_t)n_chan + 1 &&
                      std::equal(ctx->chan_off.begin(), ctx->chan_off.end(), offsets);
    if (!same) {
        if (ctx->frozen)
            return fail(RG_ERR_PARAM,
                        "a captured graph reads this context's channel offsets: calls with other "
                        "offsets need rg_context_thaw (or another context)");
        RG_TRY(ctx->scratch[13].ensure(sizeof(int64_t) * (n_chan + 1)));
        RG_HIP(hipMemcpyAsync(ctx->scratch[13].p, offsets, sizeof(int64_t) * (n_chan + 1),
                              hipMemcpyHostToDevice, s));
        RG_HIP(hipStreamSynchronize(s)); // the host array may go away after the call
        ctx->chan_off.assign(offsets, offsets + n_chan + 1);
    }
    *d_off = ctx->scratch[13].as<int64_t>();
    return RG_OK;
}

} // namespace rg

extern "C" {

const char *rg_last_error(void) { return g_last_error.c_str(); }

int rg_abi_version(void) { return RG_ABI_VERSION; }

int rg_context_create(rg_context **ctx, int device)
{
    if (!ctx)
        return fail(RG_ERR_PARAM, "ctx is NULL");
    *ctx = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(RG_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= count)
        return fail(RG_ERR_PARAM, "device %d out of range (%d devices)", device, count);
    RG_HIP(hipSetDevice(device));
    rg_context *c = new rg_context();
    c->device = device;
    for (auto &buf : c->scratch)
        buf.frozen = &c->frozen;
    int st = ensure_counters(c);
    if (st != RG_OK) {
        delete c;
        return st;
    }
    *ctx = c;
    return RG_OK;
}

int rg_context_set_timing(rg_context *ctx, int32_t enable)
{
    if (!ctx)
        return fail(RG_ERR_PARAM, "ctx is NULL");
    ctx->timing = enable != 0;
    return RG_OK;
}

int rg_context_timer(rg_context *ctx, int32_t which, double *total_ms, int64_t *launches)
{
    if (!ctx || !total_ms || !launches)
        return fail(RG_ERR_PARAM, "NULL argument");
    if (which < 0 || which >= RG_TIMER__COUNT)
        return fail(RG_ERR_PARAM, "unknown timer %d", which);
    RG_HIP(hipSetDevice(ctx->device));
    double tot = 0.0;
    int64_t cnt = 0;
    for (auto &pr : ctx->timer_events[which]) {
        RG_HIP(hipEventSynchronize(pr.second));
        float ms = 0.f;
        RG_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
        tot += ms;
        ++cnt;
        ctx->event_pool.push_back(pr.first);
        ctx->event_pool.push_back(pr.second);
    }
    ctx->timer_events[which].clear();
    *total_ms = tot;
    *launches = cnt;
    return RG_OK;
}

int rg_context_destroy(rg_context *ctx)
{
    if (!ctx)
        return RG_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (auto &v : ctx->timer_events)
        for (auto &pr : v) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (auto e : ctx->event_pool)
        (void)hipEventDestroy(e);
    ctx->counters.release();
    ctx->counters_host.release();
    for (auto &b : ctx->scratch)
        b.release();
    for (auto &kv : ctx->twiddles)
        kv.second.release();
    delete ctx;
    return RG_OK;
}

int rg_context_stats(rg_context *ctx, rg_grid_stats *out)
{
    if (!ctx || !out)
        return fail(RG_ERR_PARAM, "NULL argument");
    RG_HIP(hipSetDevice(ctx->device));
    RG_HIP(hipMemcpyAsync(ctx->counters_host.p, ctx->counters.p,
                          sizeof(unsigned long long) * CNT__N, hipMemcpyDeviceToHost,
                          ctx->last_stream));
    RG_HIP(hipStreamSynchronize(ctx->last_stream));
    const unsigned long long *c = (const unsigned long long *)ctx->counters_host.p;
    out->n_vis = ctx->last_nvis;
    out->n_dropped = (int64_t)c[CNT_DROPPED];
    out->n_bad_plane = (int64_t)c[CNT_BAD_PLANE];
    out->n_items = (int64_t)c[CNT_ITEMS];
    out->n_tasks = (int64_t)c[CNT_TASKS];
    out->n_towers = (int64_t)c[CNT_TOWERS];
    out->n_records = (int64_t)c[CNT_RECORDS];
    out->n_bands = (int64_t)c[CNT_BANDS];
    out->n_chunks = (int64_t)c[CNT_CHUNKS];
    // every visibility is gridded, dropped (off the grid / tower set) or has
    // no W-plane
    out->n_gridded = ctx->last_nvis - out->n_dropped - out->n_bad_plane;
    if (c[CNT_OVERFLOW])
        return fail(RG_ERR_RANGE,
                    "the last call overran %llu internal capacity bound(s): its result is "
                    "incomplete", (unsigned long long)c[CNT_OVERFLOW]);
    return RG_OK;
}

int rg_context_thaw(rg_context *ctx)
{
    if (!ctx)
        return fail(RG_ERR_PARAM, "ctx is NULL");
    ctx->frozen = false;
    return RG_OK;
}

static int grid_common(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
                       const int64_t *offsets, int32_t n_chan, const rg_wkern *wk, void *grid,
                       int32_t grid_type, void *stream)
{
    if (!ctx)
        return fail(RG_ERR_PARAM, "context is NULL");
    RG_TRY(check_vis(vis));
    RG_TRY(check_wkern(wk));
    RG_TRY(check_params(p, wk));
    if (!grid)
        return fail(RG_ERR_PARAM, "grid is NULL");
    if (grid_type != RG_C128 && grid_type != RG_C64)
        return fail(RG_ERR_PARAM, "unknown grid type %d", grid_type);
    hipStream_t s = (hipStream_t)stream;
    RG_HIP(hipSetDevice(ctx->device));
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    RG_HIP(hipStreamIsCapturing(s, &cap));
    const bool capturing = cap == hipStreamCaptureStatusActive;
    if (capturing)
        ctx->frozen = true; // the graph will hold this context's scratch addresses
    RG_TRY(zero_fill(ctx->counters.p, sizeof(unsigned long long) * CNT__N, s));
    ctx->last_stream = s;
    ctx->last_nvis = vis->n;
    if (p->mode == RG_MODE_PLAIN && n_chan != 1)
        return fail(RG_ERR_UNSUPPORTED, "plain mode grids one channel per call");
    const int64_t *d_off = nullptr;
    RG_TRY(channel_offsets(ctx, offsets, n_chan, s, &d_off));
    if (p->subgrid_size == 0) {
        ctx->last_kind = 1;
        if (p->mode == RG_MODE_PLAIN) {
            if (grid_type != RG_C128)
                return fail(RG_ERR_PARAM, "plain mode writes an RG_C128 grid");
            return wproj_plain(ctx, p, vis, wk, grid, s);
        }
        return wproj_batched(ctx, p, vis, d_off, n_chan, wk, grid, grid_type, s);
    }
    ctx->last_kind = 2;
    return wtowers_run(ctx, p, vis, d_off, n_chan, wk, grid, grid_type, s, capturing);
}

int rg_grid(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis, const rg_wkern *wk,
            void *grid, int32_t grid_type, void *stream)
{
    int64_t off[2] = {0, vis ? vis->n : 0};
    return grid_common(ctx, p, vis, off, 1, wk, grid, grid_type, stream);
}

int rg_grid_channels(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
                     const int64_t *offsets, int32_t n_chan, const rg_wkern *wk, void *grids,
                     int32_t grid_type, void *stream)
{
    if (!offsets || n_chan < 1)
        return fail(RG_ERR_PARAM, "need n_chan >= 1 and channel offsets");
    if (!vis)
        return fail(RG_ERR_PARAM, "visibility descriptor is NULL");
    if (offsets[0] != 0 || offsets[n_chan] != vis->n)
        return fail(RG_ERR_PARAM, "channel offsets must span [0, n)");
    for (int c = 0; c < n_chan; ++c)
        if (offsets[c + 1] < offsets[c])
            return fail(RG_ERR_PARAM, "channel offsets must be non-decreasing");
    if (p && p->mode != RG_MODE_BATCHED)
        return fail(RG_ERR_UNSUPPORTED, "multi-channel gridding is a batched-mode entry point");
    return grid_common(ctx, p, vis, offsets, n_chan, wk, grids, grid_type, stream);
}

int rg_grid_noalloc(const rg_grid_params *p, const rg_flat_vis *vis, const rg_wkern *wk,
                    void *grid, int32_t grid_type, int32_t accumulate, void *stream)
{
    RG_TRY(check_vis(vis));
    RG_TRY(check_wkern(wk));
    RG_TRY(check_params(p, wk));
    if (!grid)
        return fail(RG_ERR_PARAM, "grid is NULL");
    if (grid_type != RG_C128 && grid_type != RG_C64)
        return fail(RG_ERR_PARAM, "unknown grid type %d", grid_type);
    if (p->subgrid_size != 0 || p->mode != RG_MODE_BATCHED)
        return fail(RG_ERR_UNSUPPORTED,
                    "rg_grid_noalloc grids batched W-projection only (subgrid_size 0, "
                    "RG_MODE_BATCHED)");
    return wproj_noalloc(p, vis, wk, grid, grid_type, accumulate != 0, (hipStream_t)stream);
}

int rg_grid_to_image(rg_context *ctx, void *grid, int32_t type, int32_t n, void *stream)
{
    if (!ctx || !grid)
        return fail(RG_ERR_PARAM, "NULL argument");
    if (type != RG_C128 && type != RG_C64)
        return fail(RG_ERR_PARAM, "unknown grid type %d", type);
    RG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    RG_TRY(fft_shift(grid, type, n, 1, s));
    RG_TRY(fft2d(ctx, grid, type, n, 1, +1, s));
    return fft_shift(grid, type, n, 1, s);
}

int rg_fft2d(rg_context *ctx, void *data, int32_t type, int32_t n, int32_t batch, int32_t sign,
             void *stream)
{
    if (!ctx || !data)
        return fail(RG_ERR_PARAM, "NULL argument");
    if (type != RG_C128 && type != RG_C64)
        return fail(RG_ERR_PARAM, "unknown type %d", type);
    RG_HIP(hipSetDevice(ctx->device));
    return fft2d(ctx, data, type, n, batch, sign, (hipStream_t)stream);
}

int rg_fft_shift(void *data, int32_t type, int32_t n, void *stream)
{
    if (!data)
        return fail(RG_ERR_PARAM, "NULL argument");
    if (type != RG_C128 && type != RG_C64)
        return fail(RG_ERR_PARAM, "unknown type %d", type);
    return fft_shift(data, type, n, 1, (hipStream_t)stream);
}

int rg_weight(rg_context *ctx, const double *u, const double *v, int64_t n, void *vis,
              int32_t vis_type, int32_t grid_size, double theta, uint32_t *counts, void *stream)
{
    if (!ctx)
        return fail(RG_ERR_PARAM, "NULL context");
    if (n < 0 || (n > 0 && (!u || !v || !vis)))
        return fail(RG_ERR_PARAM, "invalid visibility arrays");
    if (vis_type != RG_C128 && vis_type != RG_C64)
        return fail(RG_ERR_PARAM, "unknown visibility type %d", vis_type);
    if (grid_size <= 0 || grid_size > 65536 || !(theta > 0))
        return fail(RG_ERR_PARAM, "invalid grid geometry (N %d, theta %g)", grid_size, theta);
    RG_HIP(hipSetDevice(ctx->device));
    return weight_vis(ctx, u, v, n, vis, vis_type, grid_size, theta, counts, (hipStream_t)stream);
}

int rg_make_hermitian(void *grid, int32_t type, int32_t n, void *stream)
{
    if (!grid)
        return fail(RG_ERR_PARAM, "NULL argument");
    if (type != RG_C128 && type != RG_C64)
        return fail(RG_ERR_PARAM, "unknown type %d", type);
    if (n <= 0 || n > 65536)
        return fail(RG_ERR_PARAM, "invalid grid size %d", n);
    return make_hermitian(grid, type, n, (hipStream_t)stream);
}

int rg_gcf_correct(void *image, int32_t type, int32_t n, const double *taper, void *stream)
{
    if (!image || !taper)
        return fail(RG_ERR_PARAM, "NULL argument");
    if (type != RG_C128 && type != RG_C64)
        return fail(RG_ERR_PARAM, "unknown type %d", type);
    if (n <= 0 || n > 65536)
        return fail(RG_ERR_PARAM, "invalid grid size %d", n);
    return gcf_correct(image, type, n, taper, (hipStream_t)stream);
}

static int check_sep(const rg_sep_kern *k, const char *what, int support)
{
    if (!k || !k->data || k->size <= 0 || k->oversampling <= 0)
        return fail(RG_ERR_PARAM, "invalid %s kernel", what);
    if (support > 0 && k->size != support)
        return fail(RG_ERR_PARAM, "%s kernel size %d differs from its support %d", what, k->size,
                    support);
    return RG_OK;
}

static int check_wstack(const rg_wstack_params *p)
{
    if (!p)
        return fail(RG_ERR_PARAM, "NULL parameters");
    if (!(p->theta > 0) || !(p->lambda > 0) || !(p->du > 0) || !(p->dw > 0) || !(p->x0 > 0) ||
        p->aa_support_uv < 2 || p->aa_support_uv > 64 ||
        (p->aa_support_uv & (p->aa_support_uv - 1)) != 0 || p->aa_support_w < 2 ||
        p->aa_support_w > 64 || p->aa_support_w % 2)
        return fail(RG_ERR_PARAM, "invalid W-stack parameters (uv support a power of two in "
                                  "2..64, even w support <= 64)");
    return RG_OK;
}

int rg_wstack_build(rg_context *ctx, const rg_wstack_params *p, const double *l, const double *m,
                    const double *flux, int32_t npts, const rg_sep_kern *corr_lm,
                    const rg_sep_kern *corr_n, int32_t *og_out, int32_t *planes_out,
                    int64_t *skipped_out, void *stream)
{
    if (!ctx)
        return fail(RG_ERR_PARAM, "NULL context");
    RG_TRY(check_wstack(p));
    RG_TRY(check_sep(corr_lm, "l/m correction", 0));
    RG_TRY(check_sep(corr_n, "n correction", 0));
    if (npts < 0 || (npts > 0 && (!l || !m)))
        return fail(RG_ERR_PARAM, "invalid sky arrays");
    RG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    ctx->last_stream = s;
    RG_TRY(wstack_build(ctx, p, l, m, flux, npts, corr_lm, corr_n, s));
    if (og_out)
        *og_out = ctx->ws_og;
    if (planes_out)
        *planes_out = ctx->ws_planes;
    if (skipped_out) {
        unsigned long long c = 0;
        RG_HIP(hipMemcpyAsync(&c, ctx->counters.as<unsigned long long>() + CNT_SKY_SKIPPED,
                              sizeof(c), hipMemcpyDeviceToHost, s));
        RG_HIP(hipStreamSynchronize(s));
        *skipped_out = (int64_t)c;
    }
    return RG_OK;
}

int rg_wstack_degrid(rg_context *ctx, const rg_wstack_params *p, const double *u, const double *v,
                     const double *w, int64_t nvis, const rg_sep_kern *conv_uv,
                     const rg_sep_kern *conv_w, void *out, void *stream)
{
    if (!ctx)
        return fail(RG_ERR_PARAM, "NULL context");
    RG_TRY(check_wstack(p));
    RG_TRY(check_sep(conv_uv, "uv", p->aa_support_uv));
    RG_TRY(check_sep(conv_w, "w", p->aa_support_w));
    if (nvis < 0 || (nvis > 0 && (!u || !v || !w || !out)))
        return fail(RG_ERR_PARAM, "invalid visibility arrays");
    RG_HIP(hipSetDevice(ctx->device));
    ctx->last_stream = (hipStream_t)stream;
    return wstack_degrid(ctx, p, u, v, w, nvis, conv_uv, conv_w, out, (hipStream_t)stream);
}

int rg_wstack_copy(rg_context *ctx, void *dst, int64_t capacity_bytes, void *stream)
{
    if (!ctx || !dst)
        return fail(RG_ERR_PARAM, "NULL argument");
    if (ctx->ws_og == 0)
        return fail(RG_ERR_PARAM, "no W-stack has been built in the context");
    const int64_t bytes = (int64_t)sizeof(double2) * ctx->ws_og * ctx->ws_og * ctx->ws_planes;
    if (capacity_bytes < bytes)
        return fail(RG_ERR_PARAM, "destination holds %lld bytes, the stack %lld",
                    (long long)capacity_bytes, (long long)bytes);
    RG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    RG_HIP(hipMemcpyAsync(dst, ctx->scratch[17].p, (size_t)bytes, hipMemcpyDefault, s));
    RG_HIP(hipStreamSynchronize(s));
    return RG_OK;
}

int rg_wstack_predict(rg_context *ctx, const rg_wstack_params *p, const double *l, const double *m,
                      const double *flux, int32_t npts, const double *u, const double *v,
                      const double *w, int64_t nvis, const rg_sep_kern *conv_uv,
                      const rg_sep_kern *conv_w, const rg_sep_kern *corr_lm,
                      const rg_sep_kern *corr_n, void *out, void *stream)
{
    RG_TRY(rg_wstack_build(ctx, p, l, m, flux, npts, corr_lm, corr_n, nullptr, nullptr, nullptr,
                           stream));
    return rg_wstack_degrid(ctx, p, u, v, w, nvis, conv_uv, conv_w, out, stream);
}

int rg_epic_grid(rg_context *ctx, const void *fdata, const int32_t *x, const int32_t *y,
                 const int32_t *z, int32_t n_ant, int32_t nbatch, const void *illum,
                 int32_t support, int32_t grid_size, void *grids, void *stream)
{
    (void)z;
    if (!ctx || !grids || !illum)
        return fail(RG_ERR_PARAM, "NULL argument");
    if (n_ant < 0 || nbatch < 0 || (n_ant > 0 && nbatch > 0 && (!fdata || !x || !y)))
        return fail(RG_ERR_PARAM, "invalid E-field arrays");
    if (support < 1 || support > 64 || grid_size < 1 || grid_size > 4096 ||
        (int64_t)n_ant * support * support > ((int64_t)1 << 30))
        return fail(RG_ERR_PARAM, "invalid EPIC geometry (support %d, grid %d)", support,
                    grid_size);
    RG_HIP(hipSetDevice(ctx->device));
    if (nbatch == 0)
        return RG_OK;
    return epic_grid(fdata, x, y, n_ant, nbatch, illum, support, grid_size, grids,
                     (hipStream_t)stream);
}

int rg_dft(rg_context *ctx, const rg_flat_vis *vis, const int32_t *px, const int32_t *py,
           int64_t npix, int32_t grid_size, double lambda, void *out, int32_t mode, void *stream)
{
    if (!ctx || !out || !px || !py)
        return fail(RG_ERR_PARAM, "NULL argument");
    RG_TRY(check_vis(vis));
    if (grid_size <= 0 || !(lambda > 0) || npix < 0)
        return fail(RG_ERR_PARAM, "invalid DFT geometry");
    if (mode != RG_DFT_FP64 && mode != RG_DFT_MIXED)
        return fail(RG_ERR_PARAM, "unknown DFT mode %d", mode);
    RG_HIP(hipSetDevice(ctx->device));
    return dft_run(ctx, vis, px, py, npix, grid_size, lambda, out, mode, (hipStream_t)stream);
}

int rg_dft_image(rg_context *ctx, const rg_flat_vis *vis, int32_t grid_size, double lambda,
                 void *out, int32_t blocks, int32_t threads_block, int32_t mode, void *stream)
{
    if (!ctx || !out)
        return fail(RG_ERR_PARAM, "NULL argument");
    RG_TRY(check_vis(vis));
    if (grid_size <= 0 || !(lambda > 0))
        return fail(RG_ERR_PARAM, "invalid DFT geometry");
    // README.md:84-86: blocks * threads must cover the image exactly.
    if ((int64_t)blocks * threads_block != (int64_t)grid_size * grid_size)
        return fail(RG_ERR_PARAM, "blocks * threads_block (%lld) must equal grid_size^2 (%lld)",
                    (long long)blocks * threads_block, (long long)grid_size * grid_size);
    if (mode != RG_DFT_FP64 && mode != RG_DFT_MIXED)
        return fail(RG_ERR_PARAM, "unknown DFT mode %d", mode);
    RG_HIP(hipSetDevice(ctx->device));
    return dft_run(ctx, vis, nullptr, nullptr, (int64_t)grid_size * grid_size, grid_size, lambda,
                   out, mode, (hipStream_t)stream);
}

int rg_predict_points(const double *u, const double *v, const double *w, int64_t n,
                      const double *l, const double *m, const double *flux, int32_t n_src,
                      void *vis_out, int32_t vis_type, void *stream)
{
    if (n < 0 || n_src < 0)
        return fail(RG_ERR_PARAM, "negative sizes");
    if (n > 0 && (!u || !v || !w || !vis_out))
        return fail(RG_ERR_PARAM, "NULL argument");
    if (n_src > 0 && (!l || !m || !flux))
        return fail(RG_ERR_PARAM, "NULL source arrays");
    if (vis_type != RG_C128 && vis_type != RG_C64)
        return fail(RG_ERR_PARAM, "unknown visibility type %d", vis_type);
    return predict_run(u, v, w, n, l, m, flux, n_src, vis_out, vis_type, (hipStream_t)stream);
}

// ---- prolate spheroidal taper -------------------------------------------------------------
// psi_0(c, x) = sum_i d_i Pbar_{2i}(x), Pbar_k = sqrt(k + 1/2) P_k, where d is the
// eigenvector of the smallest eigenvalue of the symmetric tridiagonal matrix of
// L = -d/dx (1 - x^2) d/dx + c^2 x^2 in the even normalised Legendre basis:
//   T[i][i]   = k(k+1) + c^2 (2k^2 + 2k - 1) / ((2k - 1)(2k + 3)),   k = 2i
//   T[i][i+1] = c^2 (k+1)(k+2) / ((2k+3) sqrt((2k+1)(2k+5))).
// Eigenpairs by cyclic Jacobi rotations (the matrix is small).

static void pswf_coeffs(double c, std::vector<double> &d)
{
    const int K = 24 + (int)(1.5 * c);
    std::vector<double> A((size_t)K * K, 0.0), V((size_t)K * K, 0.0);
    for (int i = 0; i < K; ++i) {
        const double k = 2.0 * i;
        A[(size_t)i * K + i] = k * (k + 1) + c * c * (2 * k * k + 2 * k - 1) / ((2 * k - 1) * (2 * k + 3));
        if (i + 1 < K) {
            const double off = c * c * (k + 1) * (k + 2) / ((2 * k + 3) * std::sqrt((2 * k + 1) * (2 * k + 5)));
            A[(size_t)i * K + i + 1] = off;
            A[(size_t)(i + 1) * K + i] = off;
        }
        V[(size_t)i * K + i] = 1.0;
    }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double offn = 0;
        for (int p = 0; p < K; ++p)
            for (int q = p + 1; q < K; ++q)
                offn += A[(size_t)p * K + q] * A[(size_t)p * K + q];
        if (offn < 1e-30)
            break;
        for (int p = 0; p < K; ++p) {
            for (int q = p + 1; q < K; ++q) {
                const double apq = A[(size_t)p * K + q];
                if (std::fabs(apq) < 1e-300)
                    continue;
                const double app = A[(size_t)p * K + p], aqq = A[(size_t)q * K + q];
                const double tau = (aqq - app) / (2 * apq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1 + tau * tau));
                const double cs = 1 / std::sqrt(1 + t * t), sn = t * cs;
                for (int r = 0; r < K; ++r) {
                    const double arp = A[(size_t)r * K + p], arq = A[(size_t)r * K + q];
                    A[(size_t)r * K + p] = cs * arp - sn * arq;
                    A[(size_t)r * K + q] = sn * arp + cs * arq;
                }
                for (int r = 0; r < K; ++r) {
                    const double apr = A[(size_t)p * K + r], aqr = A[(size_t)q * K + r];
                    A[(size_t)p * K + r] = cs * apr - sn * aqr;
                    A[(size_t)q * K + r] = sn * apr + cs * aqr;
                }
                for (int r = 0; r < K; ++r) {
                    const double vrp = V[(size_t)r * K + p], vrq = V[(size_t)r * K + q];
                    V[(size_t)r * K + p] = cs * vrp - sn * vrq;
                    V[(size_t)r * K + q] = sn * vrp + cs * vrq;
                }
            }
        }
    }
    int best = 0;
    for (int i = 1; i < K; ++i)
        if (A[(size_t)i * K + i] < A[(size_t)best * K + best])
            best = i;
    d.resize(K);
    for (int i = 0; i < K; ++i)
        d[i] = V[(size_t)i * K + best];
}

static double pswf_eval(const std::vector<double> &d, double x)
{
    // sum_i d_i sqrt(2i + 1/2) P_{2i}(x), Legendre recurrence
    double p0 = 1.0, p1 = x, s = d[0] * std::sqrt(0.5) * p0;
    const int K = (int)d.size();
    for (int k = 1; k < 2 * K - 1; ++k) {
        const double p2 = ((2.0 * k + 1) * x * p1 - k * p0) / (k + 1);
        p0 = p1;
        p1 = p2;
        if ((k + 1) % 2 == 0)
            s += d[(k + 1) / 2] * std::sqrt(k + 1 + 0.5) * p1;
    }
    return s;
}

double rg_pswf(double c, double x)
{
    std::vector<double> d;
    pswf_coeffs(c, d);
    return pswf_eval(d, x) / pswf_eval(d, 0.0);
}

int rg_pswf_eval(double c, const double *x, int64_t n, double *out)
{
    if (n < 0 || (n > 0 && (!x || !out)))
        return fail(RG_ERR_PARAM, "invalid arguments");
    std::vector<double> d;
    pswf_coeffs(c, d);
    const double a0 = pswf_eval(d, 0.0);
    for (int64_t i = 0; i < n; ++i)
        out[i] = pswf_eval(d, x[i]) / a0;
    return RG_OK;
}

int rg_wkern_generate(void *out_host, int32_t type, int32_t planes, int32_t R, int32_t S,
                      double theta, double w_min, double w_step, int32_t M, double c,
                      int32_t nthreads)
{
    if (!out_host)
        return fail(RG_ERR_PARAM, "output is NULL");
    if (type != RG_C128 && type != RG_C64)
        return fail(RG_ERR_PARAM, "unknown type %d", type);
    if (planes <= 0 || R <= 0 || S <= 0 || !(theta > 0))
        return fail(RG_ERR_PARAM, "invalid W-kernel geometry");
    if (M <= 0)
        M = 128;
    if (c <= 0)
        c = M_PI * S / 4.0;
    if (nthreads <= 0)
        nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    using cd = std::complex<double>;
    std::vector<double> coef;
    pswf_coeffs(c, coef);
    const double a0 = pswf_eval(coef, 0.0);
    std::vector<double> A(M);
    // image samples at the cell centres l_j = theta (j + 1/2 - M/2) / M, symmetric about 0
    std::vector<double> xs(M);
    for (int j = 0; j < M; ++j)
        xs[j] = ((double)j + 0.5 - 0.5 * M) / M; // l_j / theta
    for (int j = 0; j < M; ++j)
        A[j] = pswf_eval(coef, 2.0 * xs[j]) / a0;
    const int SR = S * R;
    // E[t][j] = exp(+2 pi i d_t l_j / theta), t = xf*S + x, d_t = x - S/2 - xf/R
    std::vector<cd> E((size_t)SR * M);
    for (int xf = 0; xf < R; ++xf)
        for (int x = 0; x < S; ++x) {
            const double d = (double)(x - S / 2) - (double)xf / R;
            for (int j = 0; j < M; ++j) {
                const double ph = 2.0 * M_PI * d * xs[j];
                E[((size_t)xf * S + x) * M + j] = cd(std::cos(ph), std::sin(ph));
            }
        }
    std::vector<double> lsq(M);
    for (int j = 0; j < M; ++j) {
        const double l = theta * xs[j];
        lsq[j] = l * l;
    }
    const size_t plane_elems = (size_t)R * R * S * S;
    auto work = [&](int p0, int p1) {
        std::vector<cd> img((size_t)M * M), T((size_t)M * SR);
        for (int p = p0; p < p1; ++p) {
            const double w = w_min + p * w_step;
            for (int k = 0; k < M; ++k)
                for (int j = 0; j < M; ++j) {
                    const double nn = std::sqrt(1.0 - lsq[j] - lsq[k]);
                    const double ph = -2.0 * M_PI * w * (nn - 1.0);
                    img[(size_t)k * M + j] = cd(std::cos(ph), std::sin(ph)) * (A[j] * A[k] / ((double)M * M));
                }
            // T[k][t] = sum_j img[k][j] E[t][j]
            for (int k = 0; k < M; ++k)
                for (int t = 0; t < SR; ++t) {
                    cd acc = 0;
                    const cd *e = &E[(size_t)t * M];
                    const cd *im = &img[(size_t)k * M];
                    for (int j = 0; j < M; ++j)
                        acc += im[j] * e[j];
                    T[(size_t)k * SR + t] = acc;
                }
            // K[ty][tx] = sum_k E[ty][k] T[k][tx]
            for (int ty = 0; ty < SR; ++ty) {
                const int yf = ty / S, y = ty % S;
                const cd *e = &E[(size_t)ty * M];
                for (int tx = 0; tx < SR; ++tx) {
                    const int xf = tx / S, x = tx % S;
                    cd acc = 0;
                    for (int k = 0; k < M; ++k)
                        acc += e[k] * T[(size_t)k * SR + tx];
                    const size_t o = (size_t)p * plane_elems +
                                     (((size_t)yf * R + xf) * S + y) * S + x;
                    if (type == RG_C128) {
                        ((double *)out_host)[2 * o] = acc.real();
                        ((double *)out_host)[2 * o + 1] = acc.imag();
                    } else {
                        ((float *)out_host)[2 * o] = (float)acc.real();
                        ((float *)out_host)[2 * o + 1] = (float)acc.imag();
                    }
                }
            }
        }
    };
    const int nt = std::min(nthreads, planes);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        const int p0 = (int)((int64_t)planes * t / nt), p1 = (int)((int64_t)planes * (t + 1) / nt);
        th.emplace_back(work, p0, p1);
    }
    for (auto &t : th)
        t.join();
    return RG_OK;
}

} // extern "C"
