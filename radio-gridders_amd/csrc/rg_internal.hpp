// rg_internal.hpp -- shared definitions of librg (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "rg.h"

namespace rg {

// ---- errors --------------------------------------------------------------

void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

#define RG_HIP(expr)                                                                       \
    do {                                                                                   \
        hipError_t rg_e_ = (expr);                                                         \
        if (rg_e_ != hipSuccess)                                                           \
            return ::rg::fail(RG_ERR_HIP, "%s:%d: %s: %s", __FILE__, __LINE__, #expr,      \
                              hipGetErrorString(rg_e_));                                   \
    } while (0)

#define RG_TRY(expr)                                                                       \
    do {                                                                                   \
        int rg_s_ = (expr);                                                                \
        if (rg_s_ != RG_OK)                                                                \
            return rg_s_;                                                                  \
    } while (0)

// ---- coordinates -----------------------------------------------------------
// FP64 index math, written exactly as the reference so that host and device
// round identically: contraction is off inside these functions.

// frac_coord (src/reference/grid.c:37-57, src/common/radio.cuh:156-176):
// footprint top-left (gx, gy) and oversample indices (xf, yf).
__host__ __device__ inline void frac_coord(int N, int S, int R, double theta, double u,
                                           double v, double du, double dv, int &gx, int &gy,
                                           int &xf, int &yf)
{
#pragma clang fp contract(off)
    double x = theta * (u - du);
    double y = theta * (v - dv);
    int flx = (int)floor(x + .5 / R);
    int fly = (int)floor(y + .5 / R);
    xf = (int)floor((x - (double)flx) * R + .5);
    yf = (int)floor((y - (double)fly) * R + .5);
    // x + .5/R can round to the next integer, leaving (x - flx) R + .5 just
    // below 0 (or reaching R): keep the oversample index inside the table.
    xf = xf < 0 ? 0 : (xf >= R ? R - 1 : xf);
    yf = yf < 0 ? 0 : (yf >= R ? R - 1 : yf);
    gx = flx + N / 2 - S / 2;
    gy = fly + N / 2 - S / 2;
}

// W-plane index (grid.c:168-169, grid_cuda.cu:84).
__host__ __device__ inline int w_plane(double w, double w_min, double w_step)
{
#pragma clang fp contract(off)
    return (int)floor((w - w_min) / w_step + .5);
}

// W-towers chunk index (radio.cuh:361-362): double division, truncated.
__host__ __device__ inline int tower_index(double u, double theta, int N, int chunk)
{
#pragma clang fp contract(off)
    double q = (floor(u * theta + 0.5) + N / 2) / chunk;
    return (int)q;
}

// W-towers floor index (grid.c:453-455).
__host__ __device__ inline int floor_index(double w, double winc)
{
#pragma clang fp contract(off)
    return (int)floor(w / winc + 0.5);
}

// ---- device buffers ----------------------------------------------------------

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    // set for a context's scratch slots: while *frozen (a captured graph holds
    // the slot's address) ensure() fails instead of reallocating
    const bool *frozen = nullptr;
    int ensure(size_t n);
    void release();
    template <typename T> T *as() const { return reinterpret_cast<T *>(p); }
};

struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t n);
    void release();
};

// Twiddle table key: (n, type, sign)
using TwKey = std::tuple<int, int, int>;

// Capacities of batched W-towers calls of one shape (csrc/wtowers.hip).
struct WtPlan {
    bool valid = false;
    int N = 0, n = 0, M = 0, S = 0, R = 0, P = 0, H = 0, type = 0, nchan = 0;
    uint32_t K = 0;
    int floor_batch = 0;
    double theta = 0, winc = 0, w_min = 0, w_step = 0;
    int64_t nvis = 0;   // longest stream the plan was made for
    int nwp_cap = 0;    // floors per tower slot
    int ntasks_cap = 0, nchunks_cap = 0, ntowers_cap = 0;
    int nitems_cap = 0; // band work items
};

} // namespace rg

// Context: scratch memory, twiddle tables and the last call's counters.
struct rg_context {
    int device = 0;
    // counters[0..7] written by kernels (int64): see Counter enum in gridding code
    rg::DevBuf counters;
    rg::PinnedBuf counters_host;
    hipStream_t last_stream = nullptr;
    int last_kind = 0; // 0 none, 1 wproj, 2 wtowers
    int64_t last_nvis = 0;
    // geometry of the W-stack held in scratch[17] (0 = none)
    int ws_g = 0, ws_og = 0, ws_planes = 0;
    // Scratch arrays, reused by every call on the context (calls on one
    // context run one after another on their stream):
    //   0-6, 15  visibility sort (binning.hpp; W-stack sort: 0-5, 7, 8, 15)
    //   7-12     W-towers batch buffers (12: the padded kernel table); 9-10
    //            also the W-stack build's row flags / row buffer, 11 the FFT
    //            transposes
    //   13       channel offsets, 14 DFT, 16 uniform-weighting counts
    //   17       the W-stack, kept between calls
    //   8        W-towers floor range; 18 W-towers task tables, 19 band masks
    rg::DevBuf scratch[20];
    // channel offsets resident in scratch[13] (uploaded only when they change,
    // so a repeated multi-channel call issues no copy and can be captured)
    std::vector<int64_t> chan_off;
    // set by a call on a capturing stream: the graph holds scratch addresses,
    // so no slot may be reallocated and the offsets not rewritten until
    // rg_context_thaw
    bool frozen = false;
    // batched W-towers capacities of the last shape
    rg::WtPlan wt_plan;
    std::map<rg::TwKey, rg::DevBuf> twiddles;
    // kernel timers (RG_TIMER_*): recorded (start, stop) event pairs
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> timer_events[RG_TIMER__COUNT];
    std::vector<hipEvent_t> event_pool;
};

namespace rg {

enum Counter {
    CNT_DROPPED = 0,
    CNT_BAD_PLANE = 1,
    CNT_ITEMS = 2,
    CNT_TASKS = 3,
    CNT_TOWERS = 4,
    CNT_RECORDS = 5, // records written by the sort (one per key a visibility reaches)
    CNT_MAX_ITEMS = 6,
    CNT_OVERFLOW = 7,    // capacity overruns inside a call (results incomplete): must stay 0
    CNT_BANDS = 8,       // W-towers band windows with records (bands written)
    CNT_CHUNKS = 9,      // W-towers column-pass chunks
    CNT_SKY_SKIPPED = 10, // W-stack predict: sky points outside the image or tables
    CNT__N = 11
};

int ensure_counters(rg_context *ctx);

// Records a (start, stop) event pair around the launches issued while it is
// alive, when the context's timing is enabled.
struct KernelTimer {
    rg_context *ctx;
    int which;
    hipStream_t s;
    hipEvent_t start = nullptr;
    KernelTimer(rg_context *c, int w, hipStream_t st);
    ~KernelTimer();
};
// twiddle table exp(sign * 2 pi i k / n), k < n, of `type`
int get_twiddles(rg_context *ctx, int n, int type, int sign, const void **out);
// the same twiddles in the stage layout of fft_band_rows (fft_core.hpp)
int get_stage_twiddles(rg_context *ctx, int n, int type, int sign, const void **out);

// internal entry points shared between translation units
int wproj_plain(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
                const rg_wkern *wk, void *grid, hipStream_t s);
int wproj_batched(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
                  const int64_t *d_off, int n_chan, const rg_wkern *wk, void *grid,
                  int grid_type, hipStream_t s);
// batched W-projection into the library's static scratch (no allocation)
int wproj_noalloc(const rg_grid_params *p, const rg_flat_vis *vis, const rg_wkern *wk, void *grid,
                  int grid_type, int accumulate, hipStream_t s);
size_t noalloc_scratch_bytes();
// d_off: device channel offsets (nchan + 1, NULL for one channel); capturing:
// the stream is being captured (no read-backs, no allocation)
int wtowers_run(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
                const int64_t *d_off, int nchan, const rg_wkern *wk, void *grid, int grid_type,
                hipStream_t s, bool capturing);
// device copy of host channel offsets, kept in the context (scratch[13])
int channel_offsets(rg_context *ctx, const int64_t *offsets, int n_chan, hipStream_t s,
                    const int64_t **d_off);
// in-place FFT of nrows contiguous rows of length n (power of two 8..4096)
int fft_rows(rg_context *ctx, void *data, int type, int n, int64_t nrows, int sign, hipStream_t s);
// row_flags (optional): bit r set for every row r (of the n * batch rows) that
// may be nonzero on input; the first row pass skips the others, which must be 0.
// dev_batch (optional, device int): only the first *dev_batch <= batch arrays
// are transformed (the launch is sized for batch).
int fft2d(rg_context *ctx, void *data, int type, int n, int batch, int sign, hipStream_t s,
          const uint32_t *row_flags = nullptr, const int32_t *dev_batch = nullptr);
int fft_shift(void *data, int type, int n, int batch, hipStream_t s);
int zero_fill(void *p, size_t bytes, hipStream_t s);
int weight_vis(rg_context *ctx, const double *u, const double *v, int64_t n, void *vis,
               int vis_type, int N, double theta, uint32_t *counts, hipStream_t s);
int make_hermitian(void *grid, int type, int n, hipStream_t s);
int gcf_correct(void *img, int type, int n, const double *taper, hipStream_t s);
int wstack_geometry(double theta, double lambda, double dw, int aa_w, double x0, int &g, int &og,
                    int &w_planes);
int epic_grid(const void *f, const int32_t *x, const int32_t *y, int n_ant, int nbatch,
              const void *illum, int supp, int g, void *grids, hipStream_t s);
int wstack_build(rg_context *ctx, const rg_wstack_params *p, const double *l, const double *m,
                 const double *flux, int32_t npts, const rg_sep_kern *corr_lm,
                 const rg_sep_kern *corr_n, hipStream_t s);
int wstack_degrid(rg_context *ctx, const rg_wstack_params *p, const double *u, const double *v,
                  const double *w, int64_t nvis, const rg_sep_kern *conv_uv,
                  const rg_sep_kern *conv_w, void *out, hipStream_t s);

} // namespace rg
