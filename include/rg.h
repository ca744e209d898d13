/*
 * rg.h -- C ABI of the MI355X gridding engine (librg.so).
 *
 * Drop-in boundary for the Radio-Gridders hot path: Romein W-projection,
 * W-towers and the direct DFT.  Plain C types only (pointers, sizes,
 * doubles); every function returns an int status (RG_OK or a negative
 * RG_ERR_* code) and never exits the process.  rg_last_error() returns the
 * message of the most recent failure on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the Radio-Gridders tree):
 *   src/grid/grid_common.cuh:45-64   wtowers_CUDA, wtowers_CUDA_flat,
 *                                    wproj_CUDA, wproj_CUDA_flat
 *   src/dft/dft_common.cuh:12-18     image_dft_host, image_dft_host_flat
 *   src/common/hdf5.cuh:31-37, 88-96 struct flat_vis_data, struct w_kernel_data
 *   src/grid/grid.cpp, src/dft/dft.cpp  the grid.out / dft.out CLIs
 *
 * Data conventions (identical to the reference):
 *   - u, v, w in wavelengths (FP64) -- the flat SoA of hdf5.cuh:31-37;
 *   - complex values interleaved (re, im);
 *   - W-kernel table [P][R][R][S][S] (plane, y-oversample, x-oversample,
 *     y-tap, x-tap), W-plane index floor((w - w_min)/w_step + .5);
 *   - grids row-major [v][u], N x N, zero frequency at (N/2, N/2).
 * Array-level entry points take DEVICE pointers and a hipStream_t passed as
 * void* (NULL = default stream) and are asynchronous unless stated.
 */
#ifndef RG_H
#define RG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RG_ABI_VERSION 2

/* status codes */
#define RG_OK 0
#define RG_ERR_PARAM (-1)       /* invalid argument or configuration */
#define RG_ERR_IO (-2)          /* file missing, unreadable or malformed */
#define RG_ERR_SHAPE (-3)       /* inconsistent dataset shapes */
#define RG_ERR_HIP (-4)         /* HIP runtime error */
#define RG_ERR_RANGE (-5)       /* data outside the W-kernel table's w range */
#define RG_ERR_NOMEM (-6)       /* allocation failed */
#define RG_ERR_UNSUPPORTED (-7) /* configuration not implemented */

/* element types of visibilities, kernels and grids */
#define RG_C128 0 /* complex double (re, im) */
#define RG_C64 1  /* complex float  (re, im) */

/* gridder modes */
#define RG_MODE_PLAIN 0   /* the reference's launch structure, FP64 (grid_cuda.cu:826-921, 522-746) */
#define RG_MODE_BATCHED 1 /* tiled Romein work queue + batched subgrid FFTs */

/* DFT modes */
#define RG_DFT_FP64 0  /* FP64 phase and sincos (dft_cuda.cu:15-39) */
#define RG_DFT_MIXED 1 /* FP64 phase reduced mod 1, FP32 sincos, FP64 sums */

/* Flat SoA visibility stream (struct flat_vis_data, hdf5.cuh:31-37). */
typedef struct rg_flat_vis {
    const double *u;
    const double *v;
    const double *w;
    const void *vis;  /* RG_C128 or RG_C64 */
    int64_t n;
    int32_t vis_type;
    int32_t reserved;
} rg_flat_vis;

/* W-kernel table (struct w_kernel_data, hdf5.cuh:88-96). */
typedef struct rg_wkern {
    const void *data; /* [P][R][R][S][S], RG_C128 or RG_C64 */
    int32_t type;
    int32_t plane_count;  /* P */
    int32_t oversampling; /* R */
    int32_t size;         /* S (size_x == size_y) */
    double w_min;
    double w_step;
} rg_wkern;

/* Gridding parameters.  subgrid_size == 0 selects W-projection, otherwise
 * W-towers with chunk = subgrid_size - subgrid_margin. */
typedef struct rg_grid_params {
    int32_t grid_size;      /* N */
    int32_t mode;           /* RG_MODE_* */
    double theta;           /* field of view (radians) */
    int32_t subgrid_size;   /* W-towers n (0 = W-projection) */
    int32_t subgrid_margin; /* W-towers margin M */
    double winc;            /* W-towers floor spacing (wavelengths) */
    int32_t tile;           /* batched W-projection: tile edge in cells (32 or 48, 0 = 32);
                               batched W-towers: band height in rows (4, 8 or 16 (16 in
                               FP32 only), 0 = 8) */
    int32_t vis_per_item;   /* batched visibilities per work item (0 = default) */
    int32_t compensated;    /* batched FP32: 0 plain FP32 sums (default), 1 two-sum
                               compensated sums: W-projection tile sums (worth it only into
                               an RG_C128 grid), W-towers band lanes' register sums */
    int32_t floor_batch;    /* batched W-towers: (tower, floor) tasks per batch (0 = as many as
                               fit 16 GiB of subgrids) */
    int32_t fft_radix;      /* batched W-towers: radix of the column pass's FFT stages (8 or
                               4, 0 = 8; the last stage takes what is left of log2 n) */
} rg_grid_params;

/* Counters of the last rg_grid call (read back synchronously). */
typedef struct rg_grid_stats {
    int64_t n_vis;       /* visibilities in the stream */
    int64_t n_gridded;   /* visibilities gridded */
    int64_t n_dropped;   /* outside the grid / tower set (dropped, as the reference's bounds check) */
    int64_t n_bad_plane; /* W-plane index outside the table (skipped) */
    int64_t n_items;     /* batched work items (W-projection tiles, W-towers bands) */
    int64_t n_tasks;     /* non-empty (tower, floor) tasks (W-towers) */
    int64_t n_towers;    /* non-empty towers (W-towers) */
    int64_t n_records;   /* records of the batched work queue (W-projection: sorted
                            visibilities; W-towers: band records, one per band window a
                            visibility falls in) */
    int64_t n_bands;     /* batched W-towers: subgrid bands written (windows with records) */
    int64_t n_chunks;    /* batched W-towers: column-pass chunks (runs of <= 32 floors of
                            one tower), each writing one partial subimage */
} rg_grid_stats;

typedef struct rg_context rg_context;

/* ---- library / context ---- */
const char *rg_last_error(void);
int rg_abi_version(void);
/* Creates a context on HIP device `device`.  A context owns reusable scratch
 * memory and twiddle tables; one context per host thread. */
int rg_context_create(rg_context **ctx, int device);
int rg_context_destroy(rg_context *ctx);
/* Stats of the last rg_grid on this context (synchronises the stream of that call).
 * Returns RG_ERR_RANGE (with `out` filled) when that call overran one of its
 * internal capacity bounds, i.e. its result is incomplete. */
int rg_context_stats(rg_context *ctx, rg_grid_stats *out);
/* Graph capture: a gridding call made on a capturing stream freezes the
 * context, since the captured graph holds the addresses of its scratch
 * memory and channel offsets.  While frozen, a call that would need more
 * scratch memory or other channel offsets fails with RG_ERR_PARAM instead of
 * reallocating or rewriting what the graph reads.  rg_context_thaw declares
 * that no graph captured from this context will be replayed again and lifts
 * the restriction (no counterpart in the reference, which has no graphs). */
int rg_context_thaw(rg_context *ctx);

/* Kernel timers: when enabled, librg records HIP events on the launch stream
 * around each launch of the timed kernels below.  rg_context_timer
 * synchronises, returns the summed duration (ms) and launch count of one
 * timer since it was last read, and resets it. */
#define RG_TIMER_WPROJ_TILES 0  /* batched W-projection tile kernel */
#define RG_TIMER_WPROJ_PLAIN 1  /* plain W-projection scatter */
#define RG_TIMER_BINNING 2      /* count + scan + scatter of the sort */
#define RG_TIMER_TOWER_TILES 3  /* batched W-towers band pass (k_tower_band_items): scatter + row IFFT */
#define RG_TIMER_TOWER_REDUCE 4 /* batched W-towers sum of partial images into subimages */
#define RG_TIMER_TOWER_COLS 5   /* batched W-towers column IFFT + screen pass */
#define RG_TIMER_TOWER_FINISH 6 /* W-towers forward FFTs + fold */
#define RG_TIMER_TOWER_COMBINE 7 /* batched W-towers sum of bands split over work items */
#define RG_TIMER__COUNT 8
int rg_context_set_timing(rg_context *ctx, int32_t enable);
int rg_context_timer(rg_context *ctx, int32_t which, double *total_ms, int64_t *launches);

/* ---- gridding (array level) ----
 * Zeroes `grid` (N x N of grid_type) and grids every visibility of `vis`
 * with the W-kernel table `wk`: Romein W-projection (scatter_grid_kernel_flat,
 * grid_cuda.cu:253-280) or W-towers (wtowers_CUDA_flat, grid_cuda.cu:522-746,
 * including the fold add_subs2main_kernel, common_kernels.cu:43-79).  The
 * result is the uv grid, before the final shift / FFT.
 * RG_MODE_PLAIN requires RG_C128 everywhere. */
int rg_grid(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
            const rg_wkern *wk, void *grid, int32_t grid_type, void *stream);

/* Multi-channel batched gridding: channel c owns visibilities
 * [offsets[c], offsets[c+1]) of `vis` and the grid at
 * grid + c * N * N elements.  All channels are gridded by one work queue:
 * W-projection tiles, or W-towers (subgrid_size > 0) tasks of every
 * (channel, tower, floor) with one RG_C128 master grid per channel.  offsets
 * is a HOST array of n_chan + 1 entries.
 *
 * Batched W-towers calls (one or many channels) size their key space and
 * buffers by a plan kept in the context: the first call of a shape reads the
 * floor range and task counts back; later calls of the same shape (grid,
 * subgrid, kernels, parameters, channel count, no longer stream) issue no
 * host read-back before their end, where an eager call checks that its data
 * fitted the plan (and re-plans and redoes itself otherwise).  Such a call
 * can be captured in a HIP graph after an eager call of its shape; a replay
 * whose data outgrew the plan is reported by rg_context_stats
 * (RG_ERR_RANGE). */
int rg_grid_channels(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
                     const int64_t *offsets, int32_t n_chan, const rg_wkern *wk,
                     void *grids, int32_t grid_type, void *stream);

/* Allocation-free batched W-projection (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc.  Same parameters and result
 * as rg_grid with RG_MODE_BATCHED and subgrid_size 0 (FP32 or FP64 kernels),
 * on the current HIP device, with no context: the sort's working memory is a
 * 64 MiB device array inside librg (one per device, held until the process
 * ends), enough for about 2.0 M visibilities (FP32 kernels) or 1.4 M (FP64)
 * per call on a 2048^2 grid; a call that needs more returns RG_ERR_PARAM.
 * accumulate = 0 zeroes the grid first (as rg_grid), 1 adds into it, so a
 * longer stream can be split over calls.  Calls on one device share that
 * array: calls that overlap in time on one device (other streams or host
 * threads) race.  No call counters are reported (rg_context_stats does not
 * see these calls); visibilities off the grid or outside the table's w range
 * are skipped as in rg_grid. */
int rg_grid_noalloc(const rg_grid_params *p, const rg_flat_vis *vis, const rg_wkern *wk,
                    void *grid, int32_t grid_type, int32_t accumulate, void *stream);

/* shift -> unnormalised inverse FFT -> shift, in place (grid_cuda.cu:907-918,
 * fft_shift_kernel common_kernels.cuh:17-32). */
int rg_grid_to_image(rg_context *ctx, void *grid, int32_t type, int32_t n, void *stream);

/* In-place unnormalised 2-D FFT of `batch` n x n arrays (n a power of two,
 * 4 <= n <= 4096); sign +1 = inverse (cuFFT INVERSE), -1 = forward. */
int rg_fft2d(rg_context *ctx, void *data, int32_t type, int32_t n, int32_t batch,
             int32_t sign, void *stream);

/* In-place quadrant swap of an even-sized n x n array (fft_shift_kernel). */
int rg_fft_shift(void *data, int32_t type, int32_t n, void *stream);

/* ---- image conventions of the CPU reference pipeline (main.c:151-154, 258-281) ----
 * Uniform weighting (weight, src/reference/grid.c:59-81): every visibility is
 * divided by the number of visibilities in its cell (y + N/2) N + (x + N/2),
 * x = floor(theta u + 1/2), y = floor(theta v + 1/2) (coord, grid.c:25-35).
 * vis (device, RG_C128 or RG_C64) is modified in place; the quotient is taken
 * in FP64.  A visibility whose cell lies outside the grid (the reference
 * indexes out of bounds) is neither counted nor changed.  counts: optional
 * device array of N * N uint32 that receives the per-cell counts (NULL: the
 * context's scratch memory). */
int rg_weight(rg_context *ctx, const double *u, const double *v, int64_t n, void *vis,
              int32_t vis_type, int32_t grid_size, double theta, uint32_t *counts,
              void *stream);

/* Hermitian completion of an n x n uv grid in place (make_hermitian,
 * src/reference/grid.c:639-662): cells base + k and n^2 - 1 - k (base = n + 1
 * for even n, 0 for odd n) each add the other's conjugate, for every k up to
 * the zero frequency (n + 1)(n / 2), which adds its own conjugate. */
int rg_make_hermitian(void *grid, int32_t type, int32_t n, void *stream);

/* Grid correction: image[r][c] /= taper[r] * taper[c] (FP64 quotient), taper a
 * device array of n doubles, e.g. the kernel's image-plane taper A at the
 * pixel positions (no counterpart in the reference, whose images keep the
 * taper). */
int rg_gcf_correct(void *image, int32_t type, int32_t n, const double *taper, void *stream);

/* ---- direct DFT (array level) ----
 * out[i] (RG_C128) = sum_v V exp(-2 pi i (l u + m v + (1 - sqrt(1-l^2-m^2)) w))
 * with l = (py[i] - N/2)/lambda, m = (px[i] - N/2)/lambda
 * (calculate_dft_sum_flat / image_dft_flat, dft_cuda.cu:15-39, 98-112).
 * px, py are device int32 arrays (column, row). */
int rg_dft(rg_context *ctx, const rg_flat_vis *vis, const int32_t *px, const int32_t *py,
           int64_t npix, int32_t grid_size, double lambda, void *out, int32_t mode,
           void *stream);

/* Full N x N DFT image in the image_dft_flat layout (index = row * N + col).
 * blocks * threads_block must equal N * N (README.md:84-86 contract). */
int rg_dft_image(rg_context *ctx, const rg_flat_vis *vis, int32_t grid_size, double lambda,
                 void *out, int32_t blocks, int32_t threads_block, int32_t mode, void *stream);

/* ---- W-stacking predict (sky -> visibilities), separable kernels ----
 * wstack_predict_cu_3D (src/wstack/predict.cuh:33-47; CPU semantics
 * wstack_predict, src/wstack/predict.cpp:396-560).  Kernel tables are device
 * arrays in the layout of the reference's loaders: convolution kernels
 * data[oversampling][size] (row ov holds the size taps), correction
 * functions size * oversampling samples. */
typedef struct rg_sep_kern {
    const double *data;   /* device */
    int32_t size;
    int32_t oversampling;
} rg_sep_kern;

typedef struct rg_wstack_params {
    double theta;          /* field of view (radians) */
    double lambda;         /* uv extent (wavelengths); g = floor(theta * lambda) */
    double du, dw;         /* uv cell and w-plane spacing of the kernels */
    double x0;             /* image fraction of the kernels: og = round(0.5 / x0) * g */
    int32_t aa_support_uv; /* uv kernel support (its size) */
    int32_t aa_support_w;  /* w kernel support (its size) */
} rg_wstack_params;

/* Builds the W-stack of a point-source sky in the context: w_planes planes
 * of og x og RG_C128, plane k the fft-shifted forward FFT of the sky (flux
 * (NULL: 1) / correction at each source's quantised pixel, generate_sky
 * predict.cpp:219-278) times F^(k - w_planes/2), F the w-screen of spacing
 * dw (generate_fresnel, predict.cpp:182-216).  l, m, flux are device arrays
 * of npts.  Sources whose pixel or correction index falls outside its table
 * are skipped (the reference indexes out of bounds).  og_out, planes_out,
 * skipped_out: optional host outputs (skipped_out synchronises the stream). */
int rg_wstack_build(rg_context *ctx, const rg_wstack_params *p, const double *l, const double *m,
                    const double *flux, int32_t npts, const rg_sep_kern *corr_lm,
                    const rg_sep_kern *corr_n, int32_t *og_out, int32_t *planes_out,
                    int64_t *skipped_out, void *stream);

/* Predicts nvis visibilities (device u, v, w in wavelengths; out RG_C128)
 * from the context's W-stack with the separable uv / w kernels
 * (deconvolve_visibility_, src/wstack/wstack_common.h:420-470).  A
 * visibility whose taps leave the stack is predicted as 0 and counted in
 * rg_context_stats' n_dropped. */
int rg_wstack_degrid(rg_context *ctx, const rg_wstack_params *p, const double *u, const double *v,
                     const double *w, int64_t nvis, const rg_sep_kern *conv_uv,
                     const rg_sep_kern *conv_w, void *out, void *stream);

/* Copies the context's W-stack ([planes][og][og] RG_C128, row v, column u)
 * into dst (host or device memory of capacity_bytes) and waits for the copy.
 * No counterpart in the reference (its stack is internal); for inspection
 * and tests. */
int rg_wstack_copy(rg_context *ctx, void *dst, int64_t capacity_bytes, void *stream);

/* rg_wstack_build then rg_wstack_degrid (wstack_predict_cu_3D). */
int rg_wstack_predict(rg_context *ctx, const rg_wstack_params *p, const double *l, const double *m,
                      const double *flux, int32_t npts, const double *u, const double *v,
                      const double *w, int64_t nvis, const rg_sep_kern *conv_uv,
                      const rg_sep_kern *conv_w, const rg_sep_kern *corr_lm,
                      const rg_sep_kern *corr_n, void *out, void *stream);

/* ---- EPIC E-field gridding ----
 * epic_romein / scatter_grid_kernel (src/epic/epic_grid.cu:54-218): batch b
 * holds n_ant antenna E-fields f[b][a] (RG_C64) at integer grid positions
 * (x, y)[b][a]; grids[b] (g x g RG_C64, row y, column x) =
 * sum_a conj(illum[ty][tx]) f[b][a] at (y + ty, x + tx) over the
 * support x support illumination (RG_C64); taps outside the grid are dropped.
 * z is accepted for the reference's signature and not read (nor is it by the
 * reference's kernel).  All arrays are device arrays. */
int rg_epic_grid(rg_context *ctx, const void *fdata, const int32_t *x, const int32_t *y,
                 const int32_t *z, int32_t n_ant, int32_t nbatch, const void *illum,
                 int32_t support, int32_t grid_size, void *grids, void *stream);

/* ---- input generation ---- */
/* Point-source visibilities V = sum_s flux_s exp(-2 pi i (u l_s + v m_s + w (n_s - 1)))
 * (predict.cpp:27-45 convention) evaluated in FP64 on the device.
 * u, v, w, l, m, flux are device arrays; vis_out is RG_C128 or RG_C64. */
int rg_predict_points(const double *u, const double *v, const double *w, int64_t n,
                      const double *l, const double *m, const double *flux, int32_t n_src,
                      void *vis_out, int32_t vis_type, void *stream);

/* Normalised zeroth prolate spheroidal wave function psi_0(c, x) / psi_0(c, 0),
 * |x| <= 1, from its Legendre expansion (host). */
double rg_pswf(double c, double x);
/* rg_pswf at n points (the expansion is computed once). */
int rg_pswf_eval(double c, const double *x, int64_t n, double *out);

/* W-kernel table generator (host, FP64): plane p holds the oversampled
 * kernel of the image-plane function A(l) A(m) exp(-2 pi i w (n - 1)),
 * w = w_min + p * w_step, A = rg_pswf(pswf_c, 2 l / theta), sampled at the
 * image_samples^2 cell centres of the field and transformed to the taps
 * d = x - S/2 - xf/R:  K(d) = M^-2 sum A A exp(-2 pi i w (n-1)) exp(2 pi i d l / theta).
 * out_host is P*R*R*S*S elements of `type`.
 * pswf_c <= 0 selects pi * S / 4.  nthreads <= 0 uses all cores. */
int rg_wkern_generate(void *out_host, int32_t type, int32_t planes, int32_t oversampling,
                      int32_t size, double theta, double w_min, double w_step,
                      int32_t image_samples, double pswf_c, int32_t nthreads);

/* ---- file level (reference wrappers) ----
 * Visibility files and W-kernel files use the reference's HDF5 layouts
 * (load_vis hdf5.c:64-369, load_wkern_CUDA hdf5.c:1362-1520) when libhdf5
 * can be loaded at run time.  `grid` is a caller-allocated device array of
 * grid_size^2 complex128; on return it holds the fft-shifted, inverse-FFT'd
 * image like the reference.  `gridh` is accepted and unused (as in the
 * reference).  Non-flat variants behave like the flat ones. */
int rg_wtowers_flat(const char *visfile, const char *wkernfile, void *grid, void *gridh,
                    int grid_size, double theta, double lambda, double bl_min, double bl_max,
                    int subgrid_size, int subgrid_margin, double winc);
int rg_wtowers(const char *visfile, const char *wkernfile, void *grid, void *gridh,
               int grid_size, double theta, double lambda, double bl_min, double bl_max,
               int subgrid_size, int subgrid_margin, double winc);
int rg_wproj_flat(const char *visfile, const char *wkernfile, void *grid, void *gridh,
                  int grid_size, double theta, double lambda, double bl_min, double bl_max,
                  int threads_per_block);
int rg_wproj(const char *visfile, const char *wkernfile, void *grid, void *gridh,
             int grid_size, double theta, double lambda, double bl_min, double bl_max,
             int threads_per_block);
int rg_image_dft_host_flat(const char *visfile, void *grid_host, void *grid_dev, int grid_size,
                           double theta, double lambda, double bl_min, double bl_max,
                           int blocks, int threads_block);
int rg_image_dft_host(const char *visfile, void *grid_host, void *grid_dev, int grid_size,
                      double theta, double lambda, double bl_min, double bl_max,
                      int blocks, int threads_block);

/* Selects the gridding mode the file-level wrappers use (default RG_MODE_BATCHED
 * with FP64 data, i.e. the same results as the plain path to FP64 rounding). */
int rg_set_file_mode(int32_t mode);

/* Image convention of the file-level gridding wrappers:
 * RG_IMAGE_GPU (default): the GPU reference's image (grid_cuda.cu:907-918),
 * no weighting, shift -> inverse FFT -> shift;
 * RG_IMAGE_CPU: the CPU reference's image (src/reference/main.c:151-154,
 * 258-281): uniform weighting before gridding, hermitian completion, then
 * shift -> forward FFT -> shift. */
#define RG_IMAGE_GPU 0
#define RG_IMAGE_CPU 1
int rg_set_image_convention(int32_t convention);

/* ---- HDF5 files (host; libhdf5 >= 1.10 loaded at run time) ---- */
/* 1 when libhdf5 could be loaded (RG_HDF5_LIB overrides the search). */
int rg_hdf5_available(void);
/* load_vis + flatten_visibilities_CUDA (hdf5.c:64-369, 827-872): baselines whose
 * first uvw sample has length in [bl_min, bl_max) metres, flattened
 * baseline -> time -> frequency into wavelengths.  Arrays are malloc'ed;
 * release each with rg_free_host.  vis is interleaved complex128. */
int rg_load_flat_vis(const char *visfile, double bl_min, double bl_max, int64_t *n, double **u,
                     double **v, double **w, double **vis);
/* load_wkern_CUDA (hdf5.c:1362-1520): group wkern/<theta %g>, planes indexed by w. */
int rg_load_wkern(const char *wkernfile, double theta, int32_t *planes, int32_t *oversampling,
                  int32_t *size, double *w_min, double *w_step, double **data);
void rg_free_host(void *p);

/* Separable (Sze-Tan) kernel file (load_sep_kern / load_sep_kern_T,
 * src/common/hdf5.c:372-500): dataset "sepkern/kern" (2-D FP64) read in file
 * order into a malloc'd host array (free with rg_free_host), and the file
 * attributes du, dw, x0 (NaN when absent).  transposed = 0: dims are
 * [oversampling][size]; 1: [size][oversampling] (the _T loader). */
int rg_load_sep_kern(const char *filename, int32_t transposed, int32_t *size,
                     int32_t *oversampling, double *du, double *dw, double *x0, double **data);
/* Writers of the same layouts: vis/<a1>/<a2>/{frequency, time, uvw, vis} with
 * uvw [n_bl][T][3] metres and vis [n_bl][T][F] complex128; and
 * wkern/<theta %g>/<w>/kern [R][R][S][S] complex128 for w = w_min + p w_step. */
int rg_write_vis_hdf5(const char *filename, int32_t n_bl, const int32_t *a1, const int32_t *a2,
                      int32_t T, int32_t F, const double *time, const double *freq,
                      const double *uvw, const double *vis);
int rg_write_wkern_hdf5(const char *filename, double theta, int32_t planes, int32_t oversampling,
                        int32_t size, double w_min, double w_step, const double *data);

#ifdef __cplusplus
}
#endif

#endif
