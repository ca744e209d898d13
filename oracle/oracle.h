/*
 * oracle.h -- CPU FP64 restatement of the Radio-Gridders gridding hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and by nothing else.  The product library (radio-gridders_amd) never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the Radio-Gridders source tree).  Deliberate deviations from the reference
 * (bounds checks, race-free parallel loops, the flat binning rule) are stated
 * at each function.
 *
 * Conventions shared with the product:
 *   - complex values are interleaved (re, im) FP64 pairs;
 *   - grids are row-major [v][u], N x N, zero frequency at (N/2, N/2);
 *   - W-kernel tables are [P][R][R][S][S]: plane, y-oversample, x-oversample,
 *     y-tap, x-tap (radio.cuh:175, grid_cuda.cu:118);
 *   - u, v, w are in wavelengths (the flat SoA of hdf5.cuh:31-37).
 */
#ifndef RG_ORACLE_H
#define RG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } orc_c128;

/* frac_coord (src/reference/grid.c:37-57 == src/common/radio.cuh:156-176):
 * footprint top-left (gx, gy) in grid cells and oversample indices (xf, yf). */
void orc_frac_coord(int grid_size, int kernel_size, int oversample, double theta,
                    double u, double v, double d_u, double d_v,
                    int *gx, int *gy, int *xf, int *yf);

/* The reference's packed form: grid_offset = gy*N + gx, sub_offset = S*S*(yf*R + xf). */
void orc_frac_coord_packed(int grid_size, int kernel_size, int oversample, double theta,
                           double u, double v, double d_u, double d_v,
                           int *grid_offset, int *sub_offset);

/* W-plane index floor((w - w_min)/w_step + .5)  (grid.c:168-169, grid_cuda.cu:84). */
int orc_w_plane(double w, double w_min, double w_step);

/* Direct scatter W-projection over a flat SoA stream (grid.c:157-202 over the
 * flat layout of grid_cuda.cu:81-124).  Adds into grid (not zeroed here).
 * Taps outside [0,N)^2 are dropped (the GPU reference's scatter_grid_add
 * bounds check, grid_cuda.cu:29-31).  Visibilities whose W-plane index falls
 * outside [0,P) are skipped and counted in *n_bad.  Returns the flop count
 * 8*S*S per gridded visibility (grid.c:184). */
uint64_t orc_grid_wprojection_flat(orc_c128 *grid, int grid_size, double theta,
                                   const double *u, const double *v, const double *w,
                                   const orc_c128 *vis, int64_t nvis,
                                   const orc_c128 *kern, int planes, int oversample,
                                   int support, double w_min, double w_step,
                                   int64_t *n_bad);

/* Same result with the grid rows partitioned over nthreads OpenMP threads
 * (race-free: each thread writes only the rows it owns).  The reference has
 * no parallel W-projection; this is the OpenMP CPU baseline variant. */
uint64_t orc_grid_wprojection_flat_omp(orc_c128 *grid, int grid_size, double theta,
                                       const double *u, const double *v, const double *w,
                                       const orc_c128 *vis, int64_t nvis,
                                       const orc_c128 *kern, int planes, int oversample,
                                       int support, double w_min, double w_step,
                                       int nthreads, int64_t *n_bad);

/* W-towers chunk of a visibility: (floor(u*theta+.5) + N/2) / chunk_size with
 * the double division truncated to int (radio.cuh:361-362). */
int orc_tower_index(double u, double theta, int grid_size, int chunk_size);

/* W-towers floor of a visibility: floor(w/winc + .5) (grid.c:453-455). */
int orc_floor_index(double w, double winc);

/* W-towers over a flat SoA stream (grid.c:295-555, flat binning of
 * radio.cuh:340-419).  Adds the folded towers into grid (not zeroed here).
 * nthreads > 1 runs towers in parallel with private subgrids and a
 * row-owner fold (the reference's OMP_TOWERS mode without its races).
 * Returns the flop count; *n_bad counts visibilities dropped because their
 * tower lies outside the chunk_count^2 set or their residual W-plane index
 * falls outside the table. */
uint64_t orc_grid_wtowers_flat(orc_c128 *grid, int grid_size, double theta,
                               const double *u, const double *v, const double *w,
                               const orc_c128 *vis, int64_t nvis,
                               const orc_c128 *kern, int planes, int oversample,
                               int support, double w_min, double w_step,
                               int subgrid_size, int subgrid_margin, double winc,
                               int nthreads, int64_t *n_bad);

/* Fresnel transfer pattern exp(2 pi i winc (1 - sqrt(1-l^2-m^2))), fft-shifted
 * so the zero image position is at (0,0) (grid.c:304-319). */
void orc_wtransfer(orc_c128 *out, int subgrid_size, double theta, double winc);

/* Integer power by squaring, 1/cipow(b,-e) for e < 0 (grid.c:280-293). */
orc_c128 orc_cipow(orc_c128 base, int e);

/* Unnormalised in-place 2-D FFT, sign +1 = inverse (FFTW +1 / cuFFT INVERSE),
 * sign -1 = forward.  n must be a power of two.  Radix-2, twiddles from
 * cos/sin of each index. */
void orc_fft2d(orc_c128 *a, int n, int sign);
void orc_fft1d(orc_c128 *a, int n, int sign);

/* Quadrant swap for even n (grid.c:665-680). */
void orc_fft_shift(orc_c128 *a, int n);

/* Hermitian completion (grid.c:639-662). */
void orc_make_hermitian(orc_c128 *a, int n);

/* Uniform weighting on a flat stream: counts per coarse cell
 * floor(theta*u + .5) then divides each visibility by its cell count
 * (grid.c:59-81, radio.cuh:272-288).  Uses a caller-provided count grid. */
void orc_weight_flat(unsigned int *wgrid, int grid_size, double theta,
                     const double *u, const double *v, orc_c128 *vis, int64_t nvis);

/* shift -> inverse FFT -> shift of an N x N grid in place (grid_cuda.cu:907-918). */
void orc_grid_to_image(orc_c128 *grid, int grid_size);

/* Direct DFT of the flat stream at pixels (px[i], py[i]) (column, row)
 * (dft_cuda.cu:15-39, 98-112):
 *   l = (row - N/2)/lambda, m = (col - N/2)/lambda,
 *   out = sum V * exp(-2 pi i (l u + m v + (1 - sqrt(1-l^2-m^2)) w)). */
void orc_dft_flat(orc_c128 *out, const int *px, const int *py, int64_t npix,
                  int grid_size, double lambda,
                  const double *u, const double *v, const double *w,
                  const orc_c128 *vis, int64_t nvis, int nthreads);

/* W-stacking predict (src/wstack/predict.cpp, wstack_common.h). */
int orc_wstack_geometry(double theta, double lam, double dw, int aa_w, double x0, int *og,
                        int *w_planes);
int orc_wstack_build(orc_c128 *stack, double theta, double lam, const double *pl, const double *pm,
                     const double *flux, int npts, double du, double dw, int aa_w, double x0,
                     const double *corr_lm, int corr_lm_n, const double *corr_n, int corr_n_n);
int64_t orc_wstack_degrid(orc_c128 *out, const double *u, const double *v, const double *w,
                          int64_t nvis, const orc_c128 *stack, int og, int w_planes, int g,
                          double du, double dw, int aa_uv, int aa_w, const double *conv_uv,
                          int ov_uv, const double *conv_w, int ov_w, int nthreads);
void orc_predict_quantized(orc_c128 *out, const double *pl, const double *pm, const double *flux,
                           int npts, double theta, double lam, const double *u, const double *v,
                           const double *w, int64_t nvis, int nthreads);

/* EPIC E-field gridding (src/epic/epic_grid.cu:54-179, 337-359). */
void orc_epic_grid(orc_c128 *grids, const orc_c128 *f, const int *x, const int *y, int n_ant,
                   int nbatch, const orc_c128 *illum, int supp, int g);

#ifdef __cplusplus
}
#endif

#endif
