/*
 * oracle.c -- CPU FP64 restatement of the Radio-Gridders gridding hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py; never by the product library.
 *
 * Built without fast-math and with -ffp-contract=off so the coordinate and
 * index formulas round exactly as written (see oracle/Makefile).
 */
#include "oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double TWO_PI = 2.0 * M_PI;

/* ------------------------------------------------------------------ */
/* coordinates                                                          */
/* ------------------------------------------------------------------ */

/* src/reference/grid.c:37-57; src/common/radio.cuh:156-176. */
void orc_frac_coord(int grid_size, int kernel_size, int oversample, double theta,
                    double u, double v, double d_u, double d_v,
                    int *gx, int *gy, int *xf, int *yf)
{
    double x = theta * (u - d_u);
    double y = theta * (v - d_v);
    int flx = (int)floor(x + .5 / oversample);
    int fly = (int)floor(y + .5 / oversample);
    *xf = (int)floor((x - (double)flx) * oversample + .5);
    *yf = (int)floor((y - (double)fly) * oversample + .5);
    /* x + .5/R can round up to the next integer, leaving the oversample
     * index at -1 (or R); the reference then reads outside the sub-kernel.
     * Both this restatement and the GPU path clamp to [0, R). */
    if (*xf < 0) *xf = 0;
    if (*xf >= oversample) *xf = oversample - 1;
    if (*yf < 0) *yf = 0;
    if (*yf >= oversample) *yf = oversample - 1;
    *gx = flx + grid_size / 2 - kernel_size / 2;
    *gy = fly + grid_size / 2 - kernel_size / 2;
}

void orc_frac_coord_packed(int grid_size, int kernel_size, int oversample, double theta,
                           double u, double v, double d_u, double d_v,
                           int *grid_offset, int *sub_offset)
{
    int gx, gy, xf, yf;
    orc_frac_coord(grid_size, kernel_size, oversample, theta, u, v, d_u, d_v,
                   &gx, &gy, &xf, &yf);
    *grid_offset = gy * grid_size + gx;
    *sub_offset = kernel_size * kernel_size * (yf * oversample + xf);
}

/* grid.c:168-169. */
int orc_w_plane(double w, double w_min, double w_step)
{
    return (int)floor((w - w_min) / w_step + .5);
}

/* ------------------------------------------------------------------ */
/* W-projection                                                         */
/* ------------------------------------------------------------------ */

/* One visibility, grid.c:157-185 (w_project), restricted to rows
 * [row0, row1) and with the scatter_grid_add bounds check
 * (grid_cuda.cu:29-31).  Returns 0 when the W-plane is out of range. */
static int w_project_rows(orc_c128 *grid, int N, double theta,
                          double u, double v, double w, orc_c128 vis,
                          double d_u, double d_v, double d_w,
                          const orc_c128 *kern, int P, int R, int S,
                          double w_min, double w_step, int row0, int row1)
{
    int gx, gy, xf, yf;
    orc_frac_coord(N, S, R, theta, u, v, d_u, d_v, &gx, &gy, &xf, &yf);
    int plane = orc_w_plane(w - d_w, w_min, w_step);
    if (plane < 0 || plane >= P)
        return 0;
    const orc_c128 *wk = kern + (size_t)plane * R * R * S * S
                              + (size_t)S * S * (yf * R + xf);
    for (int y = 0; y < S; y++) {
        int gyy = gy + y;
        if (gyy < row0 || gyy >= row1)
            continue;
        for (int x = 0; x < S; x++) {
            int gxx = gx + x;
            if (gxx < 0 || gxx >= N)
                continue;
            orc_c128 k = wk[y * S + x];
            /* v * conj(k) */
            orc_c128 *g = &grid[(size_t)gyy * N + gxx];
            g->re += vis.re * k.re + vis.im * k.im;
            g->im += vis.im * k.re - vis.re * k.im;
        }
    }
    return 1;
}

/* grid.c:187-202 over the flat stream. */
uint64_t orc_grid_wprojection_flat(orc_c128 *grid, int grid_size, double theta,
                                   const double *u, const double *v, const double *w,
                                   const orc_c128 *vis, int64_t nvis,
                                   const orc_c128 *kern, int planes, int oversample,
                                   int support, double w_min, double w_step,
                                   int64_t *n_bad)
{
    uint64_t flops = 0;
    int64_t bad = 0;
    for (int64_t i = 0; i < nvis; i++) {
        if (w_project_rows(grid, grid_size, theta, u[i], v[i], w[i], vis[i], 0, 0, 0,
                           kern, planes, oversample, support, w_min, w_step,
                           0, grid_size))
            flops += 8ull * support * support;
        else
            bad++;
    }
    if (n_bad)
        *n_bad = bad;
    return flops;
}

uint64_t orc_grid_wprojection_flat_omp(orc_c128 *grid, int grid_size, double theta,
                                       const double *u, const double *v, const double *w,
                                       const orc_c128 *vis, int64_t nvis,
                                       const orc_c128 *kern, int planes, int oversample,
                                       int support, double w_min, double w_step,
                                       int nthreads, int64_t *n_bad)
{
    int64_t bad = 0;
    if (nthreads < 1)
        nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(+ : bad)
    {
        int t = 0, nt = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        nt = omp_get_num_threads();
#endif
        int row0 = (int)((int64_t)grid_size * t / nt);
        int row1 = (int)((int64_t)grid_size * (t + 1) / nt);
        for (int64_t i = 0; i < nvis; i++) {
            int ok = w_project_rows(grid, grid_size, theta, u[i], v[i], w[i], vis[i],
                                    0, 0, 0, kern, planes, oversample, support,
                                    w_min, w_step, row0, row1);
            if (t == 0 && !ok)
                bad++;
        }
    }
    if (n_bad)
        *n_bad = bad;
    return (uint64_t)(nvis - bad) * 8ull * support * support;
}

/* ------------------------------------------------------------------ */
/* FFT                                                                  */
/* ------------------------------------------------------------------ */

static int ilog2(int n)
{
    int l = 0;
    while ((1 << l) < n)
        l++;
    return l;
}

/* Iterative radix-2 decimation-in-time, unnormalised,
 * out[k] = sum_j a[j] exp(sign * 2 pi i j k / n). */
static void fft1d_strided(orc_c128 *a, int n, int stride, int sign, const orc_c128 *tw)
{
    int lg = ilog2(n);
    for (int i = 0; i < n; i++) {
        int r = 0;
        for (int b = 0; b < lg; b++)
            r |= ((i >> b) & 1) << (lg - 1 - b);
        if (r > i) {
            orc_c128 t = a[(size_t)i * stride];
            a[(size_t)i * stride] = a[(size_t)r * stride];
            a[(size_t)r * stride] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int s = 0; s < n; s += len) {
            for (int j = 0; j < half; j++) {
                orc_c128 wv = tw[j * step];
                if (sign < 0)
                    wv.im = -wv.im;
                orc_c128 *p = &a[(size_t)(s + j) * stride];
                orc_c128 *q = &a[(size_t)(s + j + half) * stride];
                orc_c128 t = { q->re * wv.re - q->im * wv.im, q->re * wv.im + q->im * wv.re };
                q->re = p->re - t.re;
                q->im = p->im - t.im;
                p->re += t.re;
                p->im += t.im;
            }
        }
    }
}

/* tw[k] = exp(+2 pi i k / n) for k < n/2 */
static orc_c128 *make_twiddles(int n)
{
    orc_c128 *tw = (orc_c128 *)malloc(sizeof(orc_c128) * (size_t)(n / 2 + 1));
    for (int k = 0; k <= n / 2; k++) {
        double a = TWO_PI * (double)k / (double)n;
        tw[k].re = cos(a);
        tw[k].im = sin(a);
    }
    return tw;
}

void orc_fft1d(orc_c128 *a, int n, int sign)
{
    orc_c128 *tw = make_twiddles(n);
    fft1d_strided(a, n, 1, sign, tw);
    free(tw);
}

void orc_fft2d(orc_c128 *a, int n, int sign)
{
    orc_c128 *tw = make_twiddles(n);
    for (int r = 0; r < n; r++)
        fft1d_strided(a + (size_t)r * n, n, 1, sign, tw);
    for (int c = 0; c < n; c++)
        fft1d_strided(a + c, n, n, sign, tw);
    free(tw);
}

/* grid.c:665-680 */
void orc_fft_shift(orc_c128 *a, int n)
{
    for (int y = 0; y < n; y++) {
        for (int x = 0; x < n / 2; x++) {
            size_t ix0 = (size_t)y * n + x;
            size_t ix1 = (ix0 + (size_t)(n + 1) * (n / 2)) % ((size_t)n * n);
            orc_c128 t = a[ix0];
            a[ix0] = a[ix1];
            a[ix1] = t;
        }
    }
}

/* grid.c:639-662 */
void orc_make_hermitian(orc_c128 *a, int n)
{
    orc_c128 *p0 = (n % 2 == 0) ? a + n + 1 : a;
    orc_c128 *p1 = a + (size_t)n * n - 1;
    while (p0 < p1) {
        orc_c128 g0 = *p0;
        p0->re += p1->re;
        p0->im -= p1->im;
        p0++;
        p1->re += g0.re;
        p1->im -= g0.im;
        p1--;
    }
    /* p0 == p1 == zero frequency */
    p0->re += p0->re;
    p0->im -= p0->im;
}

/* grid.c:59-81 over the flat stream (radio.cuh:272-288). */
void orc_weight_flat(unsigned int *wgrid, int grid_size, double theta,
                     const double *u, const double *v, orc_c128 *vis, int64_t nvis)
{
    memset(wgrid, 0, sizeof(unsigned int) * (size_t)grid_size * grid_size);
    for (int64_t i = 0; i < nvis; i++) {
        int x = (int)floor(theta * u[i] + .5);
        int y = (int)floor(theta * v[i] + .5);
        wgrid[(size_t)(y + grid_size / 2) * grid_size + (x + grid_size / 2)]++;
    }
    for (int64_t i = 0; i < nvis; i++) {
        int x = (int)floor(theta * u[i] + .5);
        int y = (int)floor(theta * v[i] + .5);
        unsigned int c = wgrid[(size_t)(y + grid_size / 2) * grid_size + (x + grid_size / 2)];
        vis[i].re /= c;
        vis[i].im /= c;
    }
}

/* grid_cuda.cu:907-918 */
void orc_grid_to_image(orc_c128 *grid, int grid_size)
{
    orc_fft_shift(grid, grid_size);
    orc_fft2d(grid, grid_size, +1);
    orc_fft_shift(grid, grid_size);
}

/* ------------------------------------------------------------------ */
/* W-towers                                                             */
/* ------------------------------------------------------------------ */

static orc_c128 cmul(orc_c128 a, orc_c128 b)
{
    orc_c128 r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
    return r;
}

static orc_c128 cdiv(orc_c128 a, orc_c128 b)
{
    double d = b.re * b.re + b.im * b.im;
    orc_c128 r = { (a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d };
    return r;
}

/* grid.c:280-293 */
orc_c128 orc_cipow(orc_c128 base, int e)
{
    orc_c128 result = { 1.0, 0.0 };
    if (e < 0) {
        orc_c128 one = { 1.0, 0.0 };
        return cdiv(one, orc_cipow(base, -e));
    }
    if (e == 1)
        return base;
    while (e) {
        if (e & 1)
            result = cmul(result, base);
        e >>= 1;
        base = cmul(base, base);
    }
    return result;
}

/* grid.c:304-319 */
void orc_wtransfer(orc_c128 *out, int n, double theta, double winc)
{
    for (int y = 0; y < n; y++) {
        for (int x = 0; x < n; x++) {
            double l = theta * (double)(x - n / 2) / n;
            double m = theta * (double)(y - n / 2) / n;
            double ph = winc * (1 - sqrt(1 - l * l - m * m));
            double a = TWO_PI * ph;
            out[(size_t)y * n + x].re = cos(a);
            out[(size_t)y * n + x].im = sin(a);
        }
    }
    orc_fft_shift(out, n);
}

/* radio.cuh:361-362: (floor(u*theta + 0.5) + grid_size/2) / chunk_size is a
 * double division, truncated on assignment to int. */
int orc_tower_index(double u, double theta, int grid_size, int chunk_size)
{
    double q = (floor(u * theta + 0.5) + grid_size / 2) / chunk_size;
    return (int)q;
}

/* grid.c:453-455: visibility w belongs to floor wp when
 * (wp - .5) winc <= w < (wp + .5) winc, i.e. wp = floor(w/winc + .5). */
int orc_floor_index(double w, double winc)
{
    return (int)floor(w / winc + 0.5);
}

typedef struct {
    int64_t idx;
    int tower;
    int wp;
} vis_key;

static int cmp_key(const void *a, const void *b)
{
    const vis_key *x = (const vis_key *)a, *y = (const vis_key *)b;
    if (x->tower != y->tower)
        return x->tower < y->tower ? -1 : 1;
    if (x->wp != y->wp)
        return x->wp < y->wp ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* grid.c:295-555 (grid_wtowers) over the flat stream.  Towers (chunks) and
 * floors are assigned per visibility with the flat binning rules above;
 * within each (tower, floor) visibilities keep their stream order.  Empty
 * floors are skipped and the Fresnel transfer uses cipow(F, wp - last_wp)
 * (grid.c:473-480), then the w = 0 transfer (grid.c:489-496), forward FFT
 * (grid.c:499) and fold with 1/n^2 (grid.c:503-516).  Taps outside the
 * subgrid are dropped (the reference's CPU scatter has no bounds check). */
uint64_t orc_grid_wtowers_flat(orc_c128 *grid, int N, double theta,
                               const double *u, const double *v, const double *w,
                               const orc_c128 *vis, int64_t nvis,
                               const orc_c128 *kern, int P, int R, int S,
                               double w_min, double w_step,
                               int n, int margin, double winc,
                               int nthreads, int64_t *n_bad)
{
    int chunk = n - margin;
    int cc = N / chunk + 1;
    int64_t bad = 0;
    uint64_t flops = 0;
    if (nthreads < 1)
        nthreads = 1;

    orc_c128 *wtransfer = (orc_c128 *)malloc(sizeof(orc_c128) * (size_t)n * n);
    orc_wtransfer(wtransfer, n, theta, winc);

    /* Bin: sort (tower, floor, stream index). */
    vis_key *keys = (vis_key *)malloc(sizeof(vis_key) * (size_t)(nvis > 0 ? nvis : 1));
    int64_t nk = 0;
    for (int64_t i = 0; i < nvis; i++) {
        int cx = orc_tower_index(u[i], theta, N, chunk);
        int cy = orc_tower_index(v[i], theta, N, chunk);
        if (cx < 0 || cx >= cc || cy < 0 || cy >= cc) {
            bad++;
            continue;
        }
        keys[nk].idx = i;
        keys[nk].tower = cy * cc + cx;
        keys[nk].wp = orc_floor_index(w[i], winc);
        nk++;
    }
    qsort(keys, (size_t)nk, sizeof(vis_key), cmp_key);

    /* Start of each tower's run in keys[]. */
    int64_t *tstart = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cc * cc + 1));
    {
        int64_t k = 0;
        for (int t = 0; t <= cc * cc; t++) {
            while (k < nk && keys[k].tower < t)
                k++;
            tstart[t] = k;
        }
    }

    orc_c128 *subimgs = (orc_c128 *)calloc((size_t)cc * cc * n * n, sizeof(orc_c128));
    char *have = (char *)calloc((size_t)cc * cc, 1);
    int64_t bad_plane = 0;

#pragma omp parallel num_threads(nthreads) reduction(+ : flops, bad_plane)
    {
        orc_c128 *subgrid = (orc_c128 *)malloc(sizeof(orc_c128) * (size_t)n * n);
#pragma omp for schedule(dynamic)
        for (int t = 0; t < cc * cc; t++) {
            if (tstart[t] == tstart[t + 1])
                continue;
            int cx = t % cc, cy = t / cc;
            int x_min = chunk * cx - N / 2;
            int y_min = chunk * cy - N / 2;
            double u_mid = (double)(x_min + chunk / 2) / theta;
            double v_mid = (double)(y_min + chunk / 2) / theta;
            orc_c128 *subimg = subimgs + (size_t)t * n * n;
            int last_wp = 0, first = 1;
            int64_t k = tstart[t];
            while (k < tstart[t + 1]) {
                int wp = keys[k].wp;
                double w_mid = (double)wp * winc;
                memset(subgrid, 0, sizeof(orc_c128) * (size_t)n * n);
                int64_t k1 = k;
                while (k1 < tstart[t + 1] && keys[k1].wp == wp) {
                    int64_t i = keys[k1].idx;
                    if (w_project_rows(subgrid, n, theta, u[i], v[i], w[i], vis[i],
                                       u_mid, v_mid, w_mid, kern, P, R, S, w_min, w_step,
                                       0, n))
                        flops += 8ull * S * S;
                    else
                        bad_plane++;
                    k1++;
                }
                k = k1;
                orc_fft2d(subgrid, n, +1);
                if (first) {
                    last_wp = wp;
                    first = 0;
                }
                for (size_t p = 0; p < (size_t)n * n; p++) {
                    orc_c128 wt = orc_cipow(wtransfer[p], wp - last_wp);
                    orc_c128 s = cmul(wt, subimg[p]);
                    subimg[p].re = s.re + subgrid[p].re;
                    subimg[p].im = s.im + subgrid[p].im;
                }
                last_wp = wp;
            }
            if (last_wp != 0) {
                for (size_t p = 0; p < (size_t)n * n; p++)
                    subimg[p] = cdiv(subimg[p], orc_cipow(wtransfer[p], last_wp));
            }
            orc_fft2d(subimg, n, -1);
            have[t] = 1;
        }
        free(subgrid);
    }

    /* Fold (grid.c:503-516), row-owner so it is race-free in parallel. */
    double sg_norm = (double)n * n;
    orc_c128 *mid = grid + (size_t)(N + 1) * (N / 2);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int y = -N / 2; y < N / 2; y++) {
        for (int t = 0; t < cc * cc; t++) {
            if (!have[t])
                continue;
            int cx = t % cc, cy = t / cc;
            int x_min = chunk * cx - N / 2;
            int y_min = chunk * cy - N / 2;
            int x0 = x_min - margin / 2, x1 = x0 + n;
            int y0 = y_min - margin / 2, y1 = y0 + n;
            if (x0 < -N / 2) x0 = -N / 2;
            if (y0 < -N / 2) y0 = -N / 2;
            if (x1 > N / 2) x1 = N / 2;
            if (y1 > N / 2) y1 = N / 2;
            if (y < y0 || y >= y1)
                continue;
            const orc_c128 *subimg = subimgs + (size_t)t * n * n;
            for (int x = x0; x < x1; x++) {
                const orc_c128 s = subimg[(size_t)(x - x_min + margin / 2) +
                                          (size_t)(y - y_min + margin / 2) * n];
                mid[(ptrdiff_t)x + (ptrdiff_t)y * N].re += s.re / sg_norm;
                mid[(ptrdiff_t)x + (ptrdiff_t)y * N].im += s.im / sg_norm;
            }
        }
    }

    free(have);
    free(subimgs);
    free(tstart);
    free(keys);
    free(wtransfer);
    if (n_bad)
        *n_bad = bad + bad_plane;
    return flops;
}

/* ------------------------------------------------------------------ */
/* Direct DFT                                                           */
/* ------------------------------------------------------------------ */

/* dft_cuda.cu:15-39 (calculate_dft_sum_flat) at the pixel map of
 * dft_cuda.cu:98-112 (image_dft_flat, resolution = 1, hdf5.cuh:19). */
void orc_dft_flat(orc_c128 *out, const int *px, const int *py, int64_t npix,
                  int grid_size, double lambda,
                  const double *u, const double *v, const double *w,
                  const orc_c128 *vis, int64_t nvis, int nthreads)
{
    if (nthreads < 1)
        nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
    for (int64_t p = 0; p < npix; p++) {
        double l = ((double)(py[p] - grid_size / 2) / lambda) * 1.0;
        double m = ((double)(px[p] - grid_size / 2) / lambda) * 1.0;
        double nn = 1 - sqrt(1 - l * l - m * m);
        double re = 0, im = 0;
        for (int64_t i = 0; i < nvis; i++) {
            double angle = 2 * M_PI * (l * u[i] + m * v[i] + nn * w[i]);
            double c = cos(angle), s = sin(angle);
            re += vis[i].re * c + vis[i].im * s;
            im += -vis[i].re * s + vis[i].im * c;
        }
        out[p].re = re;
        out[p].im = im;
    }
}

/* ------------------------------------------------------------------ */
/* W-stacking predict (src/wstack/predict.cpp, wstack_common.h)         */
/* ------------------------------------------------------------------ */

/* wstack_predict (predict.cpp:410-428): grid g = floor(theta lam), oversampled
 * grid og = round(0.5 / x0) g, w_planes = ceil(sin(theta/2) sqrt(2 g^2) / dw)
 * + aa_w + 1. */
int orc_wstack_geometry(double theta, double lam, double dw, int aa_w, double x0, int *og,
                        int *w_planes)
{
    const int g = (int)floor(theta * lam);
    const double x0ih = round(0.5 / x0);
    *og = (int)(x0ih * g);
    /* generate_sky (predict.cpp:228-231) rounds after the product */
    const int og_sky = (int)round(0.5 / x0 * g);
    const double max_w = sin(theta / 2) * sqrt(2.0 * ((double)g * g));
    *w_planes = (int)ceil(max_w / dw) + aa_w + 1;
    return (*og > g && og_sky == *og) ? 0 : -1;
}

/* Quantised sky position of a point (generate_sky, predict.cpp:236-246). */
static int wstack_quantise(double l, double m, double theta, double lam, int g, int *lc, int *mc,
                           double *lq, double *mq)
{
    const double flc = floor((l / theta + 0.5) * (double)g);
    const double fmc = floor((m / theta + 0.5) * (double)g);
    if (!(flc >= 0 && flc < g && fmc >= 0 && fmc < g))
        return -1; /* outside the sky (or NaN) */
    *lc = (int)flc;
    *mc = (int)fmc;
    *lq = (double)*lc / lam - theta / 2;
    *mq = (double)*mc / lam - theta / 2;
    return 0;
}

/* The W-stack (predict.cpp:182-216, 219-278, 438-497): the sky of point
 * sources (flux / correction at the quantised pixel (row mc, column lc) of
 * the inner g x g region), fft-shifted, times F^(t0 + k) with
 * F = exp(2 pi i dw (1 - sqrt(1 - l^2 - m^2))) on the inner region and
 * t0 = -(w_planes / 2), forward 2-D FFT, fft-shifted: plane k of
 * stack[w_planes][og][og] (row v, column u).  F^(t0 + k) is evaluated as
 * exp(2 pi i (t0 + k) ph) (the reference takes an integer power of F and
 * then multiplies by F per plane: equal to rounding).  Points whose pixel or
 * correction index falls outside its table are skipped (the reference
 * indexes out of bounds); returns their number. */
int orc_wstack_build(orc_c128 *stack, double theta, double lam, const double *pl, const double *pm,
                     const double *flux, int npts, double du, double dw, int aa_w, double x0,
                     const double *corr_lm, int corr_lm_n, const double *corr_n, int corr_n_n)
{
    int og, w_planes;
    if (orc_wstack_geometry(theta, lam, dw, aa_w, x0, &og, &w_planes) != 0)
        return -1;
    const int g = (int)floor(theta * lam);
    const int gd = (og - g) / 2;
    const size_t plane = (size_t)og * og;
    orc_c128 *sky = calloc(plane, sizeof(orc_c128));
    double *ph = calloc(plane, sizeof(double));
    char *inner = calloc(plane, 1);
    int dropped = 0;
    for (int i = 0; i < npts; i++) {
        int lc, mc;
        double lq, mq;
        if (wstack_quantise(pl[i], pm[i], theta, lam, g, &lc, &mc, &lq, &mq) != 0) {
            dropped++;
            continue;
        }
        const double n = sqrt(1.0 - lq * lq - mq * mq) - 1.0;
        const double lm_step = 1.0 / (double)corr_lm_n, n_step = 1.0 / (double)corr_n_n;
        const int aau = (int)floor((du * lq) / lm_step) + corr_lm_n / 2;
        const int aav = (int)floor((du * mq) / lm_step) + corr_lm_n / 2;
        const int aaw = (int)floor((dw * n) / n_step) + corr_n_n / 2;
        if (aau < 0 || aau >= corr_lm_n || aav < 0 ||
            aav >= corr_lm_n || aaw < 0 || aaw >= corr_n_n) {
            dropped++;
            continue;
        }
        double a = 1.0;
        a *= corr_lm[aau];
        a *= corr_lm[aav];
        a *= corr_n[aaw];
        const double f = flux ? flux[i] : 1.0;
        orc_c128 *s = &sky[(size_t)(gd + mc) * og + (gd + lc)];
        s->re += f / a;
        s->im += 0.0 / a;
    }
    for (int y = 0; y < g; y++) {
        for (int x = 0; x < g; x++) {
            const double l = theta * ((double)x - g / 2) / g;
            const double m = theta * ((double)y - g / 2) / g;
            const size_t o = (size_t)(y + gd) * og + (x + gd);
            ph[o] = dw * (1.0 - sqrt(1.0 - l * l - m * m));
            inner[o] = 1;
        }
    }
    /* fft_shift_2Darray of the sky and of the screen (quadrant swap) */
    orc_fft_shift(sky, og);
    for (int y = 0; y < og / 2; y++) {
        for (int x = 0; x < og; x++) {
            const size_t a0 = (size_t)y * og + x;
            const size_t a1 = (size_t)(y + og / 2) * og + (x + og / 2) % og;
            double t = ph[a0];
            ph[a0] = ph[a1];
            ph[a1] = t;
            char c = inner[a0];
            inner[a0] = inner[a1];
            inner[a1] = c;
        }
    }
    const int t0 = -(w_planes / 2);
    for (int k = 0; k < w_planes; k++) {
        orc_c128 *p = stack + (size_t)k * plane;
        const double e = (double)(t0 + k);
        for (size_t o = 0; o < plane; o++) {
            if (!inner[o]) {
                p[o].re = p[o].im = 0.0;
                continue;
            }
            const double c = cos(TWO_PI * e * ph[o]), s = sin(TWO_PI * e * ph[o]);
            p[o].re = sky[o].re * c - sky[o].im * s;
            p[o].im = sky[o].re * s + sky[o].im * c;
        }
        orc_fft2d(p, og, -1);
        orc_fft_shift(p, og);
    }
    free(sky);
    free(ph);
    free(inner);
    return dropped;
}

/* deconvolve_visibility_ (wstack_common.h:420-470): separable uv / w kernel
 * taps over the stack,
 *   vis = sum_{dwi, dvi, dui} stack[dws][dvs][dus] (cw cv) cu,
 * dus = ceil(u/du) + g + dui, dvs = ceil(v/du) + g + dvi,
 * dws = ceil(w/dw) + w_planes/2 + dwi, dui, dvi in [-aa_uv/2, aa_uv/2),
 * dwi in [-aa_w/2, aa_w/2), tap rows aa * ov + (d + aa/2) with
 * ov = floor(|x - ceil(x/d) d| / d * oversampling) (clamped to
 * oversampling - 1).  A visibility whose taps leave the stack is not predicted
 * (0) and counted. */
int64_t orc_wstack_degrid(orc_c128 *out, const double *u, const double *v, const double *w,
                          int64_t nvis, const orc_c128 *stack, int og, int w_planes, int g,
                          double du, double dw, int aa_uv, int aa_w, const double *conv_uv,
                          int ov_uv, const double *conv_w, int ov_w, int nthreads)
{
    int64_t dropped = 0;
    const int aa_h = aa_uv / 2, aaw_h = aa_w / 2;
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static, 1000) reduction(+ : dropped) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < nvis; i++) {
        const double cu = ceil(u[i] / du), cv = ceil(v[i] / du), cw = ceil(w[i] / dw);
        const double flu = u[i] - cu * du, flv = v[i] - cv * du, flw = w[i] - cw * dw;
        int ovu = (int)floor(fabs(flu) / du * ov_uv);
        int ovv = (int)floor(fabs(flv) / du * ov_uv);
        int ovw = (int)floor(fabs(flw) / dw * ov_w);
        ovu = ovu < ov_uv ? ovu : ov_uv - 1;
        ovv = ovv < ov_uv ? ovv : ov_uv - 1;
        ovw = ovw < ov_w ? ovw : ov_w - 1;
        const int us = (int)cu + g, vs = (int)cv + g, ws = (int)cw + w_planes / 2;
        orc_c128 acc = { 0.0, 0.0 };
        if (!(cu > -1e9 && cu < 1e9 && cv > -1e9 && cv < 1e9 && cw > -1e9 && cw < 1e9) ||
            us - aa_h < 0 || us + aa_h > og || vs - aa_h < 0 || vs + aa_h > og ||
            ws - aaw_h < 0 || ws + aaw_h > w_planes) {
            dropped++;
            out[i] = acc;
            continue;
        }
        for (int dwi = -aaw_h; dwi < aaw_h; dwi++) {
            const double gw = conv_w[aa_w * ovw + (dwi + aaw_h)];
            for (int dvi = -aa_h; dvi < aa_h; dvi++) {
                const double gv = gw * conv_uv[aa_uv * ovv + (dvi + aa_h)];
                const orc_c128 *row = stack + ((size_t)(ws + dwi) * og + (vs + dvi)) * og;
                for (int dui = -aa_h; dui < aa_h; dui++) {
                    const double gu = gv * conv_uv[aa_uv * ovu + (dui + aa_h)];
                    acc.re += row[us + dui].re * gu;
                    acc.im += row[us + dui].im * gu;
                }
            }
        }
        out[i] = acc;
    }
    return dropped;
}

/* predict_visibility_quantized_vec (predict.cpp:88-131): the direct sum over
 * the quantised points, sum flux exp(-2 pi i (u lq + v mq + w (n - 1))). */
void orc_predict_quantized(orc_c128 *out, const double *pl, const double *pm, const double *flux,
                           int npts, double theta, double lam, const double *u, const double *v,
                           const double *w, int64_t nvis, int nthreads)
{
    const int g = (int)floor(theta * lam);
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < nvis; i++) {
        double re = 0.0, im = 0.0;
        for (int p = 0; p < npts; p++) {
            int lc, mc;
            double lq, mq;
            if (wstack_quantise(pl[p], pm[p], theta, lam, g, &lc, &mc, &lq, &mq) != 0)
                continue; /* outside the sky */
            const double n = sqrt(1.0 - lq * lq - mq * mq) - 1.0;
            const double ang = -TWO_PI * (u[i] * lq + v[i] * mq + w[i] * n);
            const double f = flux ? flux[p] : 1.0;
            re += f * cos(ang);
            im += f * sin(ang);
        }
        out[i].re = re;
        out[i].im = im;
    }
}

/* ------------------------------------------------------------------ */
/* EPIC E-field gridding (src/epic/epic_grid.cu)                        */
/* ------------------------------------------------------------------ */

/* scatter_grid_point / scatter_grid_kernel (epic_grid.cu:54-179) and
 * epic_naive (epic_grid.cu:337-359): for every batch b and antenna a,
 * grid[b][y_a + ty][x_a + tx] += conj(illum[ty][tx]) f[b][a] over the
 * support x support illumination; taps outside [0, g)^2 are dropped
 * (scatter_grid_add's bounds check).  The naive loop multiplies by illum
 * itself, which is the same for the real top-hat the reference uses. */
void orc_epic_grid(orc_c128 *grids, const orc_c128 *f, const int *x, const int *y, int n_ant,
                   int nbatch, const orc_c128 *illum, int supp, int g)
{
    memset(grids, 0, sizeof(orc_c128) * (size_t)nbatch * g * g);
    for (int b = 0; b < nbatch; b++) {
        orc_c128 *grid = grids + (size_t)b * g * g;
        for (int a = 0; a < n_ant; a++) {
            const size_t i = (size_t)b * n_ant + a;
            for (int ty = 0; ty < supp; ty++) {
                for (int tx = 0; tx < supp; tx++) {
                    const int X = x[i] + tx, Y = y[i] + ty;
                    if (X < 0 || X >= g || Y < 0 || Y >= g)
                        continue;
                    const orc_c128 k = illum[ty * supp + tx];
                    orc_c128 *c = &grid[(size_t)Y * g + X];
                    c->re += k.re * f[i].re + k.im * f[i].im;
                    c->im += k.re * f[i].im - k.im * f[i].re;
                }
            }
        }
    }
}
