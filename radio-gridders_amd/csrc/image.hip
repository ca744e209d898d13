// image.hip -- the image conventions of the CPU reference pipeline
// (src/reference/main.c:151-154, 258-281): uniform weighting of the
// visibilities before gridding (weight, src/reference/grid.c:59-81), hermitian
// completion of the uv grid (make_hermitian, grid.c:639-662), and a separable
// grid-correction division of the image by the kernel taper A(l) A(m).
#include "binning.hpp"

namespace rg {

// ---- uniform weighting --------------------------------------------------------------------
// Cell of a visibility (coord, grid.c:25-35): x = floor(theta u + 1/2),
// y = floor(theta v + 1/2) with the product and the sum rounded separately
// (no contraction), index (y + N/2) N + (x + N/2).  A cell outside the
// grid (which the reference indexes out of bounds) gives false: such a
// visibility is neither counted nor weighted.
__device__ inline bool weight_cell(double u, double v, double theta, int N, uint32_t &cell)
{
#pragma clang fp contract(off)
    const double fx = floor(theta * u + .5), fy = floor(theta * v + .5);
    const double lo = -(double)(N / 2), hi = (double)(N - N / 2);
    if (!(fx >= lo && fx < hi && fy >= lo && fy < hi))
        return false;
    cell = (uint32_t)((int)fy + N / 2) * (uint32_t)N + (uint32_t)((int)fx + N / 2);
    return true;
}

// visibilities per cell (wavefront runs of one cell add once)
__global__ __launch_bounds__(256) void k_weight_count(const double *__restrict__ u,
                                                      const double *__restrict__ v, int64_t n,
                                                      double theta, int N, uint32_t *count)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t cell = 0;
    const bool ok = i < n && weight_cell(u[i], v[i], theta, N, cell);
    wave_agg_add(count, cell, ok, false);
}

// vis /= count of its cell, in FP64 (the FP32 form rounds the FP64 quotient)
template <typename T2>
__global__ __launch_bounds__(256) void k_weight_apply(const double *__restrict__ u,
                                                      const double *__restrict__ v, int64_t n,
                                                      double theta, int N,
                                                      const uint32_t *__restrict__ count, T2 *vis)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t cell;
    if (i >= n || !weight_cell(u[i], v[i], theta, N, cell))
        return;
    using T = decltype(vis[0].x);
    const double c = (double)count[cell];
    T2 a = vis[i];
    a.x = (T)((double)a.x / c);
    a.y = (T)((double)a.y / c);
    vis[i] = a;
}

int weight_vis(rg_context *ctx, const double *u, const double *v, int64_t n, void *vis,
               int vis_type, int N, double theta, uint32_t *counts, hipStream_t s)
{
    const size_t cells = (size_t)N * N;
    if (!counts) {
        RG_TRY(ctx->scratch[16].ensure(sizeof(uint32_t) * cells));
        counts = ctx->scratch[16].as<uint32_t>();
    }
    RG_TRY(zero_fill(counts, sizeof(uint32_t) * cells, s));
    if (n == 0)
        return RG_OK;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_weight_count, dim3(blocks), dim3(256), 0, s, u, v, n, theta, N, counts);
    RG_HIP(hipGetLastError());
    if (vis_type == RG_C128)
        hipLaunchKernelGGL(k_weight_apply<double2>, dim3(blocks), dim3(256), 0, s, u, v, n, theta, N,
                           (const uint32_t *)counts, (double2 *)vis);
    else
        hipLaunchKernelGGL(k_weight_apply<float2>, dim3(blocks), dim3(256), 0, s, u, v, n, theta, N,
                           (const uint32_t *)counts, (float2 *)vis);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

// ---- hermitian completion -------------------------------------------------------------------
// The reference walks p0 up from `base` (n + 1 for even n, 0 for odd n) and
// p1 down from n^2 - 1, adding each one's conjugate into the other, until
// they meet on the zero frequency (n + 1)(n / 2), which then gets its own
// conjugate added.  Pair k is (base + k, n^2 - 1 - k); the pairs are
// disjoint, so one thread per pair updates both cells.
template <typename T2>
__global__ __launch_bounds__(256) void k_make_hermitian(T2 *a, int64_t base, int64_t npairs,
                                                        int64_t last, int64_t center)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < npairs;
         k += (int64_t)gridDim.x * blockDim.x) {
        const T2 g0 = a[base + k], g1 = a[last - k];
        T2 r0, r1;
        r0.x = g0.x + g1.x;
        r0.y = g0.y - g1.y;
        r1.x = g1.x + g0.x;
        r1.y = g1.y - g0.y;
        a[base + k] = r0;
        a[last - k] = r1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        T2 c = a[center];
        c.x = c.x + c.x;
        c.y = c.y - c.y;
        a[center] = c;
    }
}

int make_hermitian(void *grid, int type, int n, hipStream_t s)
{
    const int64_t nn = (int64_t)n * n;
    const int64_t base = n % 2 == 0 ? n + 1 : 0;
    const int64_t center = (int64_t)(n + 1) * (n / 2);
    const int64_t npairs = center - base;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((npairs + 255) / 256, 65536));
    if (type == RG_C128)
        hipLaunchKernelGGL(k_make_hermitian<double2>, dim3((unsigned)blocks), dim3(256), 0, s,
                           (double2 *)grid, base, npairs, nn - 1, center);
    else
        hipLaunchKernelGGL(k_make_hermitian<float2>, dim3((unsigned)blocks), dim3(256), 0, s,
                           (float2 *)grid, base, npairs, nn - 1, center);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

// ---- grid correction ----------------------------------------------------------------------
// image[r][c] /= taper[r] taper[c] (FP64 quotient)
template <typename T2>
__global__ __launch_bounds__(256) void k_gcf_correct(T2 *img, int n,
                                                     const double *__restrict__ taper)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * n)
        return;
    using T = decltype(img[0].x);
    const int r = (int)(i / n), c = (int)(i - (int64_t)r * n);
    const double t = taper[r] * taper[c];
    T2 a = img[i];
    a.x = (T)((double)a.x / t);
    a.y = (T)((double)a.y / t);
    img[i] = a;
}

int gcf_correct(void *img, int type, int n, const double *taper, hipStream_t s)
{
    const int64_t nn = (int64_t)n * n;
    const unsigned blocks = (unsigned)((nn + 255) / 256);
    if (type == RG_C128)
        hipLaunchKernelGGL(k_gcf_correct<double2>, dim3(blocks), dim3(256), 0, s, (double2 *)img, n,
                           taper);
    else
        hipLaunchKernelGGL(k_gcf_correct<float2>, dim3(blocks), dim3(256), 0, s, (float2 *)img, n,
                           taper);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

} // namespace rg
