// dft.hip -- direct DFT (the accuracy oracle of the reference, dft_cuda.cu)
// and point-source prediction of synthetic visibilities.
#include <algorithm>

#include "binning.hpp"

namespace rg {

constexpr int DFT_THREADS = 256;

// One thread per pixel; visibilities are staged through LDS in chunks of
// DFT_THREADS.  FP64 mode restates calculate_dft_sum_flat (dft_cuda.cu:15-39)
// at the pixel map of image_dft_flat (dft_cuda.cu:98-112).  MIXED mode keeps
// the phase l u + m v + (1 - n) w in FP64, reduces it mod 1, evaluates the
// rotation in FP32 and accumulates each chunk in FP32 into an FP64 sum.
// With few pixels the visibility stream is also split over gridDim.y ranges:
// range r writes its partial sums to out + r * npix and k_dft_sum adds the
// ranges in order (the same summation order on every run).
template <int MODE, bool LIST, typename VT>
__global__ __launch_bounds__(DFT_THREADS) void k_dft(const double *__restrict__ u,
                                                     const double *__restrict__ v,
                                                     const double *__restrict__ w,
                                                     const VT *__restrict__ vis, int64_t nvis,
                                                     const int32_t *__restrict__ px,
                                                     const int32_t *__restrict__ py, int64_t npix,
                                                     int N, double lambda, double2 *out)
{
    const int64_t per = (nvis + gridDim.y - 1) / gridDim.y;
    const int64_t v0 = per * blockIdx.y, v1 = min(nvis, v0 + per);
    out += (int64_t)blockIdx.y * npix;
    __shared__ double su[DFT_THREADS], sv[DFT_THREADS], sw[DFT_THREADS];
    __shared__ double2 svis[DFT_THREADS];
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int x = 0, y = 0;
    if (p < npix) {
        if (LIST) {
            x = px[p];
            y = py[p];
        } else {
            y = (int)(p / N);
            x = (int)(p - (int64_t)y * N);
        }
    }
    const double l = ((double)(y - N / 2) / lambda) * 1.0;
    const double m = ((double)(x - N / 2) / lambda) * 1.0;
    const double nn = 1 - sqrt(1 - l * l - m * m);
    double re = 0.0, im = 0.0;
    for (int64_t base = v0; base < v1; base += DFT_THREADS) {
        const int nloc = (int)min((int64_t)DFT_THREADS, v1 - base);
        __syncthreads();
        if ((int)threadIdx.x < nloc) {
            const int64_t i = base + threadIdx.x;
            su[threadIdx.x] = u[i];
            sv[threadIdx.x] = v[i];
            sw[threadIdx.x] = w[i];
            svis[threadIdx.x] = make_double2((double)vis[i].x, (double)vis[i].y);
        }
        __syncthreads();
        if (MODE == RG_DFT_FP64) {
            for (int j = 0; j < nloc; ++j) {
                const double angle = 2 * M_PI * (l * su[j] + m * sv[j] + nn * sw[j]);
                double s, c;
                sincos(angle, &s, &c);
                const double2 vv = svis[j];
                re += vv.x * c + vv.y * s;
                im += -vv.x * s + vv.y * c;
            }
        } else {
            float fre = 0.f, fim = 0.f;
            for (int j = 0; j < nloc; ++j) {
                const double ph = l * su[j] + m * sv[j] + nn * sw[j];
                const float f = (float)(ph - rint(ph)); // revolutions in [-1/2, 1/2]
                float s, c;
                sincospif(2.0f * f, &s, &c);
                const float vr = (float)svis[j].x, vi = (float)svis[j].y;
                fre += vr * c + vi * s;
                fim += -vr * s + vi * c;
            }
            re += fre;
            im += fim;
        }
    }
    if (p < npix)
        out[p] = make_double2(re, im);
}

// out[p] = sum over ranges r of part[r][p], r in order
__global__ __launch_bounds__(256) void k_dft_sum(const double2 *__restrict__ part, int nr,
                                                 int64_t npix, double2 *out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix)
        return;
    double2 a = part[p];
    for (int r = 1; r < nr; ++r) {
        const double2 b = part[(int64_t)r * npix + p];
        a.x += b.x;
        a.y += b.y;
    }
    out[p] = a;
}

template <int MODE, bool LIST>
static int launch_dft(rg_context *ctx, const rg_flat_vis *vis, const int32_t *px,
                      const int32_t *py, int64_t npix, int N, double lambda, double2 *out,
                      hipStream_t s)
{
    const int64_t blocks = (npix + DFT_THREADS - 1) / DFT_THREADS;
    if (blocks == 0)
        return RG_OK;
    // visibility ranges: enough workgroups to fill the chip (~4096), ranges of
    // at least 64 K visibilities, at most 256
    int64_t nr = std::min<int64_t>(256, (4096 + blocks - 1) / blocks);
    nr = std::max<int64_t>(1, std::min<int64_t>(nr, vis->n / 65536));
    double2 *dst = out;
    if (nr > 1) {
        RG_TRY(ctx->scratch[14].ensure(sizeof(double2) * (size_t)(nr * npix)));
        dst = ctx->scratch[14].as<double2>();
    }
    const dim3 grid((unsigned)blocks, (unsigned)nr);
    if (vis->vis_type == RG_C128)
        hipLaunchKernelGGL((k_dft<MODE, LIST, double2>), grid, dim3(DFT_THREADS), 0, s, vis->u,
                           vis->v, vis->w, (const double2 *)vis->vis, vis->n, px, py, npix, N,
                           lambda, dst);
    else
        hipLaunchKernelGGL((k_dft<MODE, LIST, float2>), grid, dim3(DFT_THREADS), 0, s, vis->u,
                           vis->v, vis->w, (const float2 *)vis->vis, vis->n, px, py, npix, N,
                           lambda, dst);
    RG_HIP(hipGetLastError());
    if (nr > 1) {
        hipLaunchKernelGGL(k_dft_sum, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s,
                           (const double2 *)dst, (int)nr, npix, out);
        RG_HIP(hipGetLastError());
    }
    return RG_OK;
}

int dft_run(rg_context *ctx, const rg_flat_vis *vis, const int32_t *px, const int32_t *py,
            int64_t npix, int N, double lambda, void *out, int mode, hipStream_t s)
{
    const bool list = px != nullptr;
    double2 *o = (double2 *)out;
    if (mode == RG_DFT_FP64)
        return list ? launch_dft<RG_DFT_FP64, true>(ctx, vis, px, py, npix, N, lambda, o, s)
                    : launch_dft<RG_DFT_FP64, false>(ctx, vis, px, py, npix, N, lambda, o, s);
    return list ? launch_dft<RG_DFT_MIXED, true>(ctx, vis, px, py, npix, N, lambda, o, s)
                : launch_dft<RG_DFT_MIXED, false>(ctx, vis, px, py, npix, N, lambda, o, s);
}

// ---- point-source prediction ----------------------------------------------------

template <typename VT>
__global__ __launch_bounds__(256) void k_predict(const double *__restrict__ u,
                                                 const double *__restrict__ v,
                                                 const double *__restrict__ w, int64_t n,
                                                 const double *__restrict__ l,
                                                 const double *__restrict__ m,
                                                 const double *__restrict__ flux, int nsrc,
                                                 VT *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double ui = u[i], vi = v[i], wi = w[i];
    double re = 0.0, im = 0.0;
    for (int k = 0; k < nsrc; ++k) {
        const double nk = sqrt(1.0 - l[k] * l[k] - m[k] * m[k]);
        const double ph = ui * l[k] + vi * m[k] + wi * (nk - 1.0);
        const double f = ph - rint(ph);
        double s, c;
        sincospi(2.0 * f, &s, &c);
        // exp(-2 pi i ph)
        re += flux[k] * c;
        im -= flux[k] * s;
    }
    out[i].x = re;
    out[i].y = im;
}

int predict_run(const double *u, const double *v, const double *w, int64_t n, const double *l,
                const double *m, const double *flux, int nsrc, void *out, int type, hipStream_t s)
{
    const int64_t blocks = (n + 255) / 256;
    if (blocks == 0)
        return RG_OK;
    if (type == RG_C128)
        hipLaunchKernelGGL(k_predict<double2>, dim3((unsigned)blocks), dim3(256), 0, s, u, v, w, n,
                           l, m, flux, nsrc, (double2 *)out);
    else
        hipLaunchKernelGGL(k_predict<float2>, dim3((unsigned)blocks), dim3(256), 0, s, u, v, w, n, l,
                           m, flux, nsrc, (float2 *)out);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

} // namespace rg
