// wstack.hip -- W-stacking predict (sky -> visibilities) with separable
// (Sze-Tan) kernels: wstack_predict_cu_3D (src/wstack/predict.cuh:33-47;
// CPU semantics wstack_predict, src/wstack/predict.cpp:396-560, and
// deconvolve_visibility_, src/wstack/wstack_common.h:420-470).
//
// Stack: w_planes planes of og x og FP64 complex (row v, column u) in HBM.
// Plane k is the fft-shifted forward FFT of the fft-shifted sky times
// F^(t0 + k), F = exp(2 pi i dw (1 - sqrt(1 - l^2 - m^2))) on the inner
// g x g region, t0 = -(w_planes / 2).  The sky is a sum of point sources at
// their quantised pixels, so each plane's input is a zero plane plus one
// value per source; one batched FFT then transforms every plane.
//
// Degridding: one wavefront per visibility, lanes spread over the kernel's
// taps so that each load instruction reads whole stack rows; the wavefront
// reduces with shuffles.
#include "binning.hpp"

namespace rg {

// ---- stack geometry -------------------------------------------------------------------------
// wstack_predict (predict.cpp:410-428) and generate_sky (predict.cpp:228-231)
int wstack_geometry(double theta, double lambda, double dw, int aa_w, double x0, int &g, int &og,
                    int &w_planes)
{
    g = (int)std::floor(theta * lambda);
    const double x0ih = std::round(0.5 / x0);
    og = (int)(x0ih * g);
    const int og_sky = (int)std::round(0.5 / x0 * g);
    const double max_w = std::sin(theta / 2) * std::sqrt(2.0 * ((double)g * g));
    w_planes = (int)std::ceil(max_w / dw) + aa_w + 1;
    if (g <= 0 || og <= g || og_sky != og)
        return fail(RG_ERR_PARAM, "inconsistent W-stack geometry (g %d, og %d, x0 %g)", g, og, x0);
    if (og < 4 || og > 4096 || (og & (og - 1)) != 0)
        return fail(RG_ERR_UNSUPPORTED, "W-stack plane size %d is not a power of two in 4..4096", og);
    return RG_OK;
}

// ---- sky into the planes --------------------------------------------------------------------
// Per source (generate_sky, predict.cpp:219-278): quantised pixel (lc, mc),
// correction a = corr_lm[aau] corr_lm[aav] corr_n[aaw]; adds flux / a times
// exp(2 pi i (t0 + k) ph) into every plane k at the source's fft-shifted
// position, ph = dw (1 - sqrt(1 - l^2 - m^2)) of the pixel (generate_fresnel,
// predict.cpp:182-216).  A source whose pixel or correction index falls
// outside its table is skipped and counted (the reference indexes out of
// bounds).
__global__ __launch_bounds__(256) void k_ws_sources(
    double2 *stack, int og, int w_planes, int g, double theta, double lambda, double du, double dw,
    const double *__restrict__ pl, const double *__restrict__ pm, const double *__restrict__ flux,
    int npts, const double *__restrict__ corr_lm, int n_lm, const double *__restrict__ corr_n,
    int n_n, unsigned long long *skipped)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts)
        return;
    const double flc = floor((pl[i] / theta + 0.5) * (double)g);
    const double fmc = floor((pm[i] / theta + 0.5) * (double)g);
    if (!(flc >= 0 && flc < g && fmc >= 0 && fmc < g)) { // also NaN
        atomicAdd(skipped, 1ull);
        return;
    }
    const int lc = (int)flc, mc = (int)fmc;
    const double lq = (double)lc / lambda - theta / 2;
    const double mq = (double)mc / lambda - theta / 2;
    const double n = sqrt(1.0 - lq * lq - mq * mq) - 1.0;
    const double lm_step = 1.0 / (double)n_lm, n_step = 1.0 / (double)n_n;
    const double fu = floor((du * lq) / lm_step), fv = floor((du * mq) / lm_step);
    const double fw = floor((dw * n) / n_step);
    const bool ok = fabs(fu) < 1e9 && fabs(fv) < 1e9 && fabs(fw) < 1e9;
    const int aau = ok ? (int)fu + n_lm / 2 : -1, aav = ok ? (int)fv + n_lm / 2 : -1;
    const int aaw = ok ? (int)fw + n_n / 2 : -1;
    if (!ok || aau < 0 || aau >= n_lm || aav < 0 || aav >= n_lm || aaw < 0 || aaw >= n_n) {
        atomicAdd(skipped, 1ull);
        return;
    }
    double a = 1.0;
    a *= corr_lm[aau];
    a *= corr_lm[aav];
    a *= corr_n[aaw];
    const double val = (flux ? flux[i] : 1.0) / a;
    // the screen at the source's pixel (x = lc, y = mc of the inner region)
    const double l = theta * ((double)lc - g / 2) / g;
    const double m = theta * ((double)mc - g / 2) / g;
    const double ph = dw * (1.0 - sqrt(1.0 - l * l - m * m));
    const int gd = (og - g) / 2;
    const int xs = (gd + lc + og / 2) % og, ys = (gd + mc + og / 2) % og;
    const int t0 = -(w_planes / 2);
    for (int k = 0; k < w_planes; ++k) {
        double s, c;
        sincospi(2.0 * (double)(t0 + k) * ph, &s, &c);
        double *dst = reinterpret_cast<double *>(stack + ((size_t)k * og + ys) * og + xs);
        atomicAdd(dst, val * c);
        atomicAdd(dst + 1, val * s);
    }
}

// ---- degridding ------------------------------------------------------------------------------
// deconvolve_visibility_ (wstack_common.h:420-470).  One wavefront per
// visibility (grid-stride over visibilities).  Lane = (pair, tap): tap
// t = lane % aa_uv along u, pair q = lane / aa_uv + k (64 / aa_uv) over the
// aa_w * aa_uv (plane, row) pairs, so one load instruction reads 64 / aa_uv
// whole kernel rows (contiguous runs of the stack).  Each lane sums
// stack[dws][dvs][dus] (cw cv) cu over its pairs; the wavefront reduces with
// shuffles.  A visibility whose taps leave the stack is predicted as 0 and
// counted.
template <int LV>
__global__ __launch_bounds__(256) void k_ws_degrid(
    const double2 *__restrict__ stack, int og, int w_planes, int g, const double *__restrict__ u,
    const double *__restrict__ v, const double *__restrict__ w, int64_t nvis, double du, double dw,
    int aa_uv, int aa_w, const double *__restrict__ conv_uv, int ov_uv,
    const double *__restrict__ conv_w, int ov_w, double2 *out, unsigned long long *dropped)
{
#pragma clang fp contract(off)
    // LV lanes per visibility (64 / LV visibilities per wavefront)
    const int lane = (int)(threadIdx.x % LV);
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LV;
    const int64_t waves = ((int64_t)gridDim.x * blockDim.x) / LV;
    const int aa_h = aa_uv / 2, aaw_h = aa_w / 2;
    const int t = lane % aa_uv, q0 = lane / aa_uv, qs = LV / aa_uv, npairs = aa_w * aa_uv;
    // every lane of a wavefront runs the same number of iterations
    const int64_t iters = (nvis + waves - 1) / waves;
    for (int64_t it = 0; it < iters; ++it) {
        const int64_t i0 = wave + it * waves;
        const bool live = i0 < nvis;
        const int64_t i = live ? i0 : 0;
        const double ui = u[i], vi = v[i], wi = w[i];
        const double cu = ceil(ui / du), cv = ceil(vi / du), cw = ceil(wi / dw);
        const bool finite =
            cu > -1e9 && cu < 1e9 && cv > -1e9 && cv < 1e9 && cw > -1e9 && cw < 1e9;
        const int us = finite ? (int)cu + g : 0, vs = finite ? (int)cv + g : 0;
        const int ws = finite ? (int)cw + w_planes / 2 : 0;
        const bool inside = finite && us - aa_h >= 0 && us + aa_h <= og && vs - aa_h >= 0 &&
                            vs + aa_h <= og && ws - aaw_h >= 0 && ws + aaw_h <= w_planes;
        double re = 0.0, im = 0.0;
        if (live && inside && t < aa_uv && q0 < qs) {
            const double flu = ui - cu * du, flv = vi - cv * du, flw = wi - cw * dw;
            const int ovu = min((int)floor(fabs(flu) / du * ov_uv), ov_uv - 1);
            const int ovv = min((int)floor(fabs(flv) / du * ov_uv), ov_uv - 1);
            const int ovw = min((int)floor(fabs(flw) / dw * ov_w), ov_w - 1);
            const double gu = conv_uv[aa_uv * ovu + t];
            const double2 *col = stack + (us - aa_h + t);
            for (int q = q0; q < npairs; q += qs) {
                const int pw = q / aa_uv, pv = q - pw * aa_uv; // dwi + aaw_h, dvi + aa_h
                const double gw = conv_w[aa_w * ovw + pw];
                const double gv = gw * conv_uv[aa_uv * ovv + pv];
                const double2 s = col[((size_t)(ws - aaw_h + pw) * og + (vs - aa_h + pv)) * og];
                const double k = gv * gu;
                re += s.x * k;
                im += s.y * k;
            }
        }
#pragma unroll
        for (int o = LV / 2; o > 0; o >>= 1) {
            re += __shfl_xor(re, o, LV);
            im += __shfl_xor(im, o, LV);
        }
        if (live && lane == 0) {
            out[i] = make_double2(inside ? re : 0.0, inside ? im : 0.0);
            if (!inside)
                atomicAdd(dropped, 1ull);
        }
    }
}

// ---- driver --------------------------------------------------------------------------------------

int wstack_build(rg_context *ctx, const rg_wstack_params *p, const double *l, const double *m,
                 const double *flux, int32_t npts, const rg_sep_kern *corr_lm,
                 const rg_sep_kern *corr_n, hipStream_t s)
{
    int g, og, wp;
    RG_TRY(wstack_geometry(p->theta, p->lambda, p->dw, p->aa_support_w, p->x0, g, og, wp));
    const size_t plane = (size_t)og * og;
    RG_TRY(ctx->scratch[17].ensure(sizeof(double2) * plane * wp));
    double2 *stack = ctx->scratch[17].as<double2>();
    RG_TRY(ensure_counters(ctx));
    RG_TRY(zero_fill(ctx->counters.p, sizeof(unsigned long long) * CNT__N, s));
    auto *cnt = ctx->counters.as<unsigned long long>();
    RG_TRY(zero_fill(stack, sizeof(double2) * plane * wp, s));
    if (npts > 0) {
        hipLaunchKernelGGL(k_ws_sources, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s,
                           stack, og, wp, g, p->theta, p->lambda, p->du, p->dw, l, m, flux, npts,
                           corr_lm->data, corr_lm->size * corr_lm->oversampling, corr_n->data,
                           corr_n->size * corr_n->oversampling, cnt + CNT_SKY_SKIPPED);
        RG_HIP(hipGetLastError());
    }
    // the sources sit at their fft-shifted positions already
    RG_TRY(fft2d(ctx, stack, RG_C128, og, wp, -1, s));
    RG_TRY(fft_shift(stack, RG_C128, og, wp, s));
    ctx->ws_g = g;
    ctx->ws_og = og;
    ctx->ws_planes = wp;
    return RG_OK;
}

int wstack_degrid(rg_context *ctx, const rg_wstack_params *p, const double *u, const double *v,
                  const double *w, int64_t nvis, const rg_sep_kern *conv_uv,
                  const rg_sep_kern *conv_w, void *out, hipStream_t s)
{
    int g, og, wp;
    RG_TRY(wstack_geometry(p->theta, p->lambda, p->dw, p->aa_support_w, p->x0, g, og, wp));
    if (ctx->ws_og != og || ctx->ws_planes != wp || ctx->ws_g != g)
        return fail(RG_ERR_PARAM, "no W-stack of this geometry has been built in the context");
    auto *cnt = ctx->counters.as<unsigned long long>();
    // per call: visibilities not predicted (the sky counter is kept)
    RG_TRY(zero_fill(cnt + CNT_DROPPED, sizeof(unsigned long long) * 2, s));
    ctx->last_nvis = nvis;
    if (nvis == 0)
        return RG_OK;
    const double2 *stack = ctx->scratch[17].as<double2>();
    // LV lanes per visibility: at least one whole kernel row per group
    constexpr int LV_DEFAULT = 64;
    const int lv = std::max(LV_DEFAULT, (int)p->aa_support_uv);
    auto launch = [&](auto lanes) {
        constexpr int LV = decltype(lanes)::value;
        const int per_block = 256 / LV;
        const int64_t blocks = std::min<int64_t>((nvis + per_block - 1) / per_block, (int64_t)1 << 20);
        hipLaunchKernelGGL(k_ws_degrid<LV>, dim3((unsigned)blocks), dim3(256), 0, s, stack, og, wp,
                           g, u, v, w, nvis, p->du, p->dw, p->aa_support_uv, p->aa_support_w,
                           conv_uv->data, conv_uv->oversampling, conv_w->data,
                           conv_w->oversampling, (double2 *)out, cnt + CNT_DROPPED);
    };
    if (lv <= 16)
        launch(std::integral_constant<int, 16>());
    else if (lv <= 32)
        launch(std::integral_constant<int, 32>());
    else
        launch(std::integral_constant<int, 64>());
    RG_HIP(hipGetLastError());
    return RG_OK;
}

} // namespace rg
