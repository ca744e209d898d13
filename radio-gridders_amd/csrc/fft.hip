// fft.hip -- batched 2-D FFTs, fft-shift and transposes.
//
// n <= 512: row pass (B rows per block, padded LDS) then column pass
// (CB columns per block, each row segment of the block is one 128-byte line).
// n > 512: row pass, tiled transpose, row pass, transpose back.
#include "fft_core.hpp"

namespace rg {

// ---- row pass ------------------------------------------------------------------
// Transforms rows [blockIdx.x*B, +B) of nrows rows of length n stored
// contiguously.  row_flags (optional): bit r set = row r non-zero; rows with
// the bit clear are left untouched (they are zero and stay zero).  dev_batch
// (optional, device): only the first *dev_batch arrays of n rows are
// transformed.
template <typename T2, int LOGN, int B, bool INV>
__global__ __launch_bounds__(B *((1 << LOGN) / 8)) void k_fft_rows(T2 *data, int64_t nrows,
                                                                   const T2 *__restrict__ twg,
                                                                   const uint32_t *row_flags,
                                                                   const int32_t *dev_batch)
{
    constexpr int n = 1 << LOGN;
    constexpr int NT = B * (n / 8);
    constexpr int PAD = B > 1 ? 1 : 0;
    constexpr int LS = B + PAD;
    constexpr bool TW_LDS = n * sizeof(T2) <= 8192;
    __shared__ T2 lds[n * LS];
    __shared__ T2 tws[TW_LDS ? n : 1];
    const T2 *tw = twg;
    if (TW_LDS) {
        for (int k = threadIdx.x; k < n; k += NT)
            tws[k] = twg[k];
        tw = tws;
    }
    const int64_t row0 = (int64_t)blockIdx.x * B;
    if (dev_batch && row0 >= (int64_t)*dev_batch * n)
        return; // the block's rows (B divides n) lie in an inactive array
    // block-uniform skip when every row of the block is empty
    bool any = true;
    if (row_flags) {
        any = false;
        for (int l = 0; l < B; ++l) {
            const int64_t r = row0 + l;
            if (r < nrows && ((row_flags[r >> 5] >> (r & 31)) & 1u))
                any = true;
        }
    }
    if (!any)
        return;
    // load: global order (line-major), element e -> (line, idx)
    for (int e = threadIdx.x; e < B * n; e += NT) {
        const int line = e / n, idx = e - line * n;
        const int64_t r = row0 + line;
        T2 val;
        val.x = 0;
        val.y = 0;
        if (r < nrows)
            val = data[r * n + idx];
        lds[idx * LS + line] = val;
    }
    __syncthreads();
    fft_block<T2, LOGN, B, LS, INV>(lds, tw);
    for (int e = threadIdx.x; e < B * n; e += NT) {
        const int line = e / n, idx = e - line * n;
        const int64_t r = row0 + line;
        if (r < nrows) {
            if (!row_flags || ((row_flags[r >> 5] >> (r & 31)) & 1u))
                data[r * n + idx] = lds[idx * LS + line];
        }
    }
}

// ---- column pass (n <= 512) ----------------------------------------------------
// Block (cb, a): columns [cb*CB, +CB) of array a (n x n, contiguous); arrays
// a >= *dev_batch (when given) are left alone.
template <typename T2, int LOGN, int CB, bool INV>
__global__ __launch_bounds__(CB *((1 << LOGN) / 8)) void k_fft_cols(T2 *data,
                                                                    const T2 *__restrict__ twg,
                                                                    const int32_t *dev_batch)
{
    if (dev_batch && (int)blockIdx.y >= *dev_batch)
        return;
    constexpr int n = 1 << LOGN;
    constexpr int NT = CB * (n / 8);
    __shared__ T2 lds[n * CB];
    __shared__ T2 tw[n];
    for (int k = threadIdx.x; k < n; k += NT)
        tw[k] = twg[k];
    T2 *a = data + (size_t)blockIdx.y * n * n + (size_t)blockIdx.x * CB;
    for (int e = threadIdx.x; e < n * CB; e += NT) {
        const int row = e / CB, col = e - row * CB;
        lds[e] = a[(size_t)row * n + col];
    }
    __syncthreads();
    fft_block<T2, LOGN, CB, CB, INV>(lds, tw);
    for (int e = threadIdx.x; e < n * CB; e += NT) {
        const int row = e / CB, col = e - row * CB;
        a[(size_t)row * n + col] = lds[e];
    }
}

// ---- transpose -------------------------------------------------------------------
template <typename T2>
__global__ __launch_bounds__(256) void k_transpose(const T2 *__restrict__ in, T2 *__restrict__ out,
                                                   int n, const int32_t *dev_batch)
{
    if (dev_batch && (int)blockIdx.z >= *dev_batch)
        return;
    __shared__ T2 tile[32][33];
    const size_t off = (size_t)blockIdx.z * n * n;
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    for (int j = ty; j < 32; j += 8)
        tile[j][tx] = in[off + (size_t)(y0 + j) * n + x0 + tx];
    __syncthreads();
    for (int j = ty; j < 32; j += 8)
        out[off + (size_t)(x0 + j) * n + y0 + tx] = tile[tx][j];
}

// ---- fft shift (fft_shift_kernel, common_kernels.cuh:17-32) ------------------------
template <typename T2>
__global__ __launch_bounds__(256) void k_fft_shift(T2 *data, int n, int batch)
{
    const int64_t half = (int64_t)n * (n / 2);
    const int64_t total = half * batch;
    const int64_t nn = (int64_t)n * n;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = t / half, e = t - b * half;
        const int64_t y = e / (n / 2), x = e - y * (n / 2);
        const int64_t ix0 = y * n + x;
        const int64_t ix1 = (ix0 + (int64_t)(n + 1) * (n / 2)) % nn;
        T2 *a = data + b * nn;
        const T2 tmp = a[ix0];
        a[ix0] = a[ix1];
        a[ix1] = tmp;
    }
}

int fft_shift(void *data, int type, int n, int batch, hipStream_t s)
{
    if (n <= 0 || (n & 1))
        return fail(RG_ERR_PARAM, "fft_shift needs an even size (got %d)", n);
    const int64_t total = (int64_t)n * (n / 2) * batch;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536)
        blocks = 65536;
    if (type == RG_C128)
        hipLaunchKernelGGL(k_fft_shift<double2>, dim3((unsigned)blocks), dim3(256), 0, s,
                           (double2 *)data, n, batch);
    else
        hipLaunchKernelGGL(k_fft_shift<float2>, dim3((unsigned)blocks), dim3(256), 0, s,
                           (float2 *)data, n, batch);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

// ---- twiddles ------------------------------------------------------------------------

int get_twiddles(rg_context *ctx, int n, int type, int sign, const void **out)
{
    TwKey key(n, type, sign);
    auto it = ctx->twiddles.find(key);
    if (it != ctx->twiddles.end()) {
        *out = it->second.p;
        return RG_OK;
    }
    DevBuf buf;
    const size_t esz = type == RG_C128 ? 16 : 8;
    RG_TRY(buf.ensure(esz * n));
    std::vector<double> hd(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        // exact quarter-wave symmetric values from one accurate evaluation
        const double a = 2.0 * M_PI * (double)k / (double)n;
        hd[2 * k] = cos(a);
        hd[2 * k + 1] = sign * sin(a);
    }
    if (type == RG_C128) {
        RG_HIP(hipMemcpy(buf.p, hd.data(), esz * n, hipMemcpyHostToDevice));
    } else {
        std::vector<float> hf(2 * (size_t)n);
        for (size_t i = 0; i < hf.size(); ++i)
            hf[i] = (float)hd[i];
        RG_HIP(hipMemcpy(buf.p, hf.data(), esz * n, hipMemcpyHostToDevice));
    }
    *out = buf.p;
    ctx->twiddles.emplace(key, buf);
    return RG_OK;
}

// Stage table of the Stockham sequence used by fft_band_rows (radix 8
// stages, then one radix-4 or radix-2 stage): entry Ns - 1 + (r - 1) Ns + k =
// tw[r k n / (Ns R)], the same values the plain table gives.
int get_stage_twiddles(rg_context *ctx, int n, int type, int sign, const void **out)
{
    TwKey key(n, type, 2 * sign);
    auto it = ctx->twiddles.find(key);
    if (it != ctx->twiddles.end()) {
        *out = it->second.p;
        return RG_OK;
    }
    int logn = 0;
    while ((1 << logn) < n)
        ++logn;
    std::vector<double> hd(2 * (size_t)n, 0.0);
    for (int logns = 0; logns < logn;) {
        const int rem = logn - logns;
        const int R = rem >= 3 ? 8 : (rem == 2 ? 4 : 2);
        const int Ns = 1 << logns;
        for (int r = 1; r < R; ++r)
            for (int k = 0; k < Ns; ++k) {
                const int m = (int)(((int64_t)r * k * (n / (Ns * R))) % n);
                const double a = 2.0 * M_PI * (double)m / (double)n;
                const size_t e = (size_t)(Ns - 1 + (r - 1) * Ns + k);
                hd[2 * e] = cos(a);
                hd[2 * e + 1] = sign * sin(a);
            }
        logns += R == 8 ? 3 : (R == 4 ? 2 : 1);
    }
    DevBuf buf;
    const size_t esz = type == RG_C128 ? 16 : 8;
    RG_TRY(buf.ensure(esz * n));
    if (type == RG_C128) {
        RG_HIP(hipMemcpy(buf.p, hd.data(), esz * n, hipMemcpyHostToDevice));
    } else {
        std::vector<float> hf(2 * (size_t)n);
        for (size_t i = 0; i < hf.size(); ++i)
            hf[i] = (float)hd[i];
        RG_HIP(hipMemcpy(buf.p, hf.data(), esz * n, hipMemcpyHostToDevice));
    }
    *out = buf.p;
    ctx->twiddles.emplace(key, buf);
    return RG_OK;
}

// ---- dispatch ----------------------------------------------------------------------

template <typename T2, int LOGN, bool INV>
static int rows_t(T2 *data, int64_t nrows, const T2 *tw, const uint32_t *flags,
                  const int32_t *dev_batch, hipStream_t s)
{
    constexpr int n = 1 << LOGN;
    constexpr int TPL = n / 8;
    constexpr int B = TPL >= 256 ? 1 : 256 / TPL;
    const int64_t blocks = (nrows + B - 1) / B;
    hipLaunchKernelGGL((k_fft_rows<T2, LOGN, B, INV>), dim3((unsigned)blocks), dim3(B * TPL), 0, s,
                       data, nrows, tw, flags, dev_batch);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

template <typename T2, int LOGN, bool INV>
static int cols_t(T2 *data, int batch, const T2 *tw, const int32_t *dev_batch, hipStream_t s)
{
    constexpr int n = 1 << LOGN;
    constexpr int CB = sizeof(T2) == 8 ? 16 : 8;
    constexpr int cb = CB > n ? n : CB;
    hipLaunchKernelGGL((k_fft_cols<T2, LOGN, cb, INV>), dim3(n / cb, batch), dim3(cb * (n / 8)), 0,
                       s, data, tw, dev_batch);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

template <typename T2, bool INV>
int fft_rows_dispatch(T2 *data, int n, int64_t nrows, const T2 *tw, const uint32_t *flags,
                      const int32_t *dev_batch, hipStream_t s)
{
    switch (n) {
    case 8: return rows_t<T2, 3, INV>(data, nrows, tw, flags, dev_batch, s);
    case 16: return rows_t<T2, 4, INV>(data, nrows, tw, flags, dev_batch, s);
    case 32: return rows_t<T2, 5, INV>(data, nrows, tw, flags, dev_batch, s);
    case 64: return rows_t<T2, 6, INV>(data, nrows, tw, flags, dev_batch, s);
    case 128: return rows_t<T2, 7, INV>(data, nrows, tw, flags, dev_batch, s);
    case 256: return rows_t<T2, 8, INV>(data, nrows, tw, flags, dev_batch, s);
    case 512: return rows_t<T2, 9, INV>(data, nrows, tw, flags, dev_batch, s);
    case 1024: return rows_t<T2, 10, INV>(data, nrows, tw, flags, dev_batch, s);
    case 2048: return rows_t<T2, 11, INV>(data, nrows, tw, flags, dev_batch, s);
    case 4096: return rows_t<T2, 12, INV>(data, nrows, tw, flags, dev_batch, s);
    default: return fail(RG_ERR_PARAM, "FFT size %d unsupported (power of two 8..4096)", n);
    }
}

template <typename T2, bool INV>
static int fft_cols_dispatch(T2 *data, int n, int batch, const T2 *tw, const int32_t *dev_batch,
                             hipStream_t s)
{
    switch (n) {
    case 8: return cols_t<T2, 3, INV>(data, batch, tw, dev_batch, s);
    case 16: return cols_t<T2, 4, INV>(data, batch, tw, dev_batch, s);
    case 32: return cols_t<T2, 5, INV>(data, batch, tw, dev_batch, s);
    case 64: return cols_t<T2, 6, INV>(data, batch, tw, dev_batch, s);
    case 128: return cols_t<T2, 7, INV>(data, batch, tw, dev_batch, s);
    case 256: return cols_t<T2, 8, INV>(data, batch, tw, dev_batch, s);
    case 512: return cols_t<T2, 9, INV>(data, batch, tw, dev_batch, s);
    default: return fail(RG_ERR_PARAM, "column pass size %d unsupported", n);
    }
}

template int fft_rows_dispatch<float2, true>(float2 *, int, int64_t, const float2 *,
                                             const uint32_t *, const int32_t *, hipStream_t);
template int fft_rows_dispatch<double2, true>(double2 *, int, int64_t, const double2 *,
                                             const uint32_t *, const int32_t *, hipStream_t);
template int fft_rows_dispatch<float2, false>(float2 *, int, int64_t, const float2 *,
                                             const uint32_t *, const int32_t *, hipStream_t);
template int fft_rows_dispatch<double2, false>(double2 *, int, int64_t, const double2 *,
                                             const uint32_t *, const int32_t *, hipStream_t);

template <typename T2>
static int fft2d_t(rg_context *ctx, T2 *data, int type, int n, int batch, int sign,
                   const uint32_t *row_flags, const int32_t *dev_batch, hipStream_t s)
{
    const void *twv = nullptr;
    RG_TRY(get_twiddles(ctx, n, type, sign, &twv));
    const T2 *tw = (const T2 *)twv;
    const bool inv = sign > 0;
    const int64_t nrows = (int64_t)n * batch;
    if (n <= 512) {
        if (inv) {
            RG_TRY((fft_rows_dispatch<T2, true>(data, n, nrows, tw, row_flags, dev_batch, s)));
            return fft_cols_dispatch<T2, true>(data, n, batch, tw, dev_batch, s);
        }
        RG_TRY((fft_rows_dispatch<T2, false>(data, n, nrows, tw, row_flags, dev_batch, s)));
        return fft_cols_dispatch<T2, false>(data, n, batch, tw, dev_batch, s);
    }
    // rows, transpose, rows, transpose back
    RG_TRY(ctx->scratch[11].ensure(sizeof(T2) * (size_t)n * n * batch));
    T2 *tmp = ctx->scratch[11].as<T2>();
    dim3 tg(n / 32, n / 32, batch);
    for (int pass = 0; pass < 2; ++pass) {
        const uint32_t *flags = pass == 0 ? row_flags : nullptr;
        if (inv)
            RG_TRY((fft_rows_dispatch<T2, true>(data, n, nrows, tw, flags, dev_batch, s)));
        else
            RG_TRY((fft_rows_dispatch<T2, false>(data, n, nrows, tw, flags, dev_batch, s)));
        hipLaunchKernelGGL(k_transpose<T2>, tg, dim3(256), 0, s, (const T2 *)data, tmp, n,
                           dev_batch);
        RG_HIP(hipGetLastError());
        std::swap(data, tmp);
    }
    // two transposes leave the result in the original array
    return RG_OK;
}

int fft_rows(rg_context *ctx, void *data, int type, int n, int64_t nrows, int sign, hipStream_t s)
{
    if (n < 8 || n > 4096 || (n & (n - 1)) || (sign != 1 && sign != -1))
        return fail(RG_ERR_PARAM, "FFT size %d / sign %d unsupported", n, sign);
    if (nrows <= 0)
        return RG_OK;
    const void *tw = nullptr;
    RG_TRY(get_twiddles(ctx, n, type, sign, &tw));
    if (type == RG_C128)
        return sign > 0 ? fft_rows_dispatch<double2, true>((double2 *)data, n, nrows,
                                                           (const double2 *)tw, nullptr, nullptr, s)
                        : fft_rows_dispatch<double2, false>((double2 *)data, n, nrows,
                                                            (const double2 *)tw, nullptr, nullptr,
                                                            s);
    return sign > 0 ? fft_rows_dispatch<float2, true>((float2 *)data, n, nrows, (const float2 *)tw,
                                                      nullptr, nullptr, s)
                    : fft_rows_dispatch<float2, false>((float2 *)data, n, nrows,
                                                       (const float2 *)tw, nullptr, nullptr, s);
}

int fft2d(rg_context *ctx, void *data, int type, int n, int batch, int sign, hipStream_t s,
          const uint32_t *row_flags, const int32_t *dev_batch)
{
    if (n < 8 || n > 4096 || (n & (n - 1)))
        return fail(RG_ERR_PARAM, "FFT size %d unsupported (power of two 8..4096)", n);
    if (batch < 1 || (sign != 1 && sign != -1))
        return fail(RG_ERR_PARAM, "invalid FFT batch/sign");
    if (type == RG_C128)
        return fft2d_t<double2>(ctx, (double2 *)data, type, n, batch, sign, row_flags, dev_batch,
                                s);
    return fft2d_t<float2>(ctx, (float2 *)data, type, n, batch, sign, row_flags, dev_batch, s);
}

} // namespace rg
