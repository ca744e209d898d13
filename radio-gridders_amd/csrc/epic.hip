// epic.hip -- EPIC E-field gridding: scatter_grid_kernel / epic_romein
// (src/epic/epic_grid.cu:54-218), batches of antenna E-fields each gridded
// with a small illumination pattern onto its own g x g grid.
//
// One workgroup per batch (timestamp).  When the grid fits 64 KiB of LDS
// (g <= 90) it is accumulated there with LDS float atomics, one thread per
// (antenna, tap) pair, and written to HBM once, coalesced: no global
// atomics, one pass of output traffic.  Larger grids add with global
// atomics into the zeroed output.
#include "rg_internal.hpp"

namespace rg {

template <bool LDS>
__global__ __launch_bounds__(256) void k_epic_grid(const float2 *__restrict__ f,
                                                   const int32_t *__restrict__ x,
                                                   const int32_t *__restrict__ y, int n_ant,
                                                   const float2 *__restrict__ illum, int supp,
                                                   int g, float2 *grids)
{
    extern __shared__ float2 sg[];
    const int b = blockIdx.x;
    const int gg = g * g;
    float2 *out = grids + (size_t)b * gg;
    if (LDS) {
        for (int e = threadIdx.x; e < gg; e += blockDim.x)
            sg[e] = make_float2(0.f, 0.f);
        __syncthreads();
    }
    const int taps = supp * supp;
    const int work = n_ant * taps;
    for (int e = threadIdx.x; e < work; e += blockDim.x) {
        const int a = e / taps, t = e - a * taps;
        const size_t i = (size_t)b * n_ant + a;
        const int X = x[i] + t % supp, Y = y[i] + t / supp;
        if (X < 0 || X >= g || Y < 0 || Y >= g) // scatter_grid_add's bounds check
            continue;
        const float2 k = illum[t], v = f[i];
        // conj(k) v (epic_grid.cu:121-123)
        const float re = k.x * v.x + k.y * v.y, im = k.x * v.y - k.y * v.x;
        float *c = reinterpret_cast<float *>(LDS ? sg : out) + 2 * (Y * g + X);
        atomicAdd(c, re);
        atomicAdd(c + 1, im);
    }
    if (LDS) {
        __syncthreads();
        for (int e = threadIdx.x; e < gg; e += blockDim.x)
            out[e] = sg[e];
    }
}

int epic_grid(const void *f, const int32_t *x, const int32_t *y, int n_ant, int nbatch,
              const void *illum, int supp, int g, void *grids, hipStream_t s)
{
    const size_t lds = sizeof(float2) * (size_t)g * g;
    if (lds <= 64 * 1024) {
        hipLaunchKernelGGL(k_epic_grid<true>, dim3((unsigned)nbatch), dim3(256), lds, s,
                           (const float2 *)f, x, y, n_ant, (const float2 *)illum, supp, g,
                           (float2 *)grids);
    } else {
        RG_TRY(zero_fill(grids, sizeof(float2) * (size_t)nbatch * g * g, s));
        hipLaunchKernelGGL(k_epic_grid<false>, dim3((unsigned)nbatch), dim3(256), 0, s,
                           (const float2 *)f, x, y, n_ant, (const float2 *)illum, supp, g,
                           (float2 *)grids);
    }
    RG_HIP(hipGetLastError());
    return RG_OK;
}

} // namespace rg
