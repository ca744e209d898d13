// fft_core.hpp -- LDS Stockham FFT of a block of lines (device code).
//
// A block holds B lines of length n = 2^LOGN in LDS, element (line, idx) at
// lds[idx * LS + line] with LS = B + PAD.  The block's threads are mapped
// line = tid % B, lt = tid / B (TPL = n / 8 threads per line), so
// neighbouring lanes touch neighbouring LDS words.  Every stage gives each
// thread 8 values: one radix-8, two radix-4 or four radix-2 butterflies
// (Stockham autosort, decimation in time:
//   X[(j/Ns)*Ns*R + j%Ns + r*Ns] = DFT_R over r' of x[j + r'*n/R] w^(r'*(j%Ns)),
//   w = exp(sign 2 pi i / (Ns R)) ).
// Twiddles come from an LDS table tw[k] = exp(sign 2 pi i k / n), k < n.
#pragma once

#include "binning.hpp"

namespace rg {

template <typename T2> __device__ inline T2 cmul(T2 a, T2 b)
{
    T2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename T2> __device__ inline T2 cadd(T2 a, T2 b)
{
    T2 r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}
template <typename T2> __device__ inline T2 csub(T2 a, T2 b)
{
    T2 r;
    r.x = a.x - b.x;
    r.y = a.y - b.y;
    return r;
}
// multiply by sign*i
template <bool INV, typename T2> __device__ inline T2 mul_si(T2 a)
{
    T2 r;
    if (INV) { // +i
        r.x = -a.y;
        r.y = a.x;
    } else { // -i
        r.x = a.y;
        r.y = -a.x;
    }
    return r;
}

// cos and sin of 2 pi t, t in turns (FP64).  t is split in FP64 into quarter
// turns q + r (q integer, |r| <= 1/2, exact), sin and cos of (pi/2) r are
// their Taylor series to degree 9 and 10 (truncation below 2e-9 for
// |r| <= 1/2) evaluated in FP32, and q mod 4 rotates the pair by q right
// angles.  Error about 1e-7, like an FP32 sincos of the reduced argument.
__device__ inline void turn_sincos_f32(double t, float &c, float &s)
{
    const double t4 = 4.0 * t;
    const double q = rint(t4);
    const float r = (float)(t4 - q);
    const int quad = (int)((long long)q & 3);
    const float r2 = r * r;
    // (pi/2)^k / k! with alternating signs
    const float sn = r * (1.5707963267948966f +
                          r2 * (-0.6459640975062462f +
                                r2 * (0.07969262624616703f +
                                      r2 * (-0.004681754135318687f +
                                            r2 * 0.00016044118478735975f))));
    const float cs = 1.0f + r2 * (-1.2337005501361697f +
                                  r2 * (0.253669507901048f +
                                        r2 * (-0.020863480763352957f +
                                              r2 * (0.0009192602748394263f +
                                                    r2 * -2.5202042373060596e-05f))));
    const bool odd = quad & 1, neg_c = quad == 1 || quad == 2, neg_s = quad >= 2;
    const float c0 = odd ? sn : cs, s0 = odd ? cs : sn;
    c = neg_c ? -c0 : c0;
    s = neg_s ? -s0 : s0;
}

template <bool INV, typename T2> __device__ inline void dft2(T2 &a, T2 &b)
{
    const T2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <bool INV, typename T2> __device__ inline void dft4(T2 &x0, T2 &x1, T2 &x2, T2 &x3)
{
    const T2 a0 = cadd(x0, x2), a1 = csub(x0, x2);
    const T2 a2 = cadd(x1, x3), a3 = mul_si<INV>(csub(x1, x3));
    x0 = cadd(a0, a2);
    x2 = csub(a0, a2);
    x1 = cadd(a1, a3);
    x3 = csub(a1, a3);
}

template <bool INV, typename T2> __device__ inline void dft8(T2 *x)
{
    using T = decltype(x[0].x);
    const T h = (T)0.70710678118654752440084436210484903928483593768847;
    // even / odd radix-4
    T2 e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    T2 o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    // o_k *= w8^k, w8 = (1 + s i)/sqrt2
    T2 t1, t3;
    if (INV) {
        t1.x = (o1.x - o1.y) * h;
        t1.y = (o1.y + o1.x) * h;
        t3.x = (-o3.x - o3.y) * h;
        t3.y = (-o3.y + o3.x) * h;
    } else {
        t1.x = (o1.x + o1.y) * h;
        t1.y = (o1.y - o1.x) * h;
        t3.x = (-o3.x + o3.y) * h;
        t3.y = (-o3.y - o3.x) * h;
    }
    const T2 t2 = mul_si<INV>(o2);
    x[0] = cadd(e0, o0);
    x[4] = csub(e0, o0);
    x[1] = cadd(e1, t1);
    x[5] = csub(e1, t1);
    x[2] = cadd(e2, t2);
    x[6] = csub(e2, t2);
    x[3] = cadd(e3, t3);
    x[7] = csub(e3, t3);
}

// radix of the stage with current sub-transform size Ns
template <int LOGN> __host__ __device__ constexpr int stage_radix(int logns)
{
    return (LOGN - logns) >= 3 ? 8 : ((LOGN - logns) == 2 ? 4 : 2);
}

// One Stockham stage with radix R on all lines of the block.
template <typename T2, int LOGN, int B, int LS, int R, bool INV>
__device__ inline void fft_stage(T2 *lds, const T2 *tw, int logns)
{
    constexpr int n = 1 << LOGN;
    constexpr int TPL = n / 8;
    constexpr int NBUT = n / R;
    constexpr int PER = NBUT / TPL; // butterflies per thread: 1, 2 or 4
    const int line = threadIdx.x % B;
    const int lt = threadIdx.x / B;
    const int Ns = 1 << logns;
    T2 v[8];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = lt + q * TPL;
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[q * R + r] = lds[(j + r * NBUT) * LS + line];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = lt + q * TPL;
        const int k = j & (Ns - 1);
        const int tstep = k * (n >> (logns + (R == 8 ? 3 : (R == 4 ? 2 : 1))));
        if (logns > 0) { // the first stage's twiddles are all 1
#pragma unroll
            for (int r = 1; r < R; ++r)
                v[q * R + r] = cmul(v[q * R + r], tw[r * tstep]);
        }
        if (R == 8)
            dft8<INV>(&v[q * R]);
        else if (R == 4)
            dft4<INV>(v[q * R], v[q * R + 1], v[q * R + 2], v[q * R + 3]);
        else
            dft2<INV>(v[q * R], v[q * R + 1]);
        const int d = (j >> logns) * (Ns * R) + k;
#pragma unroll
        for (int r = 0; r < R; ++r)
            lds[(d + r * Ns) * LS + line] = v[q * R + r];
    }
    __syncthreads();
}

// ---- block transform with register input and output --------------------------------
// Same stage sequence and thread mapping as fft_block, but the first stage
// takes its inputs from load(slot, idx) (no LDS store of the input, no barrier
// before it) and the last stage hands its outputs to emit(slot, idx, value)
// instead of storing them.  slot = q * R + r numbers a thread's values in a
// stage (compile-time after unrolling), idx is the position in the line
// (line = threadIdx.x % B).  The caller synchronises before the next use of
// lds (the last stage reads it).  Needs at least two stages (n >= 16).
// RMAX (8 or 4) is the radix of every stage but the last, which takes what
// is left of LOGN.
template <int LOGN, int RMAX = 8> __host__ __device__ constexpr int last_radix()
{
    return RMAX == 8 ? ((LOGN % 3) == 0 ? 8 : ((LOGN % 3) == 2 ? 4 : 2))
                     : ((LOGN % 2) == 0 ? 4 : 2);
}

// position in the line of the last stage's output `slot` for this thread
template <int LOGN, int B, int RMAX = 8> __device__ inline int fft_out_index(int slot)
{
    constexpr int n = 1 << LOGN;
    constexpr int R = last_radix<LOGN, RMAX>();
    const int lt = threadIdx.x / B;
    return lt + (slot / R) * (n / 8) + (slot % R) * (n / R);
}

template <typename T2, int LOGN, int B, int LS, int R, bool INV, int MODE, typename Load,
          typename Emit>
__device__ inline void fft_stage_io(T2 *lds, const T2 *tw, int logns, Load &load, Emit &emit)
{
    // MODE 1: inputs from load, outputs to lds.  MODE 2: inputs from lds,
    // outputs to emit.
    constexpr int n = 1 << LOGN;
    constexpr int TPL = n / 8;
    constexpr int NBUT = n / R;
    constexpr int PER = NBUT / TPL;
    const int line = threadIdx.x % B;
    const int lt = threadIdx.x / B;
    const int Ns = 1 << logns;
    T2 v[8];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = lt + q * TPL;
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[q * R + r] = MODE == 1 ? load(q * R + r, j + r * NBUT) : lds[(j + r * NBUT) * LS + line];
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = lt + q * TPL;
        const int k = j & (Ns - 1);
        const int tstep = k * (n >> (logns + (R == 8 ? 3 : (R == 4 ? 2 : 1))));
        if (logns > 0) {
#pragma unroll
            for (int r = 1; r < R; ++r)
                v[q * R + r] = cmul(v[q * R + r], tw[r * tstep]);
        }
        if (R == 8)
            dft8<INV>(&v[q * R]);
        else if (R == 4)
            dft4<INV>(v[q * R], v[q * R + 1], v[q * R + 2], v[q * R + 3]);
        else
            dft2<INV>(v[q * R], v[q * R + 1]);
        const int d = (j >> logns) * (Ns * R) + k;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (MODE == 1)
                lds[(d + r * Ns) * LS + line] = v[q * R + r];
            else
                emit(q * R + r, d + r * Ns, v[q * R + r]);
        }
    }
    if (MODE == 1)
        __syncthreads();
}

template <typename T2, int LOGN, int B, int LS, bool INV, int RMAX = 8, typename Load,
          typename Emit>
__device__ inline void fft_block_io(T2 *lds, const T2 *tw, Load &&load, Emit &&emit)
{
    static_assert(RMAX == 8 || RMAX == 4, "stage radix 8 or 4");
    constexpr int LR = RMAX == 8 ? 3 : 2;
    static_assert(LOGN > LR, "two or more stages");
    constexpr int NST = LOGN / LR + (LOGN % LR ? 1 : 0);
    fft_stage_io<T2, LOGN, B, LS, RMAX, INV, 1>(lds, tw, 0, load, emit);
    int logns = LR;
#pragma unroll
    for (int st = 1; st < NST - 1; ++st) {
        fft_stage<T2, LOGN, B, LS, RMAX, INV>(lds, tw, logns);
        logns += LR;
    }
    fft_stage_io<T2, LOGN, B, LS, last_radix<LOGN, RMAX>(), INV, 2>(lds, tw, logns, load, emit);
}

// ---- line-major variant ---------------------------------------------------------
// Lines stored contiguously (line l at base + l * LSTR); thread t of a round
// works on line t / TPL, position t % TPL.  `active` threads do the work, all
// threads reach the barriers.  A stage given a store functor hands it its
// results (position in the line, value) instead of writing them back to LDS.
struct NoStore {
    template <typename... A> __device__ void operator()(A &&...) const {}
};

template <typename T2, int LOGN, int R, bool INV, typename Store = NoStore>
__device__ inline void fft_stage_lm(T2 *line, const T2 *tw, int logns, int lt, bool active,
                                    const Store *out = nullptr)
{
    constexpr int n = 1 << LOGN;
    constexpr int TPL = n / 8;
    constexpr int NBUT = n / R;
    constexpr int PER = NBUT / TPL;
    const int Ns = 1 << logns;
    T2 v[8];
    if (active) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int j = lt + q * TPL;
#pragma unroll
            for (int r = 0; r < R; ++r)
                v[q * R + r] = line[j + r * NBUT];
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int j = lt + q * TPL;
            const int k = j & (Ns - 1);
            const int tstep = k * (n >> (logns + (R == 8 ? 3 : (R == 4 ? 2 : 1))));
            if (logns > 0) { // the first stage's twiddles are all 1
#pragma unroll
                for (int r = 1; r < R; ++r)
                    v[q * R + r] = cmul(v[q * R + r], tw[r * tstep]);
            }
            if (R == 8)
                dft8<INV>(&v[q * R]);
            else if (R == 4)
                dft4<INV>(v[q * R], v[q * R + 1], v[q * R + 2], v[q * R + 3]);
            else
                dft2<INV>(v[q * R], v[q * R + 1]);
            const int d = (j >> logns) * (Ns * R) + k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (out)
                    (*out)(d + r * Ns, v[q * R + r]);
                else
                    line[d + r * Ns] = v[q * R + r];
            }
        }
    }
    __syncthreads();
}

// Transforms nl contiguous lines of length n = 2^LOGN (line stride LSTR) with
// NT threads, B = NT / (n / 8) lines per round.  With `out`, the last stage
// calls out(line, position, value) instead of storing the result in LDS.
template <typename T2, int LOGN, int NT, int LSTR, bool INV, typename Out = NoStore>
__device__ inline void fft_lines_lm(T2 *lds, const T2 *tw, int nl, const Out *out = nullptr)
{
    constexpr int n = 1 << LOGN;
    constexpr int TPL = n / 8;
    static_assert(TPL <= NT, "line longer than 8 x threads");
    constexpr int B = NT / TPL;
    const int lt = threadIdx.x % TPL;
    for (int l0 = 0; l0 < nl; l0 += B) {
        const int l = l0 + (int)threadIdx.x / TPL;
        const bool active = (int)threadIdx.x < B * TPL && l < nl;
        T2 *line = lds + (size_t)(active ? l : 0) * LSTR;
        const int lo = active ? l : 0;
        auto st_line = [&](int pos, const T2 &val) { (*out)(lo, pos, val); };
        using SL = decltype(st_line);
        const SL *oline = out ? &st_line : nullptr;
        int logns = 0;
#pragma unroll
        for (int st = 0; st < 8; ++st) {
            if (logns >= LOGN)
                break;
            const int rem = LOGN - logns;
            if (rem >= 3) {
                fft_stage_lm<T2, LOGN, 8, INV>(line, tw, logns, lt, active,
                                               logns + 3 >= LOGN ? oline : nullptr);
                logns += 3;
            } else if (rem == 2) {
                fft_stage_lm<T2, LOGN, 4, INV>(line, tw, logns, lt, active,
                                               logns + 2 >= LOGN ? oline : nullptr);
                logns += 2;
            } else {
                fft_stage_lm<T2, LOGN, 2, INV>(line, tw, logns, lt, active, oline);
                logns += 1;
            }
        }
    }
}

// Full transform of the block (data already in lds, synchronised).
template <typename T2, int LOGN, int B, int LS, bool INV>
__device__ inline void fft_block(T2 *lds, const T2 *tw)
{
    int logns = 0;
#pragma unroll
    for (int st = 0; st < 8; ++st) {
        if (logns >= LOGN)
            break;
        const int rem = LOGN - logns;
        if (rem >= 3) {
            fft_stage<T2, LOGN, B, LS, 8, INV>(lds, tw, logns);
            logns += 3;
        } else if (rem == 2) {
            fft_stage<T2, LOGN, B, LS, 4, INV>(lds, tw, logns);
            logns += 2;
        } else {
            fft_stage<T2, LOGN, B, LS, 2, INV>(lds, tw, logns);
            logns += 1;
        }
    }
}

// ---- padded band rows ---------------------------------------------------------------
// Rows of a band stored with one pad element after every 16: logical index x
// at x + (x >> 4).  Row l starts at the (16-aligned) logical index rowbase_l,
// so its element e lies at pad(rowbase_l) + e + (e >> 4).  With the pads the
// strided stores of the first Stockham stages (lanes 8 or 64 elements apart)
// spread over all LDS banks.  Twiddles come from a stage table: stage Ns,
// radix R holds w^(r k n / (Ns R)) for r = 1..R-1, k < Ns at
// tws[Ns - 1 + (r - 1) Ns + k], so the lanes of a wavefront read consecutive
// words (a plain table read at r k conflicts for even r).
__host__ __device__ constexpr int pad16(int x) { return x + (x >> 4); }

// One stage: reads src (padded, or with SRC_PAD false the row's logical
// layout), writes row (padded); every read precedes every write (barrier).
template <typename T2, int LOGN, int R, bool INV, bool SRC_PAD = true>
__device__ inline void fft_stage_band(const T2 *src, T2 *row, const T2 *tws, int logns, int lt,
                                      bool active)
{
    constexpr int n = 1 << LOGN;
    constexpr int TPL = n / 8;
    constexpr int NBUT = n / R;
    constexpr int PER = NBUT / TPL;
    const int Ns = 1 << logns;
    T2 v[8];
    if (active) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int j = lt + q * TPL;
#pragma unroll
            for (int r = 0; r < R; ++r)
                v[q * R + r] = src[SRC_PAD ? pad16(j + r * NBUT) : j + r * NBUT];
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int j = lt + q * TPL;
            const int k = j & (Ns - 1);
            if (logns > 0) { // the first stage's twiddles are all 1
#pragma unroll
                for (int r = 1; r < R; ++r)
                    v[q * R + r] = cmul(v[q * R + r], tws[Ns - 1 + (r - 1) * Ns + k]);
            }
            if (R == 8)
                dft8<INV>(&v[q * R]);
            else if (R == 4)
                dft4<INV>(v[q * R], v[q * R + 1], v[q * R + 2], v[q * R + 3]);
            else
                dft2<INV>(v[q * R], v[q * R + 1]);
            const int d = (j >> logns) * (Ns * R) + k;
#pragma unroll
            for (int r = 0; r < R; ++r)
                row[pad16(d + r * Ns)] = v[q * R + r];
        }
    }
    __syncthreads();
}

// Transforms rows 0..nl-1 of a padded band with NT threads; row l begins at
// band + pad16(LEAD + l LS) (LEAD and LS multiples of 16).  UNPADDED_IN: the
// rows arrive in the logical layout (row l at band + LEAD + l LS) and the
// first stage moves them to the padded one.  Padded row l lies above logical
// row l and below padded row l + 1 (LS > n), so with the row groups taken
// last first a stage never overwrites logical rows still to be read.
template <typename T2, int LOGN, int NT, int LEAD, int LS, bool INV, bool UNPADDED_IN = false>
__device__ inline void fft_band_rows(T2 *band, const T2 *tws, int nl)
{
    static_assert(LEAD % 16 == 0 && LS % 16 == 0, "rows start on pad boundaries");
    constexpr int n = 1 << LOGN;
    constexpr int TPL = n / 8;
    static_assert(TPL <= NT, "line longer than 8 x threads");
    constexpr int B = NT / TPL;
    const int lt = threadIdx.x % TPL;
    const int ng = (nl + B - 1) / B;
    for (int g = 0; g < ng; ++g) {
        const int l0 = (UNPADDED_IN ? ng - 1 - g : g) * B;
        const int l = l0 + (int)threadIdx.x / TPL;
        const bool active = (int)threadIdx.x < B * TPL && l < nl;
        T2 *row = band + pad16(LEAD + (active ? l : 0) * LS);
        int logns = 0;
#pragma unroll
        for (int st = 0; st < 8; ++st) {
            if (logns >= LOGN)
                break;
            const int rem = LOGN - logns;
            if (UNPADDED_IN && st == 0) {
                static_assert(LOGN >= 3, "the first stage is radix 8");
                const T2 *src = band + LEAD + (active ? l : 0) * LS;
                fft_stage_band<T2, LOGN, 8, INV, false>(src, row, tws, logns, lt, active);
                logns += 3;
            } else if (rem >= 3) {
                fft_stage_band<T2, LOGN, 8, INV>(row, row, tws, logns, lt, active);
                logns += 3;
            } else if (rem == 2) {
                fft_stage_band<T2, LOGN, 4, INV>(row, row, tws, logns, lt, active);
                logns += 2;
            } else {
                fft_stage_band<T2, LOGN, 2, INV>(row, row, tws, logns, lt, active);
                logns += 1;
            }
        }
    }
}

} // namespace rg
