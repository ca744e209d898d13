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
// value per source: a few sources are transformed along their rows and then
// summed directly over the columns, more sources by one batched 2-D FFT.
//
// Degridding: visibilities counting-sorted by uv tile, one wavefront per
// visibility with lanes spread over the kernel's taps so that each load
// instruction reads whole stack rows, each wavefront walking a contiguous
// chunk of the sorted order (chunks grouped per XCD); shuffle reduction.
#include "binning_impl.hpp"

namespace rg {

// ---- visibility order ---------------------------------------------------------------------
// Degridding reads 128-byte stack rows at random addresses when visibilities
// come in random order (the L2 misses nearly every row).  Visibilities are
// therefore counting-sorted by (16 x 16-cell uv tile, w plane) of their
// centre tap (key = tile * w_planes + ws).  The
// sort's scatter pass also evaluates everything a visibility's taps depend on
// (ceil / floor of u / du, v / du, w / dw; deconvolve_visibility_,
// wstack_common.h:430-460) and stores it packed in a 16-byte record:
//   pos = visibility index, cell = us | vs << 16,
//   bits = ws | ovu << 16 | ovv << 32 | ovw << 48
// (us, vs, ws the stack cell of the centre tap, ov* the oversampling rows,
// all < 2^16).  Visibilities whose taps leave the stack are dropped here
// (counted; their output stays 0).
struct WsCell {
    int us, vs, ws, ovu, ovv, ovw;
};

struct alignas(16) WsRec {
    uint32_t pos, cell;
    unsigned long long bits;
};

struct WStackKeyer {
    const double *u, *v, *w;
    double du, dw;
    int g, og, wp, aa_h, aaw_h, ntx, ov_uv, ov_w;
    static constexpr int TILE_SHIFT = 4;

    __device__ bool cell(int64_t i, WsCell &c, bool with_ov = true) const
    {
#pragma clang fp contract(off)
        const double ui = u[i], vi = v[i], wi = w[i];
        const double cu = ceil(ui / du), cv = ceil(vi / du), cw = ceil(wi / dw);
        if (!(cu > -1e9 && cu < 1e9 && cv > -1e9 && cv < 1e9 && cw > -1e9 && cw < 1e9))
            return false;
        c.us = (int)cu + g;
        c.vs = (int)cv + g;
        c.ws = (int)cw + wp / 2;
        if (c.us - aa_h < 0 || c.us + aa_h > og || c.vs - aa_h < 0 || c.vs + aa_h > og ||
            c.ws - aaw_h < 0 || c.ws + aaw_h > wp)
            return false;
        if (!with_ov)
            return true;
        const double flu = ui - cu * du, flv = vi - cv * du, flw = wi - cw * dw;
        c.ovu = min((int)floor(fabs(flu) / du * ov_uv), ov_uv - 1);
        c.ovv = min((int)floor(fabs(flv) / du * ov_uv), ov_uv - 1);
        c.ovw = min((int)floor(fabs(flw) / dw * ov_w), ov_w - 1);
        return true;
    }

    __device__ static uint32_t key_of(int us, int vs, int ws, int ntx, int wp)
    {
        return (uint32_t)(((vs >> TILE_SHIFT) * ntx + (us >> TILE_SHIFT)) * wp + ws);
    }

    __device__ static uint32_t key_of(const WsRec &r, int ntx, int wp)
    {
        return key_of((int)(r.cell & 0xffffu), (int)(r.cell >> 16), (int)(r.bits & 0xffffu), ntx,
                      wp);
    }

    // the record of visibility i with cell c
    __device__ static WsRec pack(int64_t i, const WsCell &c)
    {
        WsRec r;
        r.pos = (uint32_t)i;
        r.cell = (uint32_t)c.us | ((uint32_t)c.vs << 16);
        r.bits = (unsigned long long)c.ws | ((unsigned long long)c.ovu << 16) |
                 ((unsigned long long)c.ovv << 32) | ((unsigned long long)c.ovw << 48);
        return r;
    }
};

// exclusive prefix of v over the 256 threads of the workgroup; total to all
__device__ inline uint32_t block_prefix256(uint32_t v, uint32_t &total, uint32_t *tmp)
{
    tmp[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t t = threadIdx.x >= (unsigned)off ? tmp[threadIdx.x - off] : 0u;
        __syncthreads();
        tmp[threadIdx.x] += t;
        __syncthreads();
    }
    total = tmp[255];
    const uint32_t incl = tmp[threadIdx.x];
    __syncthreads();
    return incl - v;
}

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
struct WsSky {
    int og, w_planes, g;
    double theta, lambda, du, dw;
    const double *pl, *pm, *flux;
    const double *corr_lm, *corr_n;
    int n_lm, n_n;
};

// Source i's position (xs, ys) in the planes, its value (with the output
// shift's (-1)^(xs + ys) modulation, og being even) and its screen phase per
// plane step; false when the source is skipped.
__device__ inline bool ws_source(const WsSky &k, int i, int &xs, int &ys, double &val, double &ph)
{
#pragma clang fp contract(off)
    const int g = k.g, og = k.og;
    const double flc = floor((k.pl[i] / k.theta + 0.5) * (double)g);
    const double fmc = floor((k.pm[i] / k.theta + 0.5) * (double)g);
    if (!(flc >= 0 && flc < g && fmc >= 0 && fmc < g)) // also NaN
        return false;
    const int lc = (int)flc, mc = (int)fmc;
    const double lq = (double)lc / k.lambda - k.theta / 2;
    const double mq = (double)mc / k.lambda - k.theta / 2;
    const double n = sqrt(1.0 - lq * lq - mq * mq) - 1.0;
    const double lm_step = 1.0 / (double)k.n_lm, n_step = 1.0 / (double)k.n_n;
    const double fu = floor((k.du * lq) / lm_step), fv = floor((k.du * mq) / lm_step);
    const double fw = floor((k.dw * n) / n_step);
    const bool ok = fabs(fu) < 1e9 && fabs(fv) < 1e9 && fabs(fw) < 1e9;
    const int aau = ok ? (int)fu + k.n_lm / 2 : -1, aav = ok ? (int)fv + k.n_lm / 2 : -1;
    const int aaw = ok ? (int)fw + k.n_n / 2 : -1;
    if (!ok || aau < 0 || aau >= k.n_lm || aav < 0 || aav >= k.n_lm || aaw < 0 || aaw >= k.n_n)
        return false;
    double a = 1.0;
    a *= k.corr_lm[aau];
    a *= k.corr_lm[aav];
    a *= k.corr_n[aaw];
    val = (k.flux ? k.flux[i] : 1.0) / a;
    // the screen at the source's pixel (x = lc, y = mc of the inner region)
    const double l = k.theta * ((double)lc - g / 2) / g;
    const double m = k.theta * ((double)mc - g / 2) / g;
    ph = k.dw * (1.0 - sqrt(1.0 - l * l - m * m));
    const int gd = (og - g) / 2;
    xs = (gd + lc + og / 2) % og;
    ys = (gd + mc + og / 2) % og;
    if ((xs + ys) & 1)
        val = -val;
    return true;
}

// value of source (val, ph) in plane k: val exp(2 pi i (t0 + k) ph)
__device__ inline double2 ws_plane_value(double val, double ph, int k, int w_planes)
{
    double s, c;
    sincospi(2.0 * (double)(k - w_planes / 2) * ph, &s, &c);
    return make_double2(val * c, val * s);
}

// Dense form: sources added into the zeroed stack; rows holding a source
// flagged (row k og + ys), so that the first FFT pass can skip the others.
__global__ __launch_bounds__(256) void k_ws_sources(WsSky sky, int npts, double2 *stack,
                                                    unsigned long long *skipped,
                                                    uint32_t *row_flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts)
        return;
    int xs, ys;
    double val, ph;
    if (!ws_source(sky, i, xs, ys, val, ph)) {
        atomicAdd(skipped, 1ull);
        return;
    }
    const int og = sky.og;
    for (int k = 0; k < sky.w_planes; ++k) {
        const double2 v = ws_plane_value(val, ph, k, sky.w_planes);
        double *dst = reinterpret_cast<double *>(stack + ((size_t)k * og + ys) * og + xs);
        atomicAdd(dst, v.x);
        atomicAdd(dst + 1, v.y);
        const uint32_t row = (uint32_t)k * (uint32_t)og + (uint32_t)ys;
        atomicOr(&row_flags[row >> 5], 1u << (row & 31));
    }
}

// ---- few sources: rows transformed in a side buffer, columns summed directly --------------
// A plane's input is nonzero only on the rows ys holding a source (the same
// rows in every plane).  With at most WS_SPARSE_ROWS sources (and og >= 8)
// those rows are numbered (slot[ys] = j, rows[j] = ys), filled into
// buf[k][j] and row transformed there; the column transform is then the
// direct sum
//   plane[k][y][x] = sum_j buf[k][j][x] w^(rows[j] y mod og),  w = exp(-2 pi i / og),
// written once over the whole stack (no zero fill, no transposes).
constexpr int WS_SPARSE_ROWS = 64;

__global__ __launch_bounds__(256) void k_ws_mark_rows(WsSky sky, int npts, uint32_t *row_used)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int xs, ys;
    double val, ph;
    if (i < npts && ws_source(sky, i, xs, ys, val, ph))
        row_used[ys] = 1u;
}

// one workgroup of 256: slot of every used row in order, the row list and its length
__global__ __launch_bounds__(256) void k_ws_row_slots(const uint32_t *row_used, int og,
                                                      int32_t *slot, int32_t *rows,
                                                      uint32_t *nrows)
{
    __shared__ uint32_t tmp[256];
    const int per = (og + 255) / 256, y0 = (int)threadIdx.x * per;
    uint32_t sum = 0;
    for (int j = 0; j < per; ++j)
        if (y0 + j < og)
            sum += row_used[y0 + j] ? 1u : 0u;
    uint32_t total;
    uint32_t run = block_prefix256(sum, total, tmp);
    for (int j = 0; j < per; ++j) {
        const int y = y0 + j;
        if (y < og) {
            if (row_used[y]) {
                slot[y] = (int32_t)run;
                rows[run] = y;
                ++run;
            } else {
                slot[y] = -1;
            }
        }
    }
    if (threadIdx.x == 0)
        *nrows = total;
}

__global__ __launch_bounds__(256) void k_ws_sources_rows(WsSky sky, int npts,
                                                         const int32_t *__restrict__ slot,
                                                         int maxr, double2 *buf,
                                                         unsigned long long *skipped)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts)
        return;
    int xs, ys;
    double val, ph;
    if (!ws_source(sky, i, xs, ys, val, ph)) {
        atomicAdd(skipped, 1ull);
        return;
    }
    const int og = sky.og, j = slot[ys];
    for (int k = 0; k < sky.w_planes; ++k) {
        const double2 v = ws_plane_value(val, ph, k, sky.w_planes);
        double *dst = reinterpret_cast<double *>(buf + ((size_t)k * maxr + j) * og + xs);
        atomicAdd(dst, v.x);
        atomicAdd(dst + 1, v.y);
    }
}

// grid (og * og / (256 WS_COL_ROWS), w_planes): thread -> rows y0 .. y0 +
// WS_COL_ROWS - 1 of column x of plane blockIdx.y; per source row the
// twiddle w^(row y) starts from the table at y0 and steps by w^row (at most
// WS_COL_ROWS - 1 FP64 roundings)
constexpr int WS_COL_ROWS = 8;

__global__ __launch_bounds__(256) void k_ws_sparse_cols(const double2 *__restrict__ buf,
                                                        const int32_t *__restrict__ rows,
                                                        const uint32_t *__restrict__ nrows_p,
                                                        int og, int maxr,
                                                        const double2 *__restrict__ tw,
                                                        double2 *stack)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)og * (og / WS_COL_ROWS))
        return;
    const int k = (int)blockIdx.y;
    const int yg = (int)(e / og), x = (int)(e - (int64_t)yg * og);
    const int y0 = yg * WS_COL_ROWS, mask = og - 1;
    const int nr = (int)*nrows_p;
    const double2 *b = buf + (size_t)k * maxr * og + x;
    double re[WS_COL_ROWS], im[WS_COL_ROWS];
#pragma unroll
    for (int i = 0; i < WS_COL_ROWS; ++i)
        re[i] = im[i] = 0.0;
    for (int j = 0; j < nr; ++j) {
        const double2 v = b[(size_t)j * og];
        const int r = rows[j];
        double2 t = tw[(r * y0) & mask];
        const double2 st = tw[r & mask];
#pragma unroll
        for (int i = 0; i < WS_COL_ROWS; ++i) {
            re[i] += v.x * t.x - v.y * t.y;
            im[i] += v.x * t.y + v.y * t.x;
            const double tx = t.x * st.x - t.y * st.y;
            t.y = t.x * st.y + t.y * st.x;
            t.x = tx;
        }
    }
    double2 *dst = stack + ((size_t)k * og + y0) * og + x;
#pragma unroll
    for (int i = 0; i < WS_COL_ROWS; ++i)
        dst[(size_t)i * og] = make_double2(re[i], im[i]);
}

// Sum over the 16 lanes of a DPP row (lanes 16 r .. 16 r + 15), returned in
// every lane of the row: rotations by 8 and 4 inside the row, then swaps of
// lanes 2 and 1 apart inside quads; register moves only (a shuffle through
// ds_bpermute would take an LDS instruction per 32-bit half).  All lanes of
// the row must be active.
__device__ inline double row16_sum(double x)
{
    auto step = [&](auto ctrl) {
        constexpr int c = decltype(ctrl)::value;
        const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), c, 0xf, 0xf, false);
        const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), c, 0xf, 0xf, false);
        x += __hiloint2double(hi, lo);
    };
    step(std::integral_constant<int, 0x128>{}); // row_ror:8
    step(std::integral_constant<int, 0x124>{}); // row_ror:4
    step(std::integral_constant<int, 0x4e>{});  // quad_perm [2, 3, 0, 1]
    step(std::integral_constant<int, 0xb1>{});  // quad_perm [1, 0, 3, 2]
    return x;
}

// ---- degridding ------------------------------------------------------------------------------
// deconvolve_visibility_ (wstack_common.h:420-470) over the tile-sorted
// records.  LV lanes per visibility: tap t = lane % aa_uv along u, pair
// q = lane / aa_uv + k (LV / aa_uv) over the aa_w * aa_uv (plane, row) pairs,
// so one load instruction reads whole kernel rows (contiguous runs of the
// stack).  Each lane sums stack[dws][dvs][dus] (cw cv) cu over its pairs; the
// group reduces with shuffles.  Every group walks one contiguous chunk of the
// sorted records; workgroup b runs on XCD b mod 8, so chunks are numbered to
// give each XCD one contiguous eighth of the sorted order and its L2 the
// neighbouring visibilities.
template <int LV>
__global__ __launch_bounds__(256) void k_ws_degrid(
    const double2 *__restrict__ stack, int og, const WsRec *__restrict__ recs,
    const uint32_t *__restrict__ nrec_p, int aa_uv, int aa_w,
    const double *__restrict__ conv_uv, const double *__restrict__ conv_w, double2 *out)
{
#pragma clang fp contract(off)
    constexpr int XCDS = 8;
    const int lane = (int)(threadIdx.x % LV);
    const int per_block = 256 / LV;
    const int nb = gridDim.x; // a multiple of XCDS
    const int vb = (blockIdx.x % XCDS) * (nb / XCDS) + blockIdx.x / XCDS;
    const int64_t group = (int64_t)vb * per_block + threadIdx.x / LV;
    const int64_t groups = (int64_t)nb * per_block;
    const int64_t nrec = *nrec_p;
    const int64_t chunk = (nrec + groups - 1) / groups;
    const int64_t r0 = min(nrec, group * chunk), r1 = min(nrec, r0 + chunk);
    const int aa_h = aa_uv / 2, aaw_h = aa_w / 2;
    const int t = lane % aa_uv, q0 = lane / aa_uv, qs = LV / aa_uv, npairs = aa_w * aa_uv;
    WsRec next{};
    if (r0 < r1)
        next = recs[r0];
    for (int64_t k = r0; k < r1; ++k) { // uniform across the group
        const WsRec rec = next;
        if (k + 1 < r1)
            next = recs[k + 1];
        const int us = (int)(rec.cell & 0xffffu), vs = (int)(rec.cell >> 16);
        const unsigned long long b = rec.bits;
        const int ws = (int)(b & 0xffffu), ovu = (int)((b >> 16) & 0xffffu);
        const int ovv = (int)((b >> 32) & 0xffffu), ovw = (int)(b >> 48);
        const double gu = conv_uv[aa_uv * ovu + t];
        const double2 *col = stack + (us - aa_h + t);
        double re = 0.0, im = 0.0;
        for (int q = q0; q < npairs; q += qs) {
            const int pw = q / aa_uv, pv = q - pw * aa_uv; // dwi + aaw_h, dvi + aa_h
            const double gw = conv_w[aa_w * ovw + pw];
            const double gv = gw * conv_uv[aa_uv * ovv + pv];
            const double2 sv = col[((size_t)(ws - aaw_h + pw) * og + (vs - aa_h + pv)) * og];
            const double kk = gv * gu;
            re += sv.x * kk;
            im += sv.y * kk;
        }
#pragma unroll
        for (int o = LV / 2; o > 0; o >>= 1) {
            re += __shfl_xor(re, o, LV);
            im += __shfl_xor(im, o, LV);
        }
        if (lane == 0)
            out[rec.pos] = make_double2(re, im);
    }
}

// Fixed 8 x 4 support with the stack staged in LDS: one work item (<= K
// records of one (uv tile, w plane) key) at a time per workgroup; the
// (16 + 8)^2 cells x 4 planes its taps can reach (36 KiB) are loaded once,
// then every record reads its 32 rows from LDS.  LV lanes per visibility,
// lane = (tap t, pair group h), each lane summing PAIRS / (LV / AUV)
// (plane, row) pairs fully unrolled; the 64 / LV visibilities of a wavefront
// share every instruction, the reduction included.
template <int AUV, int AW, int LV>
__global__ __launch_bounds__(256) void k_ws_degrid_tile(
    const double2 *__restrict__ stack, int og, int wp, int ntx,
    const WsRec *__restrict__ recs, const Item *__restrict__ items,
    const uint32_t *__restrict__ nitems_p, const double *__restrict__ conv_uv,
    const double *__restrict__ conv_w, double2 *out)
{
#pragma clang fp contract(off)
    constexpr int T = 1 << WStackKeyer::TILE_SHIFT, H = AUV / 2, R = T + AUV, P = AW;
    constexpr int GROUPS = LV / AUV, PAIRS = AUV * AW, PER_LANE = PAIRS / GROUPS;
    constexpr int GPB = 256 / LV; // groups per workgroup
    static_assert(LV % AUV == 0 && PAIRS % GROUPS == 0 && 64 % LV == 0, "lane layout");
    __shared__ double2 reg[P * R * R];
    const int lane = (int)(threadIdx.x % LV);
    const int t = lane % AUV, h = lane / AUV, grp = (int)(threadIdx.x / LV);
    const uint32_t nitems = *nitems_p;
    for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item item = items[it];
        const int ws = (int)(item.key % (uint32_t)wp);
        const int tile = (int)(item.key / (uint32_t)wp);
        const int col0 = (tile % ntx) * T - H, row0 = (tile / ntx) * T - H, pl0 = ws - AW / 2;
        __syncthreads(); // the previous item's records have been read
        for (int e = threadIdx.x; e < P * R * R; e += blockDim.x) {
            const int pp = e / (R * R), rr = (e / R) % R, cc = e % R;
            const int row = row0 + rr, col = col0 + cc;
            double2 val = make_double2(0.0, 0.0);
            if (row >= 0 && row < og && col >= 0 && col < og)
                val = stack[((size_t)(pl0 + pp) * og + row) * og + col];
            reg[e] = val;
        }
        __syncthreads();
        const int rounds = (int)((item.count + GPB - 1) / GPB);
        // two-deep pipeline: round rd sums record rd with the weights loaded
        // in round rd - 1, loads the weights of record rd + 1 (read in round
        // rd - 1) and reads record rd + 2
        // a lane's pairs are every plane pw with rows pv = h + b GROUPS (b < PV)
        constexpr int PV = AUV / GROUPS;
        static_assert(AUV % GROUPS == 0 && PER_LANE == AW * PV, "pair layout");
        struct Taps {
            uint32_t pos;
            int lr, lc; // local row / column of tap (-H, -H) + (0, t)
            double gu, gw[AW], gv[PV];
        };
        auto record = [&](int rd) {
            WsRec r{};
            const uint32_t k = (uint32_t)(rd * GPB + grp);
            if (rd < rounds && k < item.count)
                r = recs[item.start + k];
            return r;
        };
        auto weights = [&](const WsRec &rec) {
            Taps w;
            w.pos = rec.pos;
            const int us = (int)(rec.cell & 0xffffu), vs = (int)(rec.cell >> 16);
            const unsigned long long b = rec.bits;
            const int ovu = (int)((b >> 16) & 0xffffu);
            const int ovv = (int)((b >> 32) & 0xffffu), ovw = (int)(b >> 48);
            w.lr = vs - H - row0;
            w.lc = us - H - col0 + t;
            w.gu = conv_uv[AUV * ovu + t];
#pragma unroll
            for (int pw = 0; pw < AW; ++pw)
                w.gw[pw] = conv_w[AW * ovw + pw];
#pragma unroll
            for (int b = 0; b < PV; ++b)
                w.gv[b] = conv_uv[AUV * ovv + h + b * GROUPS];
            return w;
        };
        WsRec r1 = record(1);
        Taps cur = weights(record(0));
        for (int rd = 0; rd < rounds; ++rd) { // uniform across the workgroup
            const uint32_t k = (uint32_t)(rd * GPB + grp);
            const bool live = k < item.count;
            const WsRec r2 = record(rd + 2);
            const Taps nxt = weights(r1);
            double re = 0.0, im = 0.0;
            if (live) {
                // gu sum_pw gw[pw] sum_pv gv[pv] stack[pw][pv] (the weights factor)
                const double2 *base = reg + cur.lr * R + cur.lc + h * R;
#pragma unroll
                for (int pw = 0; pw < AW; ++pw) { // dwi + AW/2
                    double ir = 0.0, ii = 0.0;
#pragma unroll
                    for (int b = 0; b < PV; ++b) { // dvi + AUV/2 = h + b GROUPS
                        const double2 sv = base[(pw * R + b * GROUPS) * R];
                        ir = fma(sv.x, cur.gv[b], ir);
                        ii = fma(sv.y, cur.gv[b], ii);
                    }
                    re = fma(ir, cur.gw[pw], re);
                    im = fma(ii, cur.gw[pw], im);
                }
                re *= cur.gu;
                im *= cur.gu;
            }
            if constexpr (LV == 16) {
                re = row16_sum(re);
                im = row16_sum(im);
            } else {
#pragma unroll
                for (int o = LV / 2; o > 0; o >>= 1) {
                    re += __shfl_xor(re, o, LV);
                    im += __shfl_xor(im, o, LV);
                }
            }
            if (live && lane == 0)
                out[cur.pos] = make_double2(re, im);
            cur = nxt;
            r1 = r2;
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
    WsSky sky;
    sky.og = og;
    sky.w_planes = wp;
    sky.g = g;
    sky.theta = p->theta;
    sky.lambda = p->lambda;
    sky.du = p->du;
    sky.dw = p->dw;
    sky.pl = l;
    sky.pm = m;
    sky.flux = flux;
    sky.corr_lm = corr_lm->data;
    sky.corr_n = corr_n->data;
    sky.n_lm = corr_lm->size * corr_lm->oversampling;
    sky.n_n = corr_n->size * corr_n->oversampling;
    const unsigned sblocks = (unsigned)((npts + 255) / 256);
    if (npts <= WS_SPARSE_ROWS && og >= WS_COL_ROWS) {
        // rows used, their slots, the row buffer (at most npts rows per plane)
        const int maxr = npts > 0 ? npts : 1;
        RG_TRY(ctx->scratch[9].ensure(sizeof(uint32_t) * (3 * (size_t)og + 1)));
        uint32_t *used = ctx->scratch[9].as<uint32_t>();
        int32_t *slot = reinterpret_cast<int32_t *>(used + og), *rows = slot + og;
        uint32_t *nrows = reinterpret_cast<uint32_t *>(rows + og);
        const size_t bufn = (size_t)wp * maxr * og;
        RG_TRY(ctx->scratch[10].ensure(sizeof(double2) * bufn));
        double2 *buf = ctx->scratch[10].as<double2>();
        RG_TRY(zero_fill(used, sizeof(uint32_t) * og, s));
        RG_TRY(zero_fill(buf, sizeof(double2) * bufn, s));
        if (npts > 0)
            hipLaunchKernelGGL(k_ws_mark_rows, dim3(sblocks), dim3(256), 0, s, sky, npts, used);
        hipLaunchKernelGGL(k_ws_row_slots, dim3(1), dim3(256), 0, s, (const uint32_t *)used, og,
                           slot, rows, nrows);
        if (npts > 0)
            hipLaunchKernelGGL(k_ws_sources_rows, dim3(sblocks), dim3(256), 0, s, sky, npts,
                               (const int32_t *)slot, maxr, buf, cnt + CNT_SKY_SKIPPED);
        RG_HIP(hipGetLastError());
        RG_TRY(fft_rows(ctx, buf, RG_C128, og, (int64_t)wp * maxr, -1, s));
        const void *tw = nullptr;
        RG_TRY(get_twiddles(ctx, og, RG_C128, -1, &tw));
        hipLaunchKernelGGL(k_ws_sparse_cols,
                           dim3((unsigned)((plane / WS_COL_ROWS + 255) / 256), (unsigned)wp),
                           dim3(256),
                           0, s, (const double2 *)buf, (const int32_t *)rows,
                           (const uint32_t *)nrows, og, maxr, (const double2 *)tw, stack);
        RG_HIP(hipGetLastError());
    } else {
        RG_TRY(zero_fill(stack, sizeof(double2) * plane * wp, s));
        // rows holding a source (the first FFT pass skips the others)
        const size_t flag_words = ((size_t)og * wp + 31) / 32;
        RG_TRY(ctx->scratch[9].ensure(sizeof(uint32_t) * flag_words));
        uint32_t *row_flags = ctx->scratch[9].as<uint32_t>();
        RG_TRY(zero_fill(row_flags, sizeof(uint32_t) * flag_words, s));
        hipLaunchKernelGGL(k_ws_sources, dim3(sblocks), dim3(256), 0, s, sky, npts, stack,
                           cnt + CNT_SKY_SKIPPED, row_flags);
        RG_HIP(hipGetLastError());
        // the sources sit at their fft-shifted positions with the output
        // shift's modulation applied
        RG_TRY(fft2d(ctx, stack, RG_C128, og, wp, -1, s, row_flags));
    }
    ctx->ws_g = g;
    ctx->ws_og = og;
    ctx->ws_planes = wp;
    return RG_OK;
}

// ---- two-level sort of the visibilities --------------------------------------------------
// The device-wide counting sort spends one device-scope atomic per
// visibility when neighbouring visibilities do not share keys (random order),
// so degridding sorts in two levels with LDS counters only:
//   coarse count   : workgroup b counts its chunk of the stream per coarse
//                    bucket c = key >> fb (at most WS_NC buckets) in LDS; each
//                    visibility's record and its rank in its workgroup's share
//                    of c are written; hist[c][b] = the share
//   bucket scan    : hist[c][.] to exclusive prefixes over the workgroups,
//                    bucket totals to bucket starts
//   coarse scatter : record of visibility i to bstart[c] + hist[c][b] + rank
//   fine rank      : chunks of the bucketed records count their keys in LDS,
//                    reserve each key's share with one device atomic per
//                    (chunk, key) and give every record its rank in its key
//   items          : key starts and work items of at most K records from the
//                    per-key counts
//   fine scatter   : record to key start + rank
// Records of one key end up contiguous in key order, in no particular order
// inside the key (degridding a visibility does not depend on the others).
constexpr int WS_NC = 4096;       // coarse buckets at most
constexpr int WS_SORT_BLOCKS = 2048; // workgroups of the coarse passes at most
constexpr int WS_FB_MAX = 13;     // keys per bucket at most 2^13 (one fine window)

__global__ __launch_bounds__(256) void k_ws_coarse_count(WStackKeyer kf, int64_t n, int64_t chunk,
                                                         int nb, int fb, int nc, uint32_t *rank,
                                                         WsRec *rec, uint32_t *hist,
                                                         unsigned long long *stats)
{
    __shared__ uint32_t h[WS_NC];
    __shared__ uint32_t dropped;
    for (int c = threadIdx.x; c < nc; c += 256)
        h[c] = 0;
    if (threadIdx.x == 0)
        dropped = 0;
    __syncthreads();
    const int64_t i0 = (int64_t)blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    uint32_t nd = 0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
        WsCell cl;
        uint32_t rk = KEY_NONE;
        if (kf.cell(i, cl, true)) {
            const uint32_t key = WStackKeyer::key_of(cl.us, cl.vs, cl.ws, kf.ntx, kf.wp);
            rk = atomicAdd(&h[key >> fb], 1u);
            rec[i] = WStackKeyer::pack(i, cl);
        } else {
            ++nd;
        }
        rank[i] = rk;
    }
    if (nd)
        atomicAdd(&dropped, nd);
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += 256)
        hist[(size_t)c * nb + blockIdx.x] = h[c];
    if (threadIdx.x == 0 && dropped)
        atomicAdd(&stats[CNT_DROPPED], (unsigned long long)dropped);
}

// workgroup c: hist[c][0..nb) to its exclusive prefix, bucket total to tot[c]
__global__ __launch_bounds__(256) void k_ws_bucket_scan(uint32_t *hist, int nb, uint32_t *tot)
{
    __shared__ uint32_t tmp[256];
    uint32_t *hc = hist + (size_t)blockIdx.x * nb;
    const int per = (nb + 255) / 256, b0 = (int)threadIdx.x * per;
    uint32_t sum = 0;
    for (int j = 0; j < per; ++j)
        if (b0 + j < nb)
            sum += hc[b0 + j];
    uint32_t total;
    uint32_t run = block_prefix256(sum, total, tmp);
    for (int j = 0; j < per; ++j) {
        if (b0 + j < nb) {
            const uint32_t v = hc[b0 + j];
            hc[b0 + j] = run;
            run += v;
        }
    }
    if (threadIdx.x == 0)
        tot[blockIdx.x] = total;
}

// one workgroup: tot[0..nc) to bucket starts bstart[0..nc] (in place, nc + 1 entries)
__global__ __launch_bounds__(256) void k_ws_bucket_starts(uint32_t *bstart, int nc)
{
    __shared__ uint32_t tmp[256];
    const int per = (nc + 255) / 256, c0 = (int)threadIdx.x * per;
    uint32_t sum = 0;
    for (int j = 0; j < per; ++j)
        if (c0 + j < nc)
            sum += bstart[c0 + j];
    uint32_t total;
    uint32_t run = block_prefix256(sum, total, tmp);
    for (int j = 0; j < per; ++j) {
        if (c0 + j < nc) {
            const uint32_t v = bstart[c0 + j];
            bstart[c0 + j] = run;
            run += v;
        }
    }
    if (threadIdx.x == 0)
        bstart[nc] = total;
}

__global__ __launch_bounds__(256) void k_ws_coarse_scatter(int64_t n, int64_t chunk, int nb,
                                                           int fb, int ntx, int wp,
                                                           const uint32_t *__restrict__ rank,
                                                           const WsRec *__restrict__ rec,
                                                           const uint32_t *__restrict__ hist,
                                                           const uint32_t *__restrict__ bstart,
                                                           WsRec *tmp)
{
    const int64_t i0 = (int64_t)blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
        const uint32_t rk = rank[i];
        if (rk == KEY_NONE)
            continue;
        const WsRec r = rec[i];
        const uint32_t c = WStackKeyer::key_of(r, ntx, wp) >> fb;
        tmp[bstart[c] + hist[(size_t)c * nb + blockIdx.x] + rk] = r;
    }
}

// Fine level over the bucketed records in chunks of WS_FINE_CHUNK (a chunk
// may span buckets; buckets are key ranges in key order).  Chunk j counts
// its records per key in LDS (window of WS_FINE_WIN keys from its first
// bucket's base; a wider chunk, only possible when its buckets are sparse,
// counts with device atomics instead), reserves each key's share with one
// device atomic on cnt_key, and writes every record's rank inside its key.
// The number of records, total = bstart[nc], is read on the device.
constexpr int WS_FINE_CHUNK = 4096;
constexpr int WS_FINE_WIN = 8192;

__global__ __launch_bounds__(256) void k_ws_fine_rank(const WsRec *__restrict__ tmp,
                                                      const uint32_t *__restrict__ total_p,
                                                      int ntx, int wp, int fb, uint32_t *cnt_key,
                                                      uint32_t *rank)
{
    __shared__ uint32_t cnt[WS_FINE_WIN];
    const uint32_t total = *total_p;
    const uint32_t j0 = blockIdx.x * (uint32_t)WS_FINE_CHUNK;
    if (j0 >= total)
        return;
    const uint32_t j1 = min(total, j0 + (uint32_t)WS_FINE_CHUNK);
    auto key_of = [&](uint32_t j) { return WStackKeyer::key_of(tmp[j], ntx, wp); };
    const uint32_t kbase = (key_of(j0) >> fb) << fb;
    const uint64_t span = (uint64_t)((key_of(j1 - 1) >> fb) + 1) * (1u << fb) - kbase;
    if (span > (uint64_t)WS_FINE_WIN) {
        for (uint32_t j = j0 + threadIdx.x; j < j1; j += 256)
            rank[j] = atomicAdd(&cnt_key[key_of(j)], 1u);
        return;
    }
    const uint32_t S = (uint32_t)span;
    for (uint32_t f = threadIdx.x; f < S; f += 256)
        cnt[f] = 0;
    __syncthreads();
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += 256)
        rank[j] = atomicAdd(&cnt[key_of(j) - kbase], 1u); // rank inside the chunk's share
    __syncthreads();
    for (uint32_t f = threadIdx.x; f < S; f += 256) {
        const uint32_t c = cnt[f];
        if (c)
            cnt[f] = atomicAdd(&cnt_key[kbase + f], c); // the share's first rank
    }
    __syncthreads();
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += 256)
        rank[j] += cnt[key_of(j) - kbase];
}

// record j to key_start[key] + rank[j] (the key's records in no particular order)
__global__ __launch_bounds__(256) void k_ws_fine_scatter(const WsRec *__restrict__ tmp,
                                                         const uint32_t *__restrict__ total_p,
                                                         int ntx, int wp,
                                                         const uint32_t *__restrict__ key_start,
                                                         const uint32_t *__restrict__ rank,
                                                         WsRec *recs)
{
    const uint32_t total = *total_p;
    for (uint32_t j = blockIdx.x * 256u + threadIdx.x; j < total; j += gridDim.x * 256u) {
        const WsRec r = tmp[j];
        recs[key_start[WStackKeyer::key_of(r, ntx, wp)] + rank[j]] = r;
    }
}

static int ws_sort(rg_context *ctx, const WStackKeyer &kf, int64_t n, int64_t nkeys, uint32_t K,
                   hipStream_t s, BinResult &out)
{
    if (n >= (int64_t)0x7fffffff)
        return fail(RG_ERR_PARAM, "visibility stream too long for one call (%lld)", (long long)n);
    int fb = 0;
    while (((nkeys + (1ll << fb) - 1) >> fb) > WS_NC)
        ++fb;
    if (fb > WS_FB_MAX)
        return fail(RG_ERR_UNSUPPORTED, "W-stack key space too large (%lld)", (long long)nkeys);
    const int nc = (int)((nkeys + (1ll << fb) - 1) >> fb);
    // chunks of whole 256-visibility rounds, at most WS_SORT_BLOCKS of them
    int64_t chunk = (n + WS_SORT_BLOCKS - 1) / WS_SORT_BLOCKS;
    chunk = std::max<int64_t>(1024, (chunk + 255) / 256 * 256);
    const int nb = (int)((n + chunk - 1) / chunk);
    const int64_t max_items = nkeys + n / K + 1;

    RG_TRY(ensure_counters(ctx));
    RG_TRY(ctx->scratch[0].ensure(sizeof(uint32_t) * (size_t)nc * nb));      // hist
    RG_TRY(ctx->scratch[1].ensure(sizeof(uint32_t) * (nkeys + 1)));          // vis_start
    RG_TRY(ctx->scratch[2].ensure(sizeof(uint32_t) * (nkeys + 1)));          // item_start
    RG_TRY(ctx->scratch[3].ensure(sizeof(uint32_t) * (nc + 1)));             // bucket starts
    RG_TRY(ctx->scratch[4].ensure(sizeof(Item) * max_items));                // items
    RG_TRY(ctx->scratch[5].ensure(sizeof(WsRec) * n));                       // records
    RG_TRY(ctx->scratch[7].ensure(sizeof(WsRec) * n));                       // bucketed records
    RG_TRY(ctx->scratch[8].ensure(sizeof(uint32_t) * nkeys));                // per-key counts
    RG_TRY(ctx->scratch[15].ensure(sizeof(uint32_t) * n)); // bucket ranks, then key ranks
    KernelTimer kt(ctx, RG_TIMER_BINNING, s);
    uint32_t *hist = ctx->scratch[0].as<uint32_t>(), *bst = ctx->scratch[3].as<uint32_t>();
    auto *stats = ctx->counters.as<unsigned long long>();
    // the records in stream order go to the final record array (free until
    // the fine scatter), their bucket ranks to scratch 15
    uint32_t *rank = ctx->scratch[15].as<uint32_t>();
    WsRec *rec0 = ctx->scratch[5].as<WsRec>();
    hipLaunchKernelGGL(k_ws_coarse_count, dim3((unsigned)nb), dim3(256), 0, s, kf, n, chunk, nb,
                       fb, nc, rank, rec0, hist, stats);
    hipLaunchKernelGGL(k_ws_bucket_scan, dim3((unsigned)nc), dim3(256), 0, s, hist, nb, bst);
    hipLaunchKernelGGL(k_ws_bucket_starts, dim3(1), dim3(256), 0, s, bst, nc);
    hipLaunchKernelGGL(k_ws_coarse_scatter, dim3((unsigned)nb), dim3(256), 0, s, n, chunk, nb, fb,
                       kf.ntx, kf.wp, (const uint32_t *)rank, (const WsRec *)rec0,
                       (const uint32_t *)hist, (const uint32_t *)bst, ctx->scratch[7].as<WsRec>());
    uint32_t *cnt_key = ctx->scratch[8].as<uint32_t>();
    // bucket ranks are no longer read: the key ranks reuse their array
    const WsRec *tmp = ctx->scratch[7].as<WsRec>();
    RG_TRY(zero_fill(cnt_key, sizeof(uint32_t) * nkeys, s));
    hipLaunchKernelGGL(k_ws_fine_rank, dim3((unsigned)((n + WS_FINE_CHUNK - 1) / WS_FINE_CHUNK)),
                       dim3(256), 0, s, tmp, (const uint32_t *)(bst + nc), kf.ntx, kf.wp, fb,
                       cnt_key, rank);
    RG_HIP(hipGetLastError());
    RG_TRY(items_from_lengths(ctx, cnt_key, nkeys, K, ctx->scratch[1].as<uint32_t>(),
                              ctx->scratch[2].as<uint32_t>(), ctx->scratch[4].as<Item>(), s));
    hipLaunchKernelGGL(k_ws_fine_scatter,
                       dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65536)), dim3(256), 0, s,
                       tmp, (const uint32_t *)(bst + nc), kf.ntx, kf.wp,
                       (const uint32_t *)ctx->scratch[1].as<uint32_t>(), (const uint32_t *)rank,
                       ctx->scratch[5].as<WsRec>());
    RG_HIP(hipGetLastError());
    out.vis_start = ctx->scratch[1].as<uint32_t>();
    out.item_start = ctx->scratch[2].as<uint32_t>();
    out.items = ctx->scratch[4].as<Item>();
    out.recs = ctx->scratch[5].p;
    out.max_items = max_items;
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
    // per call: every counter but the build's skipped sources
    static_assert(CNT_SKY_SKIPPED == CNT__N - 1, "per-call counters precede the sky counter");
    RG_TRY(zero_fill(cnt, sizeof(unsigned long long) * CNT_SKY_SKIPPED, s));
    ctx->last_nvis = nvis;
    if (nvis == 0)
        return RG_OK;
    const double2 *stack = ctx->scratch[17].as<double2>();
    // sort by uv tile (dropped visibilities keep the zero written here)
    RG_TRY(zero_fill(out, sizeof(double2) * (size_t)nvis, s));
    WStackKeyer kf;
    kf.u = u;
    kf.v = v;
    kf.w = w;
    kf.du = p->du;
    kf.dw = p->dw;
    kf.g = g;
    kf.og = og;
    kf.wp = wp;
    kf.aa_h = p->aa_support_uv / 2;
    kf.aaw_h = p->aa_support_w / 2;
    kf.ntx = (og + (1 << WStackKeyer::TILE_SHIFT) - 1) >> WStackKeyer::TILE_SHIFT;
    kf.ov_uv = conv_uv->oversampling;
    kf.ov_w = conv_w->oversampling;
    if (og > 65535 || wp > 65535 || kf.ov_uv > 65535 || kf.ov_w > 65535)
        return fail(RG_ERR_UNSUPPORTED, "W-stack indices beyond 16 bits");
    BinResult br;
    const int64_t nkeys = (int64_t)kf.ntx * kf.ntx * wp;
    constexpr uint32_t K = 1024; // records per work item
    RG_TRY(ws_sort(ctx, kf, nvis, nkeys, K, s, br));
    const WsRec *recs = reinterpret_cast<const WsRec *>(br.recs);
    const uint32_t *nrec = br.vis_start + nkeys;
    // one wavefront per visibility (the uv support, a power of two <= 64,
    // divides it); a multiple of 8 workgroups (one share per XCD), at least
    // 16 records per wavefront
    auto nblocks = [&](int per_block) {
        int64_t blocks = std::min<int64_t>((nvis + 16 * per_block - 1) / (16 * per_block), 8192);
        return (blocks + 7) / 8 * 8;
    };
    if (p->aa_support_uv == 8 && p->aa_support_w == 4) {
        constexpr int LV = 16;
        // persistent workgroups over the work items (nitems read on the device)
        const int64_t blocks = std::min<int64_t>(br.max_items, 4096);
        hipLaunchKernelGGL((k_ws_degrid_tile<8, 4, LV>), dim3((unsigned)blocks), dim3(256), 0, s,
                           stack, og, wp, kf.ntx, recs, (const Item *)br.items,
                           (const uint32_t *)(br.item_start + nkeys), conv_uv->data, conv_w->data,
                           (double2 *)out);
    } else {
        constexpr int LV = 64;
        hipLaunchKernelGGL(k_ws_degrid<LV>, dim3((unsigned)nblocks(256 / LV)), dim3(256), 0, s,
                           stack, og, recs, nrec, p->aa_support_uv, p->aa_support_w, conv_uv->data,
                           conv_w->data, (double2 *)out);
    }
    RG_HIP(hipGetLastError());
    return RG_OK;
}

} // namespace rg
