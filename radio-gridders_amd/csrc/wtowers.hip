// wtowers.hip -- W-towers gridder (wtowers_CUDA_flat, grid_cuda.cu:522-746;
// CPU semantics grid_wtowers, src/reference/grid.c:295-555).
//
// A visibility belongs to tower (cx, cy) = flat chunk rule (radio.cuh:361-362)
// and floor wp = floor(w / winc + .5) (grid.c:453-455).  Each non-empty
// (tower, floor) task grids its visibilities into an n x n subgrid with
// offsets (u_mid, v_mid, w_mid = wp * winc), inverse-FFTs it and adds it to
// the tower's subimage times F^-wp, F = exp(2 pi i winc (1 - sqrt(1-l^2-m^2)))
// (the reference's fresnel_subimg_mul chain followed by the w = 0 transfer,
// common_kernels.cu:13-40, which multiply to exactly this factor).  Each
// tower's subimage is forward-FFT'd and folded / n^2 into the FP64 master grid
// (add_subs2main_kernel, common_kernels.cu:43-79).
//
// PLAIN mode keeps the reference's launch structure in FP64: per task a
// scatter_grid_kernel_flat-shaped Romein scatter <<<128, 128>>> with FP64
// atomics, a 2-D inverse FFT and a Fresnel multiply-add, one task at a time.
// BATCHED mode processes `floor_batch` tasks per round: tiled LDS Romein into
// their subgrids, one row-IFFT pass over all non-empty rows, and a column pass
// that inverse-transforms each column block, applies the w-screen and sums all
// floors of a tower in registers before one read-modify-write of the
// subimage.
#include <algorithm>

#include "binning_impl.hpp"
#include "fft_core.hpp"
#include "romein.hpp"
#include "tiles.hpp"

namespace rg {

// ---- binning -------------------------------------------------------------------------

// Tower ids: channel c, tower (cx, cy) -> (c * cc + cy) * cc + cx.  Task slots:
// tower id * nwp + (wp - wp_min), nwp the floor capacity of the call.
struct TowerKeyer {
    const double *u, *v, *w;
    const void *vis;
    int vis_type;
    int N, S, R, P;
    double theta, w_min, w_step;
    int n, chunk, cc;
    double winc;
    const int *range; // device: wp_min of the call (range[0])
    int nwp;          // floors per tower slot (capacity); farther floors overflow
    // band_h == 0: key = task.  Otherwise keys are task * nk + b for the
    // nk = n / band_h bands of band_h subgrid rows, and a visibility's key is
    // the first band its footprint rows [gy, gy + S) reach inside [0, n); the
    // footprint reaches at most mk consecutive bands.
    // pos is always the raw subgrid corner + S: (gx + S) | (gy + S) << 16.
    int band_h, nk, mk;
    const int64_t *chan_off; // channel c owns [chan_off[c], chan_off[c + 1])
    int nchan;

    __device__ int chan_of(int64_t i) const
    {
        int lo = 0, hi = nchan - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (chan_off[mid] <= i)
                lo = mid;
            else
                hi = mid - 1;
        }
        return lo;
    }

    __device__ int classify(int64_t i, uint32_t &key, uint32_t &pos, uint32_t &kidx) const
    {
        const double ui = u[i], vi = v[i], wi = w[i];
        const int cx = tower_index(ui, theta, N, chunk);
        const int cy = tower_index(vi, theta, N, chunk);
        if (cx < 0 || cx >= cc || cy < 0 || cy >= cc)
            return VIS_DROPPED;
        const int wp = floor_index(wi, winc);
        const int wp_min = range[0];
        if (wp < wp_min)
            return VIS_DROPPED;
        if (wp >= wp_min + nwp)
            return VIS_OVERFLOW;
        const int x_min = chunk * cx - N / 2;
        const int y_min = chunk * cy - N / 2;
        const double u_mid = (double)(x_min + chunk / 2) / theta;
        const double v_mid = (double)(y_min + chunk / 2) / theta;
        const double w_mid = (double)wp * winc;
        int gx, gy, xf, yf;
        frac_coord(n, S, R, theta, ui, vi, u_mid, v_mid, gx, gy, xf, yf);
        if (gx <= -S || gx >= n || gy <= -S || gy >= n)
            return VIS_DROPPED;
        const int p = w_plane(wi - w_mid, w_min, w_step);
        if (p < 0 || p >= P)
            return VIS_BAD_PLANE;
        const int c = nchan > 1 ? chan_of(i) : 0;
        const uint32_t task = (uint32_t)(((c * cc + cy) * cc + cx) * nwp + (wp - wp_min));
        key = band_h > 0 ? task * (uint32_t)nk + (uint32_t)(max(gy, 0) / band_h) : task;
        pos = (uint32_t)(gx + S) | ((uint32_t)(gy + S) << 16);
        kidx = (uint32_t)((p * R + yf) * R + xf);
        return VIS_OK;
    }

    template <typename T> __device__ typename Vec2<T>::type vis_value(int64_t i) const
    {
        typename Vec2<T>::type r;
        if (vis_type == RG_C128) {
            const double2 d = reinterpret_cast<const double2 *>(vis)[i];
            r.x = (T)d.x;
            r.y = (T)d.y;
        } else {
            const float2 f = reinterpret_cast<const float2 *>(vis)[i];
            r.x = (T)f.x;
            r.y = (T)f.y;
        }
        return r;
    }
};

// min / max floor index over visibilities that fall in the tower set; four
// visibilities per thread and pass (their loads issued together), wavefront
// then workgroup reduction into part[workgroup]; k_range_finish reduces the
// parts (atomics on the same two words serialise: ~10 ns each)
constexpr int RANGE_VPT = 4;
constexpr int RANGE_BLOCKS = 2048;
// the range (2 ints) and the workgroups' parts in one buffer
constexpr size_t RANGE_BYTES = 256 + sizeof(int2) * RANGE_BLOCKS;

__global__ __launch_bounds__(256) void k_floor_range(const double *u, const double *v,
                                                     const double *w, int64_t n, double theta,
                                                     int N, int chunk, int cc, double winc,
                                                     int2 *part)
{
    __shared__ int smin[4], smax[4];
    int lo = INT_MAX, hi = INT_MIN;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < n;
         i0 += RANGE_VPT * stride) {
        double uu[RANGE_VPT], vv[RANGE_VPT], ww[RANGE_VPT];
#pragma unroll
        for (int q = 0; q < RANGE_VPT; ++q) {
            const int64_t i = i0 + q * stride < n ? i0 + q * stride : i0;
            uu[q] = u[i];
            vv[q] = v[i];
            ww[q] = w[i];
        }
#pragma unroll
        for (int q = 0; q < RANGE_VPT; ++q) {
            const int cx = tower_index(uu[q], theta, N, chunk);
            const int cy = tower_index(vv[q], theta, N, chunk);
            if (cx < 0 || cx >= cc || cy < 0 || cy >= cc)
                continue;
            const int wp = floor_index(ww[q], winc);
            lo = min(lo, wp);
            hi = max(hi, wp);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o));
        hi = max(hi, __shfl_xor(hi, o));
    }
    const int wv = threadIdx.x / 64;
    if (__lane_id() == 0) {
        smin[wv] = lo;
        smax[wv] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)blockDim.x / 64; ++k) {
            lo = min(lo, smin[k]);
            hi = max(hi, smax[k]);
        }
        part[blockIdx.x] = make_int2(lo, hi);
    }
}

// range[0..1] = min / max over the nparts parts (INT_MAX / INT_MIN if none)
__global__ __launch_bounds__(1024) void k_range_finish(const int2 *part, int nparts, int *range)
{
    __shared__ int smin[16], smax[16];
    int lo = INT_MAX, hi = INT_MIN;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
        const int2 v = part[k];
        lo = min(lo, v.x);
        hi = max(hi, v.y);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o));
        hi = max(hi, __shfl_xor(hi, o));
    }
    if (__lane_id() == 0) {
        smin[threadIdx.x / 64] = lo;
        smax[threadIdx.x / 64] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)blockDim.x / 64; ++k) {
            lo = min(lo, smin[k]);
            hi = max(hi, smax[k]);
        }
        range[0] = lo;
        range[1] = hi;
    }
}

// per task slot: visibility count and first work item (plain mode)
__global__ __launch_bounds__(256) void k_task_info(const uint32_t *vis_start,
                                                   const uint32_t *item_start, int64_t nslots,
                                                   int ntiles, uint32_t *nvis, uint32_t *item0)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nslots)
        return;
    nvis[t] = vis_start[(t + 1) * ntiles] - vis_start[t * ntiles];
    item0[t] = item_start[t * ntiles];
}

// ---- task tables (batched mode, built on the device) ---------------------------------------
// Tasks are the non-empty slots in slot order (tower id, then floor), so the
// tasks of a tower are consecutive.  Each tower's run of tasks is cut into
// chunks of at most TASKS_PER_CHUNK (one column-pass partial image each);
// active towers are ranked in tower-id order (the subimage index).
constexpr int TASKS_PER_CHUNK = 32;

struct TaskTables {
    int32_t *tower_ntasks;   // [ntw]
    int32_t *tower_task_off; // [ntw]
    int32_t *tower_chunk_off; // [ntw]
    int32_t *rank_of_tower;  // [ntw], -1 inactive
    int32_t *tower_of_rank;  // [ntw]
    int32_t *rank_chunk0;    // [ntw] first chunk of rank r
    int32_t *rank_nchunk;    // [ntw]
    int32_t *task_slot;      // [ntasks_cap]
    int32_t *task_wp;        // [ntasks_cap]
    int32_t *task_rank;      // [ntasks_cap] tower rank of the task
    int32_t *chunk_first;    // [nchunks_cap]
    int32_t *chunk_count;    // [nchunks_cap]
    int32_t *rank_of_slot;   // [nslots] task index of a non-empty slot
    // totals[0] tasks, [1] towers, [2] chunks (each clipped to its capacity)
    int32_t *totals;
    int ntw, nwp, nk; // tower ids, floors per tower, keys per task
    int ntasks_cap, nchunks_cap, ntowers_cap;
};

__device__ inline bool slot_nonempty(const uint32_t *key_start, int64_t s, int nk)
{
    return key_start[(s + 1) * nk] > key_start[s * nk];
}

// block t: number of non-empty slots of tower t
__global__ __launch_bounds__(256) void k_tower_count(const uint32_t *__restrict__ key_start,
                                                     TaskTables tt)
{
    __shared__ int wsum[4];
    const int t = blockIdx.x;
    int c = 0;
    for (int f = threadIdx.x; f < tt.nwp; f += 256)
        c += slot_nonempty(key_start, (int64_t)t * tt.nwp + f, tt.nk) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1)
        c += __shfl_xor(c, o);
    if (__lane_id() == 0)
        wsum[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0)
        tt.tower_ntasks[t] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// one block: offsets of every tower's tasks and chunks, tower ranks, totals
__global__ __launch_bounds__(1024) void k_tower_scan(TaskTables tt, unsigned long long *stats)
{
    __shared__ int s_task[1024], s_chunk[1024], s_act[1024];
    int run_task = 0, run_chunk = 0, run_act = 0;
    for (int base = 0; base < tt.ntw; base += 1024) {
        const int t = base + (int)threadIdx.x;
        const int nt = t < tt.ntw ? tt.tower_ntasks[t] : 0;
        const int nc = (nt + TASKS_PER_CHUNK - 1) / TASKS_PER_CHUNK;
        const int act = nt > 0 ? 1 : 0;
        s_task[threadIdx.x] = nt;
        s_chunk[threadIdx.x] = nc;
        s_act[threadIdx.x] = act;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            int a = 0, c = 0, r = 0;
            if ((int)threadIdx.x >= o) {
                a = s_task[threadIdx.x - o];
                c = s_chunk[threadIdx.x - o];
                r = s_act[threadIdx.x - o];
            }
            __syncthreads();
            s_task[threadIdx.x] += a;
            s_chunk[threadIdx.x] += c;
            s_act[threadIdx.x] += r;
            __syncthreads();
        }
        if (t < tt.ntw) {
            const int toff = run_task + s_task[threadIdx.x] - nt;
            const int coff = run_chunk + s_chunk[threadIdx.x] - nc;
            const int rk = run_act + s_act[threadIdx.x] - act;
            // a tower ranked beyond the subimage capacity is left out (the
            // call reports an overrun)
            const bool keep = act && rk < tt.ntowers_cap;
            tt.tower_task_off[t] = toff;
            tt.tower_chunk_off[t] = coff;
            tt.rank_of_tower[t] = keep ? rk : -1;
            if (keep) {
                tt.tower_of_rank[rk] = t;
                tt.rank_chunk0[rk] = coff;
                tt.rank_nchunk[rk] = nc;
            }
        }
        run_task += s_task[1023];
        run_chunk += s_chunk[1023];
        run_act += s_act[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        stats[CNT_TASKS] = (unsigned long long)run_task;
        stats[CNT_TOWERS] = (unsigned long long)run_act;
        stats[CNT_CHUNKS] = (unsigned long long)run_chunk;
        const bool over = run_task > tt.ntasks_cap || run_chunk > tt.nchunks_cap ||
                          run_act > tt.ntowers_cap;
        if (over)
            atomicAdd(&stats[CNT_OVERFLOW], 1ull);
        tt.totals[0] = min(run_task, tt.ntasks_cap);
        tt.totals[1] = min(run_act, tt.ntowers_cap);
        tt.totals[2] = min(run_chunk, tt.nchunks_cap);
    }
}

// block t: the task entries and chunks of tower t (block scan over its slots)
__global__ __launch_bounds__(256) void k_tower_tasks(const uint32_t *__restrict__ key_start,
                                                     TaskTables tt,
                                                     const int *__restrict__ range)
{
    __shared__ int wsum[4];
    const int t = blockIdx.x;
    const int nt = tt.tower_ntasks[t];
    if (nt == 0)
        return;
    const int toff = tt.tower_task_off[t], rk = tt.rank_of_tower[t];
    const int wp_min = range[0];
    int run = 0;
    for (int base = 0; base < tt.nwp; base += 256) {
        const int f = base + (int)threadIdx.x;
        const int64_t s = (int64_t)t * tt.nwp + f;
        const bool ne = f < tt.nwp && slot_nonempty(key_start, s, tt.nk);
        const unsigned long long m = __ballot(ne);
        const int lane = __lane_id(), wv = threadIdx.x / 64;
        const int in_wave = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0)
            wsum[wv] = __popcll(m);
        __syncthreads();
        int before = run;
        for (int q = 0; q < wv; ++q)
            before += wsum[q];
        if (ne) {
            const int k = toff + before + in_wave;
            if (k < tt.ntasks_cap) {
                tt.task_slot[k] = (int32_t)s;
                tt.task_wp[k] = wp_min + f;
                tt.task_rank[k] = rk;
                tt.rank_of_slot[s] = k;
            }
        }
        run += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    const int coff = tt.tower_chunk_off[t];
    const int nc = (nt + TASKS_PER_CHUNK - 1) / TASKS_PER_CHUNK;
    for (int c = threadIdx.x; c < nc; c += 256) {
        if (coff + c < tt.nchunks_cap) {
            const int first = toff + c * TASKS_PER_CHUNK;
            tt.chunk_first[coff + c] = first;
            tt.chunk_count[coff + c] =
                max(0, min(min(TASKS_PER_CHUNK, nt - c * TASKS_PER_CHUNK), tt.ntasks_cap - first));
        }
    }
}

// Per task: bit b of word b / 32 set when band b has work items (the band was
// written); MW = NB / 32 words per task (rounded up).
__global__ __launch_bounds__(256) void k_band_masks(const uint32_t *__restrict__ item_start,
                                                    const int32_t *__restrict__ task_slot,
                                                    const int32_t *__restrict__ totals, int nb,
                                                    int mw, uint32_t *masks)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; // (task, word)
    const int k = (int)(e / mw), wd = (int)(e - (int64_t)k * mw);
    if (k >= totals[0])
        return;
    const uint32_t *f = item_start + (size_t)task_slot[k] * nb;
    uint32_t m = 0;
    for (int q = 0; q < 32 && wd * 32 + q < nb; ++q) {
        const int bnd = wd * 32 + q;
        m |= (f[bnd + 1] > f[bnd] ? 1u : 0u) << q;
    }
    masks[e] = m;
}

// Tasks [t0, t1) of batch b (Q tasks per batch); false when the batch is empty.
struct BatchRef {
    const int32_t *task_slot;
    const int32_t *totals;
    int Q;
};

__device__ inline bool batch_tasks(const BatchRef &br, int b, int &t0, int &t1)
{
    const int nt = br.totals[0];
    t0 = b * br.Q;
    if (t0 >= nt)
        return false;
    t1 = min(nt, t0 + br.Q);
    return true;
}

// ---- screen ------------------------------------------------------------------------------

// 1 - sqrt(1 - l^2 - m^2) at subimage pixel (x, y) in the fft-shifted layout
// (zero image position at (0, 0), grid.c:317-319).
__device__ inline double screen_base(int x, int y, int n, double theta)
{
    const int xs = x < n / 2 ? x : x - n;
    const int ys = y < n / 2 ? y : y - n;
    const double l = theta * (double)xs / n;
    const double m = theta * (double)ys / n;
    return 1 - sqrt(1 - l * l - m * m);
}

template <typename T> struct Trig;
template <> struct Trig<float> {
    __device__ static void rot(double rev, float &c, float &s) { turn_sincos_f32(rev, c, s); }
};
template <> struct Trig<double> {
    __device__ static void rot(double rev, double &c, double &s)
    {
        const double f = rev - rint(rev);
        sincospi(2.0 * f, &s, &c);
    }
};

// ---- subgrid layout in the batch buffer ---------------------------------------------------
// A task's n x n subgrid is stored in column blocks of GCB columns: element
// (row, col) at (col / GCB) * n * GCB + row * GCB + col % GCB, so the column
// pass reads each of its blocks as one contiguous run.
constexpr int GCB = 8;

__device__ inline size_t gcb_index(int row, int col, int n)
{
    return (size_t)(col / GCB) * n * GCB + (size_t)row * GCB + (col % GCB);
}

// ---- band records ------------------------------------------------------------------------
// A visibility as band b of its task sees it, with everything a band lane
// needs precomputed (z = sizeof(T2), the element size):
//   xy = gz | cy0 << 16 (FP64; FP32: xy = gz), gz = (gx + BAND_LEAD) z the
//        footprint's first column in the band's logical row layout, cy0 =
//        H b - gy (signed) the footprint row that band row 0 meets;
//   kb, byte offset of that footprint row's first tap:
//   - FP32 (padded table: H zero rows, then every S x S sub-kernel followed
//     by H zero rows, PS = S + H): kb = ((H + kidx PS + min(cy0, S)) S) z;
//     band row l reads row min(cy0, S) + l of the record's sub-kernel, in
//     [1 - H, S + H - 1]; rows outside [0, S) fall in the zero rows after the
//     sub-kernel or after the one before it (or in the leading rows);
//   - FP64 (the table as given; its padded copy would outgrow an XCD's L2 at
//     large supports): kb = (kidx S^2 + cy0 S) z, never dereferenced for a
//     lane whose row the footprint misses (that lane reads no tap).
template <typename T> constexpr bool padded_taps() { return sizeof(T) == 4; }

template <typename T> struct BandRec {
    uint32_t xy;
    int32_t kb;
    typename Vec2<T>::type vis;
};
static_assert(sizeof(BandRec<float>) == sizeof(Rec<float>) &&
                  sizeof(BandRec<double>) == sizeof(Rec<double>),
              "band records replace binned records one for one");


__host__ __device__ constexpr int padded_rows(int S, int H) { return S + H; }

// Band rows in LDS: logical index BAND_LEAD + l LS + c for row l, column c
// (LS = n + BAND_LEAD), so the BAND_LEAD >= S columns left of a row and
// right of it fall in gaps between rows.  The scatter works in this layout;
// the row IFFT's first stage reads it and writes the rows with one pad
// element per 16 (pad16, fft_core.hpp), the layout of its later stages and
// of the copy-out.
constexpr int BAND_LEAD = 16;

template <typename T>
__device__ inline BandRec<T> band_record(const Rec<T> &r, int S, int H, int Hb, int z)
{
    const int gx = (int)(r.pos & 0xffffu) - S, gy = (int)(r.pos >> 16) - S;
    const int cy0 = Hb - gy; // >= 1 - H for every record of band Hb / H's window
    BandRec<T> o;
    if constexpr (padded_taps<T>())
        o.xy = (uint32_t)((gx + BAND_LEAD) * z); // the band lanes need no cy0
    else
        o.xy = (uint32_t)((gx + BAND_LEAD) * z) | ((uint32_t)cy0 << 16);
    if constexpr (padded_taps<T>())
        o.kb = ((H + (int32_t)r.kidx * padded_rows(S, H) + min(cy0, S)) * S) * z;
    else
        o.kb = ((int32_t)r.kidx * S * S + cy0 * S) * z;
    o.vis = r.vis;
    return o;
}

// Kernel table for the band pass: H zero rows, then sub-kernel k (S x S) at
// row H + k PS followed by H zero rows (PS = S + H); (H + nsub PS) S taps.
template <typename T2>
__global__ __launch_bounds__(256) void k_pad_kernels(const T2 *__restrict__ src, T2 *dst,
                                                     int64_t nsub, int S, int H)
{
    const int PS = padded_rows(S, H);
    const int64_t total = ((int64_t)H + nsub * PS) * S;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * 256) {
        const int64_t prow = e / S - H; // row counted from the first sub-kernel
        const int col = (int)(e % S);
        T2 v;
        v.x = 0;
        v.y = 0;
        if (prow >= 0) {
            const int64_t k = prow / PS;
            const int row = (int)(prow - k * PS);
            if (row < S)
                v = src[(k * S + row) * S + col];
        }
        dst[e] = v;
    }
}


// Band records in and subgrid bands out are touched once: nontemporal loads
// and stores keep them from evicting the kernel table, which every band lane
// re-reads, from L2.
template <typename R> __device__ inline R load_once(const R *p)
{
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    static_assert(sizeof(R) % 16 == 0, "16-byte words");
    struct W {
        v4u w[sizeof(R) / 16];
    } q;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(R) / 16); ++i)
        q.w[i] = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(p) + i);
    return __builtin_bit_cast(R, q);
}
// elements a, b at d[0], d[1]
__device__ inline void store_once2(float2 *d, float2 a, float2 b)
{
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f o = {a.x, a.y, b.x, b.y};
    __builtin_nontemporal_store(o, reinterpret_cast<v4f *>(d));
}
__device__ inline void store_once2(double2 *d, double2 a, double2 b)
{
    typedef double v2d __attribute__((ext_vector_type(2)));
    const v2d o0 = {a.x, a.y}, o1 = {b.x, b.y};
    __builtin_nontemporal_store(o0, reinterpret_cast<v2d *>(d));
    __builtin_nontemporal_store(o1, reinterpret_cast<v2d *>(d + 1));
}

// ---- record order inside band items -------------------------------------------------------
// A band lane's cell depends only on the footprint column gx (its row is fixed),
// so records sorted by gx keep each lane on one cell for long runs and the
// Romein register sum rarely flushes.  Every band work item (window records
// src = wstart[key] + (start - prefix[key]) .. + count of the binned array) is
// counting-sorted by gx + S (bins < n + S <= SORT_BINS) into its own slot
// [start, start + count) of the expanded array as band records of its band
// b = key mod nb (items of <= 32 records by rank in one wavefront); the
// order inside a bin is arbitrary.  The band pass relies on every item being
// sorted.  Workgroups loop over the items.
constexpr int SORT_BINS = 1024;

template <typename T>
__global__ __launch_bounds__(256) void k_sort_items(const Rec<T> *__restrict__ in,
                                                    BandRec<T> *__restrict__ out,
                                                    const Item *__restrict__ items,
                                                    const unsigned long long *__restrict__ stats,
                                                    const uint32_t *__restrict__ prefix,
                                                    const uint32_t *__restrict__ wstart, int nbins,
                                                    int S, int H, int nb)
{
    __shared__ uint32_t bins[SORT_BINS];
    __shared__ uint32_t wsum[4];
    constexpr int z = (int)sizeof(typename Vec2<T>::type);
    const uint32_t nitems = (uint32_t)stats[CNT_ITEMS];
    for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item item = items[it];
        const Rec<T> *src = in + (item.start - prefix[item.key] + wstart[item.key]);
        BandRec<T> *dst = out + item.start;
        const int Hb = H * (int)(item.key % (uint32_t)nb);
        if (item.count <= 32) {
            // rank sort in the first wavefront: lane i counts the records
            // ordered before its own (smaller column, or equal and earlier)
            if (threadIdx.x < 64) {
                const bool have = threadIdx.x < item.count;
                Rec<T> r{};
                uint32_t key = 0xffffffffu; // after every record
                if (have) {
                    r = src[threadIdx.x];
                    key = min(r.pos & 0xffffu, (uint32_t)nbins - 1);
                }
                uint32_t rank = 0;
#pragma unroll
                for (int j = 0; j < 32; ++j) {
                    const uint32_t kj = (uint32_t)__shfl((int)key, j);
                    rank += (kj < key || (kj == key && j < (int)threadIdx.x)) ? 1u : 0u;
                }
                if (have)
                    dst[rank] = band_record(r, S, H, Hb, z);
            }
            continue;
        }
        __syncthreads(); // bins of the previous item are no longer read
        for (int b = threadIdx.x; b < SORT_BINS; b += 256)
            bins[b] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < item.count; i += 256)
            atomicAdd(&bins[min(src[i].pos & 0xffffu, (uint32_t)nbins - 1)], 1u);
        __syncthreads();
        // exclusive scan of the bins: 4 per thread, wavefront scans, wave totals
        uint32_t v[4], tot = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] = bins[threadIdx.x * 4 + q];
            tot += v[q];
        }
        const int lane = __lane_id(), wv = threadIdx.x / 64;
        uint32_t inc = tot;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)inc, o);
            if (lane >= o)
                inc += y;
        }
        if (lane == 63)
            wsum[wv] = inc;
        __syncthreads();
        uint32_t run = inc - tot;
        for (int w2 = 0; w2 < wv; ++w2)
            run += wsum[w2];
        __syncthreads(); // every thread has read wsum before the next item
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            bins[threadIdx.x * 4 + q] = run;
            run += v[q];
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < item.count; i += 256) {
            const Rec<T> r = src[i];
            const uint32_t d = atomicAdd(&bins[min(r.pos & 0xffffu, (uint32_t)nbins - 1)], 1u);
            dst[d] = band_record(r, S, H, Hb, z);
        }
    }
}

// ---- band windows ------------------------------------------------------------------------
// Band b of a task takes the records of keys b - mk + 1 .. b (every footprint
// that can reach rows [H b, H b + H)); they are contiguous in the binned
// record array.  Per band key: window length and start.
// Band windows with records are counted (the bands the pass writes).
// (grid-stride; the windows with records are counted per workgroup and added
// with one atomic: one per wavefront on the same word serialised)
__global__ __launch_bounds__(256) void k_window_len(const uint32_t *__restrict__ key_start,
                                                    int64_t nkeys, int nb, int mk, uint32_t *len,
                                                    uint32_t *wstart, unsigned long long *stats)
{
    __shared__ uint32_t s_any[4];
    uint32_t nany = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nkeys;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int b = (int)(k % nb);
        const int64_t k0 = k - min(b, mk - 1);
        const uint32_t w0 = key_start[k0];
        const uint32_t l = key_start[k + 1] - w0;
        len[k] = l;
        wstart[k] = w0;
        nany += l > 0 ? 1u : 0u;
    }
    for (int o = 32; o > 0; o >>= 1)
        nany += (uint32_t)__shfl_xor((int)nany, o);
    if (__lane_id() == 0)
        s_any[threadIdx.x / 64] = nany;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tot = s_any[0] + s_any[1] + s_any[2] + s_any[3];
        if (tot)
            atomicAdd(&stats[CNT_BANDS], (unsigned long long)tot);
    }
}

// ---- batched band pass -------------------------------------------------------------------
// One workgroup per work item: up to K band records of the window of band b of
// a task -- records of keys b - mk + 1 .. b, sorted by footprint column --
// accumulated into subgrid rows [H b, H b + H).  The workgroup is GRP groups
// of S x H lanes; the records are staged RS at a time and every stage is
// split into one contiguous run per group.  Lane (myU, myV) owns band row
// myV and the columns congruent to myU modulo S: a footprint covering that
// row has exactly one cell there.  The lane keeps the register sum of its
// current cell and adds it into the LDS band when the cell changes (lanes of
// different groups own the same cells: see flush).  Band rows lie
// LS = n + BAND_LEAD elements apart after a lead of BAND_LEAD, so a footprint
// column outside [0, n) lands in the gap between rows and is dropped without
// a test.  The band's rows are then inverse FFT'd in LDS.  A key of one item
// stores the band straight into the task's subgrid; a key split over several
// items stores each part in a spill slot and k_band_combine sums the parts in
// item order.
template <typename T, int S, int H, int LOGN, int GRP, bool COMP>
__global__ __launch_bounds__(S *H *GRP) void k_tower_band_items(
    const BandRec<T> *__restrict__ recs, const Item *__restrict__ items,
    const uint32_t *__restrict__ item_start, BatchRef bref, int batch,
    const int32_t *__restrict__ rank_of_slot,
    const typename Vec2<T>::type *__restrict__ kern, uint32_t kern_bytes,
    const typename Vec2<T>::type *__restrict__ twg, typename Vec2<T>::type *G,
    typename Vec2<T>::type *spill, uint32_t spill_cap, uint32_t *spill_ctr, uint32_t *item_spill,
    uint32_t *spill_item, unsigned long long *stats)
{
    using T2 = typename Vec2<T>::type;
    constexpr int n = 1 << LOGN;
    constexpr int NL = S * H;
    constexpr int NT = NL * GRP;
    constexpr int NB = n / H;
    constexpr int z = (int)sizeof(T2);
    // records staged RS at a time (RS / GRP per group); they share LDS with
    // the twiddles, which keeps the FP32 8-row band at four workgroups per CU
    constexpr int RS = 256;
    constexpr int RG = RS / GRP;
    static_assert(RS % GRP == 0 && RS <= NT, "record staging");
    // row l of the band at logical [BAND_LEAD + l LS, + n) (pad16 after the
    // first IFFT stage, so the buffer holds the padded extent); the
    // S - 1 columns either side of a row fall in the gaps
    static_assert(BAND_LEAD >= S, "gaps hold the footprint columns outside a row");
    constexpr int LS = n + BAND_LEAD;
    constexpr int BANDE = pad16(BAND_LEAD + H * LS);
    __shared__ __align__(16) T2 band[BANDE];
    // staged records during the scatter, the FFT's stage twiddles afterwards
    constexpr int REC_BYTES = RS * (int)sizeof(BandRec<T>), TW_BYTES = n * (int)sizeof(T2);
    __shared__ __align__(16) char s_buf[REC_BYTES > TW_BYTES ? REC_BYTES : TW_BYTES];
    BandRec<T> *s_rec = reinterpret_cast<BandRec<T> *>(s_buf);
    T2 *tw = reinterpret_cast<T2 *>(s_buf);
    __shared__ uint32_t s_slot;
    // the batch's tasks and the work items of their keys
    int t0, t1;
    if (!batch_tasks(bref, batch, t0, t1))
        return;
    const uint32_t ib = item_start[(uint32_t)bref.task_slot[t0] * (uint32_t)NB];
    const uint32_t ie = item_start[(uint32_t)(bref.task_slot[t1 - 1] + 1) * (uint32_t)NB];
    // one workgroup per item: the grid must cover the batch's items
    if (blockIdx.x == 0 && threadIdx.x == 0 && ie - ib > gridDim.x)
        atomicAdd(&stats[CNT_OVERFLOW], 1ull);
    const uint32_t it = ib + blockIdx.x;
    if (it >= ie)
        return;
    // this thread's share of the stage twiddles, stored once the scatter is done
    constexpr int TWPER = (n + NT - 1) / NT;
    T2 twr[TWPER];
#pragma unroll
    for (int q = 0; q < TWPER; ++q)
        if (threadIdx.x + q * NT < n)
            twr[q] = twg[threadIdx.x + q * NT];
    {
        const Item item = items[it];
        const uint32_t tslot = item.key / (uint32_t)NB;
        const int b = (int)(item.key - tslot * (uint32_t)NB);
        const int j = rank_of_slot[tslot] - t0;

        // U records at a time: U tap loads in flight, then the U updates in order.
        constexpr int U = 8;
        static_assert(RG % U == 0, "record batch must be a multiple of the unroll");
        // records of the stage at base taken by the first group: with two
        // groups, half of them (a multiple of U), so that a short stage
        // keeps both groups busy
        auto split = [&](uint32_t base) {
            const int nl = (int)min((uint32_t)RS, item.count - base);
            return GRP == 2 ? min(RG, ((nl + 1) / 2 + U - 1) / U * U) : RG;
        };
        // thread t < RG stages record t of each RS-record stage if t < split,
        // thread RG + u record split + u; the next stage's record is loaded
        // while the current stage is processed
        auto stage_rec = [&](uint32_t base) {
            BandRec<T> r;
            const int t = (int)threadIdx.x, h = split(base);
            const uint32_t nl = min((uint32_t)RS, item.count - base);
            const uint32_t i = GRP != 2 ? (uint32_t)t : t < RG ? (t < h ? (uint32_t)t : nl)
                                                            : (uint32_t)(h + t - RG);
            if (i < nl) {
                r = load_once(recs + item.start + base + i);
            } else {
                // padding, no visibility: padded table: column n, after every
                // footprint column (its cells lie in the row gaps); otherwise a
                // row out of reach of every band row (no tap, cell kept)
                r.xy = padded_taps<T>() ? (uint32_t)((n + BAND_LEAD) * z) : 0x40000000u;
                r.kb = 0;
                r.vis.x = T(0);
                r.vis.y = T(0);
            }
            return r;
        };
        BandRec<T> next;
        if (threadIdx.x < RS)
            next = stage_rec(0);
        {
            // zero the band, 16 bytes per store
            float4 *b4 = reinterpret_cast<float4 *>(band);
            constexpr int N4 = (BANDE * (int)sizeof(T2)) / 16;
            for (int e = threadIdx.x; e < N4; e += NT)
                b4[e] = make_float4(0.f, 0.f, 0.f, 0.f);
            if ((BANDE * (int)sizeof(T2)) % 16 != 0 && threadIdx.x == 0) {
                band[BANDE - 1].x = T(0);
                band[BANDE - 1].y = T(0);
            }
        }
        const __amdgpu_buffer_rsrc_t ktab =
            __builtin_amdgcn_make_buffer_rsrc((void *)kern, 0, (int)kern_bytes, BUF_RSRC_W3);
        const int lane = threadIdx.x % NL, grp = threadIdx.x / NL;
        const int myU = lane % S, myV = lane / S;
        // per lane, in bytes: column residue, tap row, cell row
        const int uz = myU * z, tap_row = myV * S * z, cell_row = myV * LS * z;
        char *band_b = reinterpret_cast<char *>(band);
        // cells are byte offsets in the logical (unpadded) layout
        int cur = -1, carry = -1; // carry: the kept cell of the unpadded form
        T sr = T(0), si = T(0);
        // COMP: two-sum (Knuth) compensation of the register sums; cr, ci hold
        // the rounding errors, folded into sr, si before a sum leaves the lane
        T cr = T(0), ci = T(0);
        auto two_sum = [](T &s, T &c, T x) {
            const T t = s + x;
            const T bp = t - s;
            c += (s - (t - bp)) + (x - bp);
            s = t;
        };
        auto mac = [&](const T2 &k, const T2 &v) { // sr + i si += conj(k) v
            if constexpr (COMP) {
                two_sum(sr, cr, fma(k.x, v.x, k.y * v.y));
                two_sum(si, ci, fma(k.x, v.y, -(k.y * v.x)));
            } else {
                sr = fma(k.x, v.x, sr);
                sr = fma(k.y, v.y, sr);
                si = fma(k.x, v.y, si);
                si = fma(-k.y, v.x, si);
            }
        };
        auto fold = [&]() {
            if constexpr (COMP) {
                sr += cr;
                si += ci;
                cr = T(0);
                ci = T(0);
            }
        };
        // The groups take contiguous, column-sorted runs of each record stage,
        // so the cells of two groups can meet only within one footprint width
        // of the column where their runs meet.  With two groups the second
        // one walks its run backwards: both then reach that column last, so
        // during the stage their cells never meet and every flush is a plain
        // LDS read-modify-write; at the end of the stage the first group
        // flushes its last cells, and after a barrier the second.  (LDS float
        // atomics cost about as much as the rest of the scatter.)  With more
        // groups a flush is plain in the group's own column range
        // [safe_a, safe_b) and an LDS atomic add outside it.
        constexpr bool REV = GRP == 2;
        int safe_a = INT_MIN, safe_b = INT_MAX;
        // a cell of the row's gap: flushes into it are discarded
        const int gap_cell = cell_row + (n + BAND_LEAD) * z;
        auto flush = [&]() {
            fold();
            T *c = reinterpret_cast<T *>(band_b + cur);
            const int col = cur - cell_row;
            if (!REV && GRP > 1 && (col < safe_a || col >= safe_b)) {
                atomicAdd(c, sr);
                atomicAdd(c + 1, si);
            } else {
                T2 *c2 = reinterpret_cast<T2 *>(c);
                T2 o = *c2;
                o.x += sr;
                o.y += si;
                *c2 = o;
            }
        };
        // with two groups a lane reads a cell's LDS value when it starts
        // summing into it (no other lane writes that cell before the lane's
        // flush), so the flush is a store
        T2 pre;
        pre.x = T(0);
        pre.y = T(0);
        auto cellp = [&](int cl) {
            return reinterpret_cast<T2 *>(band_b + cl);
        };
        if (REV) {
            cur = gap_cell;
            carry = gap_cell;
        }
        for (uint32_t base = 0; base < item.count; base += RS) {
            const uint32_t nloc = min((uint32_t)RS, item.count - base);
            __syncthreads();
            if (threadIdx.x < RS) {
                // the second group's half backwards (its walk descends)
                const int t = (int)threadIdx.x;
                s_rec[REV && t >= RG ? 3 * RG - 1 - t : t] = next;
                if (base + RS < item.count)
                    next = stage_rec(base + RS);
            }
            __syncthreads();
            const int h = split(base);
            const int ng = REV ? (grp == 0 ? h : max(0, (int)nloc - h))
                               : min(RG, max(0, (int)nloc - grp * RG));
            const int nbu = (ng + U - 1) / U * U; // records walked, padding included
            const BandRec<T> *gr = s_rec + grp * RG;
            if (REV && grp == 1)
                gr += RG - nbu; // this group's half is stored in reverse
            if (GRP > 2) {
                // columns reached by the previous group's last record, and the
                // next group's first footprint column
                safe_a = grp > 0 ? (int)(s_rec[grp * RG - 1].xy & 0xffffu) + S * z : INT_MIN;
                safe_b = grp < GRP - 1 ? (int)(s_rec[(grp + 1) * RG].xy & 0xffffu) : INT_MAX;
            }
            using Bt = RecBatch<T, U>;
            // A lane whose row the footprint misses reads a zero tap: from the
            // padding of the record's sub-kernel (its cell then follows the
            // footprint column like every other lane's, so it changes cell with
            // them and flushes only when it has summed something), or, unpadded,
            // through an offset past the table, keeping its previous cell.
            auto fetch = [&](uint32_t q0, Bt &bt) {
    #pragma unroll
                for (int q = 0; q < U; ++q) {
                    const BandRec<T> r = gr[q0 + q];
                    const int gz = padded_taps<T>() ? (int)r.xy : (int)(r.xy & 0xffffu);
                    const int cxz = (uz - gz) & ((S - 1) * z);
                    bt.v[q] = r.vis;
                    if constexpr (padded_taps<T>()) {
                        bt.cell[q] = cell_row + gz + cxz;
                        bt.k[q] = tap_load(ktab, (uint32_t)(r.kb + tap_row + cxz), (T2 *)nullptr);
                    } else {
                        const int cy = ((int)r.xy >> 16) + myV;
                        const bool in = (unsigned)cy < (unsigned)S;
                        carry = in ? cell_row + gz + cxz : carry;
                        bt.cell[q] = carry;
                        bt.k[q] = tap_load(ktab, in ? (uint32_t)(r.kb + tap_row + cxz) : TAP_NONE,
                                           (T2 *)nullptr);
                    }
                }
            };
            // A lane's cell never decreases along the records (they are sorted by
            // footprint column, and the cell is the first column >= gx of the
            // lane's residue), so no lane changes cell in a batch iff every lane's
            // last cell is its current one: then the batch is plain FMAs.
            auto update = [&](const Bt &bt) {
                if (__all(bt.cell[U - 1] == cur)) {
    #pragma unroll
                    for (int q = 0; q < U; ++q)
                        mac(bt.k[q], bt.v[q]);
                    return;
                }
    #pragma unroll
                for (int q = 0; q < U; ++q) {
                    if (bt.cell[q] != cur) {
                        fold();
                        if (REV) {
                            T2 o = pre;
                            o.x += sr;
                            o.y += si;
                            *cellp(cur) = o;
                            cur = bt.cell[q];
                            pre = *cellp(cur);
                        } else {
                            if (cur >= 0)
                                flush();
                            cur = bt.cell[q];
                        }
                        sr = T(0);
                        si = T(0);
                    }
                    mac(bt.k[q], bt.v[q]);
                }
            };
            pipelined_records<U, Bt>((uint32_t)nbu, fetch, update);
            if (REV) {
                // last cells: the first group's, then (after a barrier) the
                // second's, which may be the same cells
                if (grp == 0)
                    flush();
                __syncthreads();
                if (grp == 1)
                    flush();
                cur = gap_cell;
                carry = gap_cell;
                sr = T(0);
                si = T(0);
            } else if (GRP > 1) {
                // the next stage's runs start elsewhere: nothing stays in flight
                fold();
                if (cur >= 0 && (sr != T(0) || si != T(0)))
                    flush();
                cur = -1;
                carry = -1;
                sr = T(0);
                si = T(0);
            }
        }
        fold();
        if (GRP == 1 && cur >= 0 && (sr != T(0) || si != T(0)))
            flush();

        // destination: the task's subgrid band (column blocks), or a spill slot
        // (H rows of n) for a split key
        const uint32_t k0 = item_start[item.key], k1 = item_start[item.key + 1];
        T2 *dst;
        const bool whole = k1 - k0 == 1;
        if (whole) {
            dst = G + (size_t)j * n * n;
        } else {
            if (threadIdx.x == 0) {
                const uint32_t sl = atomicAdd(spill_ctr, 1u);
                item_spill[it - ib] = sl;
                if (sl < spill_cap)
                    spill_item[sl] = it - ib;
                s_slot = sl;
            }
            __syncthreads();
            // the capacity bounds every batch (at most 2 R / K parts); an overrun
            // would lose this part, so it is counted as an error of the call
            if (s_slot >= spill_cap) {
                if (threadIdx.x == 0)
                    atomicAdd(&stats[CNT_OVERFLOW], 1ull);
                return;
            }
            dst = spill + (size_t)s_slot * H * n;
        }
        __syncthreads(); // records no longer read: their space takes the twiddles
#pragma unroll
        for (int q = 0; q < TWPER; ++q)
            if (threadIdx.x + q * NT < n)
                tw[threadIdx.x + q * NT] = twr[q];
        __syncthreads();
        fft_band_rows<T2, LOGN, NT, BAND_LEAD, LS, true, true>(band, tw, H);
        // element (l, c) of the band
        auto at = [&](int l, int c) { return band[pad16(BAND_LEAD + l * LS) + pad16(c)]; };
        // two elements per thread and store (columns c, c + 1 with c even are
        // adjacent in LDS and in both destination layouts)
        auto put2 = [&](T2 *d, int l, int c) { store_once2(d, at(l, c), at(l, c + 1)); };
        if (whole) {
            // rows [H b, H b + H) of column block cb are H * GCB contiguous elements
            T2 *d0 = dst + (size_t)H * b * GCB;
            for (int e = 2 * threadIdx.x; e < H * n; e += 2 * NT) {
                const int cb = e / (H * GCB), w = e - cb * (H * GCB);
                const int l = w / GCB, c = cb * GCB + (w - l * GCB);
                put2(d0 + (size_t)cb * n * GCB + w, l, c);
            }
        } else {
            for (int e = 2 * threadIdx.x; e < H * n; e += 2 * NT) {
                const int l = e / n, c = e - l * n;
                put2(dst + e, l, c);
            }
        }
    }
}

// Sums the spilled parts of every key split over several items (in item
// order) into the task's subgrid band.  Block s takes the item that holds
// spill slot s (the band pass records it; slots past the batch's count exit),
// so the P items of a key split into P parts each sum a 1 / P share of the
// band's elements.
template <typename T2, int H, int LOGN>
__global__ __launch_bounds__(256) void k_band_combine(const Item *__restrict__ items,
                                                      const uint32_t *__restrict__ item_start,
                                                      BatchRef bref, int batch,
                                                      const int32_t *__restrict__ rank_of_slot,
                                                      T2 *G, const T2 *__restrict__ spill,
                                                      uint32_t spill_cap,
                                                      const uint32_t *__restrict__ item_spill,
                                                      const uint32_t *__restrict__ spill_ctr,
                                                      const uint32_t *__restrict__ spill_item)
{
    constexpr int n = 1 << LOGN;
    constexpr int NB = n / H;
    int t0, t1;
    if (!batch_tasks(bref, batch, t0, t1))
        return;
    if (blockIdx.x >= min(*spill_ctr, spill_cap))
        return;
    const uint32_t ib = item_start[(uint32_t)bref.task_slot[t0] * (uint32_t)NB];
    const uint32_t it = ib + spill_item[blockIdx.x];
    {
        const uint32_t key = items[it].key;
        const uint32_t k0 = item_start[key], k1 = item_start[key + 1];
        const uint32_t parts = k1 - k0;
        if (parts < 2)
            return;
        const uint32_t tslot = key / (uint32_t)NB;
        const int b = (int)(key - tslot * (uint32_t)NB);
        const int j = rank_of_slot[tslot] - t0;
        T2 *dst = G + (size_t)j * n * n;
        const uint32_t share = it - k0;
        const int e0 = (int)((uint64_t)(H * n) * share / parts);
        const int e1 = (int)((uint64_t)(H * n) * (share + 1) / parts);
        for (int e = e0 + (int)threadIdx.x; e < e1; e += blockDim.x) {
            T2 a;
            a.x = 0;
            a.y = 0;
            for (uint32_t q = k0; q < k1; ++q) {
                const uint32_t sl = item_spill[q - ib];
                if (sl < spill_cap) {
                    const T2 v = spill[(size_t)sl * H * n + e];
                    a.x += v.x;
                    a.y += v.y;
                }
            }
            dst[gcb_index(H * b + e / n, e % n, n)] = a;
        }
    }
}

// ---- batched column pass -------------------------------------------------------------------
// Block (cb, c): column block cb of chunk c (up to 32 consecutive tasks of
// one tower).  For every task the column block of its subgrid is inverse
// FFT'd with the first stage reading straight from HBM (bands without records
// are zero and are not read) and the last stage's outputs, which stay with
// their thread, multiplied by E^wp = exp(-2 pi i wp winc base) and
// accumulated in registers; the chunk's sum is written to its partial image.
// Runs of up to SCREEN_RUN consecutive floors are summed by Horner's rule,
// h = h conj(E) + val, and the run's sum h E^wp_last is added with E^wp_last
// evaluated exactly: one screen evaluation per run instead of per floor, and
// the rounding of E compounds over at most SCREEN_RUN - 1 floors.
constexpr int SCREEN_RUN = 8;

template <typename T, int LOGN, int CB, int H, int RX>
__global__ __launch_bounds__(CB *((1 << LOGN) / 8)) void k_tower_cols(
    const typename Vec2<T>::type *__restrict__ G, const uint32_t *__restrict__ band_masks,
    BatchRef bref, int batch, const int32_t *__restrict__ task_wp,
    const int32_t *__restrict__ chunk_first, const int32_t *__restrict__ chunk_count,
    typename Vec2<T>::type *P, const typename Vec2<T>::type *__restrict__ twg, double theta,
    double winc)
{
    using T2 = typename Vec2<T>::type;
    constexpr int n = 1 << LOGN;
    constexpr int NT = CB * (n / 8);
    constexpr int PER = n * CB / NT; // 8
    constexpr int NB = n / H;
    __shared__ T2 lds[n * CB];
    __shared__ T2 tw[n];
    constexpr int MW = (NB + 31) / 32; // band-mask words per task
    static_assert(MW <= 4, "at most 128 bands");
    int t0, t1;
    if (!batch_tasks(bref, batch, t0, t1))
        return;
    const int nch = bref.totals[2];
    if ((int)blockIdx.y >= nch)
        return;
    for (int k = threadIdx.x; k < n; k += NT)
        tw[k] = twg[k];
    const int line = threadIdx.x % CB;
    const int col = blockIdx.x * CB + line;

    // the thread's output pixels (row fft_out_index(slot), col): screen
    // exponent, one-floor step conj(E), total and the current run's sum
    double base[PER];
    T2 step[PER], acc[PER], h[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        base[q] = screen_base(col, fft_out_index<LOGN, CB, RX>(q), n, theta);
        Trig<T>::rot(winc * base[q], step[q].x, step[q].y);
    }
    // acc += h E^wl
    auto close_run = [&](int wl) {
        const double ph = -(double)wl * winc;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            T cs, sn;
            Trig<T>::rot(ph * base[q], cs, sn);
            acc[q].x += h[q].x * cs - h[q].y * sn;
            acc[q].y += h[q].x * sn + h[q].y * cs;
        }
    };
    static_assert(CB == GCB || (1 << LOGN) < GCB, "column pass reads whole column blocks");
    for (int c = blockIdx.y; c < nch; c += gridDim.y) {
        // the chunk's tasks inside this batch
        const int first = max(chunk_first[c], t0);
        const int last = min(chunk_first[c] + chunk_count[c], t1);
        if (first >= last)
            continue;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            acc[q].x = T(0);
            acc[q].y = T(0);
        }
        int run = 0, wp_prev = 0;
        for (int k = first; k < last; ++k) {
            const T2 *g = G + (size_t)(k - t0) * n * n + (size_t)blockIdx.x * n * CB + line;
            // band b of the task was written iff it has work items (bit b of
            // the task's mask; uniform loads)
            uint32_t bm[MW];
#pragma unroll
            for (int q = 0; q < MW; ++q)
                bm[q] = band_masks[(size_t)k * MW + q];
            const int wp = task_wp[k];
            if (run > 0 && (run == SCREEN_RUN || wp != wp_prev + 1)) {
                close_run(wp_prev);
                run = 0;
            }
            const bool extend = run > 0;
            auto load = [&](int, int row) {
                T2 val;
                val.x = T(0);
                val.y = T(0);
                const int bnd = row / H;
                if ((bm[bnd >> 5] >> (bnd & 31)) & 1u)
                    val = g[(size_t)row * CB];
                return val;
            };
            auto emit = [&](int q, int, T2 val) {
                if (extend) { // h conj(E) + val
                    const T hx = h[q].x, hy = h[q].y;
                    h[q].x = fma(hx, step[q].x, fma(-hy, step[q].y, val.x));
                    h[q].y = fma(hx, step[q].y, fma(hy, step[q].x, val.y));
                } else {
                    h[q] = val;
                }
            };
            __syncthreads(); // the previous floor's last stage has read lds
            fft_block_io<T2, LOGN, CB, CB, true, RX>(lds, tw, load, emit);
            ++run;
            wp_prev = wp;
        }
        if (run > 0)
            close_run(wp_prev);
        T2 *pp = P + (size_t)c * n * n + col;
#pragma unroll
        for (int q = 0; q < PER; ++q)
            pp[(size_t)fft_out_index<LOGN, CB, RX>(q) * n] = acc[q];
    }
}

// subimage[r] += the partial images of rank r's chunks that meet the batch
template <typename T2>
__global__ __launch_bounds__(256) void k_tower_reduce(const T2 *__restrict__ P, T2 *subimg,
                                                      int64_t nn, BatchRef bref, int batch,
                                                      const int32_t *__restrict__ rank_chunk0,
                                                      const int32_t *__restrict__ rank_nchunk,
                                                      const int32_t *__restrict__ chunk_first,
                                                      const int32_t *__restrict__ chunk_count)
{
    int t0, t1;
    if (!batch_tasks(bref, batch, t0, t1))
        return;
    const int ntow = bref.totals[1], nch = bref.totals[2];
    for (int r = blockIdx.y; r < ntow; r += gridDim.y) {
        const int c0 = rank_chunk0[r], c1 = min(nch, c0 + rank_nchunk[r]);
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn;
             i += (int64_t)gridDim.x * blockDim.x) {
            T2 a = subimg[(size_t)r * nn + i];
            for (int c = c0; c < c1; ++c) {
                if (chunk_first[c] >= t1 || chunk_first[c] + chunk_count[c] <= t0)
                    continue;
                const T2 p = P[(size_t)c * nn + i];
                a.x += p.x;
                a.y += p.y;
            }
            subimg[(size_t)r * nn + i] = a;
        }
    }
}

// ---- fold ------------------------------------------------------------------------------------
// Owner computes: each master pixel sums the towers whose window covers it.
// Channel c's master grid is the c-th N x N grid; its towers have ids
// (c cc + cy) cc + cx.
template <typename T2>
__global__ __launch_bounds__(256) void k_fold(double2 *grid, int N, int64_t total,
                                              const T2 *subimg, const int32_t *rank_of_tower, int n,
                                              int margin, int chunk, int cc)
{
    const int64_t NN = (int64_t)N * N;
    for (int64_t gidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gidx < total;
         gidx += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(gidx / NN);
        const int64_t idx = gidx - (int64_t)ch * NN;
        const int py = (int)(idx / N), px = (int)(idx - (int64_t)py * N);
        const int x = px - N / 2, y = py - N / 2;
        const double norm = (double)n * (double)n;
        double re = 0.0, im = 0.0;
        bool any = false;
        const int ax = x + N / 2 + margin / 2, ay = y + N / 2 + margin / 2;
        const int cxh = min(cc - 1, ax / chunk), cyh = min(cc - 1, ay / chunk);
        for (int cy = max(0, cyh - 2); cy <= cyh; ++cy) {
            const int y0 = chunk * cy - N / 2 - margin / 2;
            if (y < y0 || y >= y0 + n)
                continue;
            for (int cx = max(0, cxh - 2); cx <= cxh; ++cx) {
                const int x0 = chunk * cx - N / 2 - margin / 2;
                if (x < x0 || x >= x0 + n)
                    continue;
                const int rk = rank_of_tower[(ch * cc + cy) * cc + cx];
                if (rk < 0)
                    continue;
                const T2 s = subimg[(size_t)rk * n * n + (size_t)(y - y0) * n + (x - x0)];
                re += (double)s.x / norm;
                im += (double)s.y / norm;
                any = true;
            }
        }
        if (any) {
            double2 g = grid[gidx];
            g.x += re;
            g.y += im;
            grid[gidx] = g;
        }
    }
}

// ---- plain-mode kernels ---------------------------------------------------------------------

// Romein scatter of one task's records, launched <<<128, 128>>> like
// scatter_grid_kernel_flat in wtowers_CUDA_flat (grid_cuda.cu:708).
__global__ __launch_bounds__(128) void k_tower_plain_scatter(const Rec<double> *__restrict__ recs,
                                                             uint32_t start, uint32_t count,
                                                             int nbins,
                                                             const double2 *__restrict__ kern,
                                                             int S, double2 *sub, int n)
{
    const uint32_t per = count / nbins;
    const uint32_t i0 = start + per * blockIdx.x;
    const uint32_t i1 = (blockIdx.x == (unsigned)nbins - 1) ? start + count : i0 + per;
    const int SS = S * S;
    for (int pos = threadIdx.x; pos < SS; pos += blockDim.x) {
        const int myU = pos % S, myV = pos / S;
        int cu = INT_MIN, cv = INT_MIN;
        double sr = 0.0, si = 0.0;
        for (uint32_t i = i0; i < i1; ++i) {
            const Rec<double> r = recs[i];
            const int gx = (int)(r.pos & 0xffffu) - S, gy = (int)(r.pos >> 16) - S;
            int conv_u = (myU - gx) % S, conv_v = (myV - gy) % S;
            if (conv_u < 0)
                conv_u += S;
            if (conv_v < 0)
                conv_v += S;
            const int gu = gx + conv_u, gv = gy + conv_v;
            if (gu != cu || gv != cv) {
                if (cu != INT_MIN && (unsigned)cu < (unsigned)n && (unsigned)cv < (unsigned)n) {
                    double *g = reinterpret_cast<double *>(sub + (size_t)cv * n + cu);
                    atomicAdd(g, sr);
                    atomicAdd(g + 1, si);
                }
                sr = si = 0.0;
                cu = gu;
                cv = gv;
            }
            const double2 k = kern[(size_t)r.kidx * SS + (size_t)conv_v * S + conv_u];
            sr += k.x * r.vis.x + k.y * r.vis.y;
            si += k.x * r.vis.y - k.y * r.vis.x;
        }
        if (cu != INT_MIN && (unsigned)cu < (unsigned)n && (unsigned)cv < (unsigned)n) {
            double *g = reinterpret_cast<double *>(sub + (size_t)cv * n + cu);
            atomicAdd(g, sr);
            atomicAdd(g + 1, si);
        }
    }
}

// subimg = F^k * subimg + subgrid (fresnel_subimg_mul, common_kernels.cu:13-28,
// with the power taken as grid.c:475 does for skipped floors).
__global__ __launch_bounds__(256) void k_fresnel_mac(const double2 *sub, double2 *img, int n,
                                                     int k, double theta, double winc)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * n)
        return;
    const int y = (int)(idx / n), x = (int)(idx - (int64_t)y * n);
    double c, s;
    Trig<double>::rot((double)k * winc * screen_base(x, y, n, theta), c, s);
    const double2 a = img[idx], b = sub[idx];
    img[idx] = make_double2(a.x * c - a.y * s + b.x, a.x * s + a.y * c + b.y);
}

// subimg /= F^last_wp (w0_transfer_kernel, common_kernels.cu:31-40)
__global__ __launch_bounds__(256) void k_w0_transfer(double2 *img, int n, int last_wp, double theta,
                                                     double winc)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * n)
        return;
    const int y = (int)(idx / n), x = (int)(idx - (int64_t)y * n);
    double c, s;
    Trig<double>::rot(-(double)last_wp * winc * screen_base(x, y, n, theta), c, s);
    const double2 a = img[idx];
    img[idx] = make_double2(a.x * c - a.y * s, a.x * s + a.y * c);
}

// ---- host -------------------------------------------------------------------------------------

// Band height: rows of the subgrid one workgroup owns.  H rows of n complex
// values take H n sizeof(T2) bytes of LDS; 8 rows of FP32 leave room for four
// workgroups per CU.  16 rows are FP32 only.
static int band_height_of(const rg_grid_params *p, bool fp32)
{
    if (p->tile == 16 && fp32)
        return 16;
    if (p->tile == 4)
        return 4;
    return 8;
}

struct BandLaunch {
    const void *recs;
    const Item *items;
    const uint32_t *item_start;
    BatchRef bref;
    int batch;
    const int32_t *rank_of_slot;
    const void *kern, *tw;
    uint32_t kern_bytes; // kernel table size (< 2^31)
    void *G, *spill;
    uint32_t spill_cap;
    uint32_t *spill_ctr, *item_spill, *spill_item;
    unsigned long long *stats;
    uint32_t blocks; // workgroups (each loops over the batch's items)
    rg_context *timer_ctx;
    bool comp;       // FP32: two-sum compensated register sums
};

template <typename T, int S, int H, int LOGN, bool COMP>
static int launch_band_items(const BandLaunch &a, hipStream_t s)
{
    using T2 = typename Vec2<T>::type;
    constexpr int GRP = S * H >= 256 ? 1 : 256 / (S * H);
    KernelTimer kt(a.timer_ctx, RG_TIMER_TOWER_TILES, s);
    hipLaunchKernelGGL((k_tower_band_items<T, S, H, LOGN, GRP, COMP>), dim3(a.blocks),
                       dim3(S * H * GRP), 0, s, (const BandRec<T> *)a.recs, a.items,
                       a.item_start, a.bref, a.batch, a.rank_of_slot, (const T2 *)a.kern,
                       a.kern_bytes, (const T2 *)a.tw, (T2 *)a.G, (T2 *)a.spill, a.spill_cap,
                       a.spill_ctr, a.item_spill, a.spill_item, a.stats);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

template <typename T, int S, int H, int LOGN>
static int launch_bands_h(const BandLaunch &a, hipStream_t s)
{
    using T2 = typename Vec2<T>::type;
    if constexpr (sizeof(T) == 4) {
        if (a.comp)
            RG_TRY((launch_band_items<T, S, H, LOGN, true>(a, s)));
        else
            RG_TRY((launch_band_items<T, S, H, LOGN, false>(a, s)));
    } else {
        RG_TRY((launch_band_items<T, S, H, LOGN, false>(a, s)));
    }
    KernelTimer kt(a.timer_ctx, RG_TIMER_TOWER_COMBINE, s);
    hipLaunchKernelGGL((k_band_combine<T2, H, LOGN>), dim3(std::min(a.spill_cap, a.blocks)),
                       dim3(256), 0, s, a.items, a.item_start, a.bref, a.batch, a.rank_of_slot,
                       (T2 *)a.G, (const T2 *)a.spill, a.spill_cap,
                       (const uint32_t *)a.item_spill, (const uint32_t *)a.spill_ctr,
                       (const uint32_t *)a.spill_item);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

template <typename T, int S, int LOGN>
static int launch_bands_t(int H, const BandLaunch &a, hipStream_t s)
{
    if constexpr (sizeof(T) == 4) {
        if (H == 16)
            return launch_bands_h<T, S, 16, LOGN>(a, s);
    }
    if (H == 4)
        return launch_bands_h<T, S, 4, LOGN>(a, s);
    return launch_bands_h<T, S, 8, LOGN>(a, s);
}

template <typename T, int S>
static int launch_bands(int n, int H, const BandLaunch &a, hipStream_t s)
{
    switch (n) {
    case 64: return launch_bands_t<T, S, 6>(H, a, s);
    case 128: return launch_bands_t<T, S, 7>(H, a, s);
    case 256: return launch_bands_t<T, S, 8>(H, a, s);
    case 512: return launch_bands_t<T, S, 9>(H, a, s);
    default: return fail(RG_ERR_UNSUPPORTED, "batched W-towers supports subgrid sizes 64..512 (got %d)", n);
    }
}

struct ColsLaunch {
    const void *G;
    const uint32_t *band_masks; // per task, (n / H + 31) / 32 words
    BatchRef bref;
    int batch;
    const int32_t *task_wp, *chunk_first, *chunk_count;
    void *P;
    const void *tw;
    double theta, winc;
    unsigned grid_y; // workgroup rows (each loops over the chunks)
    int radix;       // FFT stage radix (8 or 4)
};

template <typename T, int LOGN, int H, int RX>
static int launch_cols_r(const ColsLaunch &a, hipStream_t s)
{
    using T2 = typename Vec2<T>::type;
    constexpr int n = 1 << LOGN;
    constexpr int CB0 = 8; // columns per workgroup (512 threads at n = 512)
    constexpr int CB = CB0 > n ? n : CB0;
    hipLaunchKernelGGL((k_tower_cols<T, LOGN, CB, H, RX>), dim3(n / CB, a.grid_y),
                       dim3(CB * (n / 8)), 0, s, (const T2 *)a.G, a.band_masks, a.bref, a.batch,
                       a.task_wp, a.chunk_first, a.chunk_count, (T2 *)a.P, (const T2 *)a.tw,
                       a.theta, a.winc);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

template <typename T, int LOGN, int H>
static int launch_cols_h(const ColsLaunch &a, hipStream_t s)
{
    return a.radix == 4 ? launch_cols_r<T, LOGN, H, 4>(a, s) : launch_cols_r<T, LOGN, H, 8>(a, s);
}

template <typename T, int LOGN>
static int launch_cols_t(int H, const ColsLaunch &a, hipStream_t s)
{
    if (H == 16)
        return launch_cols_h<T, LOGN, 16>(a, s);
    if (H == 4)
        return launch_cols_h<T, LOGN, 4>(a, s);
    return launch_cols_h<T, LOGN, 8>(a, s);
}

template <typename T>
static int launch_cols(int n, int H, const ColsLaunch &a, hipStream_t s)
{
    switch (n) {
    case 64: return launch_cols_t<T, 6>(H, a, s);
    case 128: return launch_cols_t<T, 7>(H, a, s);
    case 256: return launch_cols_t<T, 8>(H, a, s);
    case 512: return launch_cols_t<T, 9>(H, a, s);
    default: return fail(RG_ERR_UNSUPPORTED, "batched W-towers supports subgrid sizes 64..512 (got %d)", n);
    }
}

// Geometry of a W-towers call.
struct WtGeom {
    int N, n, M, S, chunk, cc, H, NB, mk, nchan;
    int64_t ntw; // tower ids: nchan cc^2
};

static WtGeom wt_geom(const rg_grid_params *p, const rg_wkern *wk, int nchan, int H)
{
    WtGeom g;
    g.N = p->grid_size;
    g.n = p->subgrid_size;
    g.M = p->subgrid_margin;
    g.S = wk->size;
    g.chunk = g.n - g.M;
    g.cc = g.N / g.chunk + 1; // grid_cuda.cu:537-539
    g.H = H;
    g.NB = H > 0 ? g.n / H : 1;
    g.mk = H > 0 ? (g.S - 2) / H + 2 : 1; // bands an S-row footprint can reach
    g.nchan = nchan;
    g.ntw = (int64_t)nchan * g.cc * g.cc;
    return g;
}

static TowerKeyer wt_keyer(const rg_grid_params *p, const rg_flat_vis *vis, const rg_wkern *wk,
                           const WtGeom &g, const int *d_range, int nwp, const int64_t *d_off)
{
    TowerKeyer kf;
    kf.u = vis->u;
    kf.v = vis->v;
    kf.w = vis->w;
    kf.vis = vis->vis;
    kf.vis_type = vis->vis_type;
    kf.N = g.N;
    kf.S = g.S;
    kf.R = wk->oversampling;
    kf.P = wk->plane_count;
    kf.theta = p->theta;
    kf.w_min = wk->w_min;
    kf.w_step = wk->w_step;
    kf.n = g.n;
    kf.chunk = g.chunk;
    kf.cc = g.cc;
    kf.winc = p->winc;
    kf.range = d_range;
    kf.nwp = nwp;
    kf.band_h = g.H;
    kf.nk = g.NB;
    kf.mk = g.mk;
    kf.chan_off = d_off;
    kf.nchan = g.nchan;
    return kf;
}

// Floor range of the call into d_range[0..1] (on the device).
static int floor_range(const rg_grid_params *p, const rg_flat_vis *vis, const WtGeom &g,
                       int *d_range, hipStream_t s)
{
    // d_range: RANGE_BYTES (the range, then the workgroups' parts)
    int2 *part = reinterpret_cast<int2 *>(reinterpret_cast<char *>(d_range) + 256);
    int64_t blocks = (vis->n + 256 * RANGE_VPT - 1) / (256 * RANGE_VPT);
    if (blocks > RANGE_BLOCKS)
        blocks = RANGE_BLOCKS;
    if (blocks > 0)
        hipLaunchKernelGGL(k_floor_range, dim3((unsigned)blocks), dim3(256), 0, s, vis->u, vis->v,
                           vis->w, vis->n, p->theta, g.N, g.chunk, g.cc, p->winc, part);
    hipLaunchKernelGGL(k_range_finish, dim3(1), dim3(1024), 0, s, (const int2 *)part,
                       (int)blocks, d_range);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

// Synchronous read of the floor range; false when no visibility falls in the
// tower set.
static int read_range(rg_context *ctx, const int *d_range, int &wp_min, int &nwp, bool &empty,
                      hipStream_t s)
{
    int h[2];
    RG_HIP(hipMemcpyAsync(h, d_range, sizeof(h), hipMemcpyDeviceToHost, s));
    RG_HIP(hipStreamSynchronize(s));
    empty = h[0] > h[1];
    wp_min = h[0];
    nwp = empty ? 0 : h[1] - h[0] + 1;
    (void)ctx;
    return RG_OK;
}

// ---- plain mode: the reference's launch structure (one task at a time) -------------------------

struct TaskList {
    std::vector<int32_t> slot, tower_rank, wp;
    std::vector<uint32_t> nvis;
    std::vector<int32_t> tower_of_rank; // tower index of each active tower
    std::vector<int32_t> rank_of_tower; // -1 when inactive
};

static int wtowers_plain(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
                         const rg_wkern *wk, void *grid, hipStream_t s)
{
    const WtGeom g = wt_geom(p, wk, 1, 0);
    const int N = g.N, n = g.n, M = g.M, S = g.S, cc = g.cc;
    RG_TRY(zero_fill(grid, sizeof(double2) * (size_t)N * N, s));
    if (vis->n == 0)
        return RG_OK;
    RG_TRY(ctx->scratch[8].ensure(RANGE_BYTES));
    int *d_range = ctx->scratch[8].as<int>();
    RG_TRY(floor_range(p, vis, g, d_range, s));
    int wp_min = 0, nwp = 0;
    bool empty = false;
    RG_TRY(read_range(ctx, d_range, wp_min, nwp, empty, s));
    if (empty) {
        unsigned long long cnt[CNT__N] = {0};
        cnt[CNT_DROPPED] = (unsigned long long)vis->n;
        RG_HIP(hipMemcpyAsync(ctx->counters.p, cnt, sizeof(cnt), hipMemcpyHostToDevice, s));
        RG_HIP(hipStreamSynchronize(s));
        return RG_OK;
    }
    const TowerKeyer kf = wt_keyer(p, vis, wk, g, d_range, nwp, nullptr);
    const int64_t nslots = (int64_t)cc * cc * nwp;
    BinResult br;
    RG_TRY((bin_visibilities<TowerKeyer, double>(ctx, kf, vis->n, nslots, 0x40000000u, s, br,
                                                 (int64_t)1 << 20)));
    RG_TRY(ctx->scratch[9].ensure(sizeof(uint32_t) * (2 * nslots + 1)));
    uint32_t *d_nvis = ctx->scratch[9].as<uint32_t>();
    uint32_t *d_item0 = d_nvis + nslots;
    hipLaunchKernelGGL(k_task_info, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, s,
                       (const uint32_t *)br.vis_start, (const uint32_t *)br.item_start, nslots, 1,
                       d_nvis, d_item0);
    RG_HIP(hipGetLastError());
    std::vector<uint32_t> h_nvis(nslots);
    RG_HIP(hipMemcpyAsync(h_nvis.data(), d_nvis, sizeof(uint32_t) * nslots, hipMemcpyDeviceToHost,
                          s));
    RG_HIP(hipStreamSynchronize(s));

    TaskList tl;
    tl.rank_of_tower.assign((size_t)cc * cc, -1);
    for (int64_t t = 0; t < nslots; ++t) {
        const uint32_t nv = h_nvis[t];
        if (!nv)
            continue;
        const int tower = (int)(t / nwp);
        if (tl.rank_of_tower[tower] < 0) {
            tl.rank_of_tower[tower] = (int32_t)tl.tower_of_rank.size();
            tl.tower_of_rank.push_back(tower);
        }
        tl.slot.push_back((int32_t)t);
        tl.tower_rank.push_back(tl.rank_of_tower[tower]);
        tl.wp.push_back(wp_min + (int)(t % nwp));
        tl.nvis.push_back(nv);
    }
    const int ntasks = (int)tl.slot.size();
    const int ntowers = (int)tl.tower_of_rank.size();
    {
        unsigned long long tc[2] = {(unsigned long long)ntasks, (unsigned long long)ntowers};
        RG_HIP(hipMemcpyAsync(ctx->counters.as<unsigned long long>() + CNT_TASKS, tc, sizeof(tc),
                              hipMemcpyHostToDevice, s));
    }
    if (ntasks == 0)
        return RG_OK;
    const size_t nn = (size_t)n * n;
    RG_TRY(ctx->scratch[10].ensure(sizeof(double2) * nn * ntowers));
    double2 *subimg = ctx->scratch[10].as<double2>();
    RG_TRY(zero_fill(subimg, sizeof(double2) * nn * ntowers, s));
    RG_TRY(ctx->scratch[18].ensure(sizeof(int32_t) * cc * cc));
    int32_t *d_rot = ctx->scratch[18].as<int32_t>();
    RG_HIP(hipMemcpyAsync(d_rot, tl.rank_of_tower.data(), sizeof(int32_t) * cc * cc,
                          hipMemcpyHostToDevice, s));
    RG_TRY(ctx->scratch[7].ensure(sizeof(double2) * nn));
    double2 *sub = ctx->scratch[7].as<double2>();
    const Rec<double> *recs = (const Rec<double> *)br.recs;
    std::vector<uint32_t> start_of(ntasks);
    {
        uint32_t run = 0;
        int k2 = 0;
        for (int64_t t = 0; t < nslots; ++t) {
            if (k2 < ntasks && tl.slot[k2] == t)
                start_of[k2++] = run;
            run += h_nvis[t];
        }
    }
    const int blocks_px = (int)((nn + 255) / 256);
    int k = 0;
    for (int r = 0; r < ntowers; ++r) {
        double2 *img = subimg + (size_t)r * nn;
        int last_wp = 0;
        bool first = true;
        for (; k < ntasks && tl.tower_rank[k] == r; ++k) {
            RG_TRY(zero_fill(sub, sizeof(double2) * nn, s));
            hipLaunchKernelGGL(k_tower_plain_scatter, dim3(128), dim3(128), 0, s, recs, start_of[k],
                               tl.nvis[k], 128, (const double2 *)wk->data, S, sub, n);
            RG_HIP(hipGetLastError());
            RG_TRY(fft2d(ctx, sub, RG_C128, n, 1, +1, s));
            if (first) {
                last_wp = tl.wp[k];
                first = false;
            }
            hipLaunchKernelGGL(k_fresnel_mac, dim3(blocks_px), dim3(256), 0, s,
                               (const double2 *)sub, img, n, tl.wp[k] - last_wp, p->theta,
                               p->winc);
            RG_HIP(hipGetLastError());
            last_wp = tl.wp[k];
        }
        if (last_wp != 0) {
            hipLaunchKernelGGL(k_w0_transfer, dim3(blocks_px), dim3(256), 0, s, img, n, last_wp,
                               p->theta, p->winc);
            RG_HIP(hipGetLastError());
        }
        RG_TRY(fft2d(ctx, img, RG_C128, n, 1, -1, s));
    }
    KernelTimer kt_finish(ctx, RG_TIMER_TOWER_FINISH, s);
    const int64_t npx = (int64_t)N * N;
    hipLaunchKernelGGL((k_fold<double2>), dim3((unsigned)std::min<int64_t>((npx + 255) / 256, 65536)),
                       dim3(256), 0, s, (double2 *)grid, N, npx, (const double2 *)subimg,
                       (const int32_t *)d_rot, n, M, g.chunk, cc);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

// ---- batched mode ----------------------------------------------------------------------------
// Every table is built on the device.  The key space and the buffers are sized
// by a plan kept in the context: the first call of a shape (or one that
// outgrows its plan) reads the floor range and the task counts back and makes
// the plan, with some headroom; later calls of that shape issue no read-back
// until their end, where an eager call checks the overrun counter and redoes
// itself with a fresh plan if a capacity was exceeded.  A call on a capturing
// stream never reads back: it needs a plan from an eager call of its shape,
// and an overrun is reported by rg_context_stats.

static bool plan_fits(const WtPlan &pl, const rg_grid_params *p, const rg_flat_vis *vis,
                      const rg_wkern *wk, int nchan, int H, int type)
{
    return pl.valid && pl.N == p->grid_size && pl.n == p->subgrid_size &&
           pl.M == p->subgrid_margin && pl.S == wk->size && pl.R == wk->oversampling &&
           pl.P == wk->plane_count && pl.H == H && pl.type == type && pl.nchan == nchan &&
           pl.K == (p->vis_per_item > 0 ? p->vis_per_item : 2048) &&
           pl.floor_batch == p->floor_batch && pl.theta == p->theta && pl.winc == p->winc &&
           pl.w_min == wk->w_min && pl.w_step == wk->w_step && vis->n <= pl.nvis;
}

template <typename T>
static int wtowers_batched_t(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
                             const int64_t *d_off, int nchan, const rg_wkern *wk, void *grid,
                             hipStream_t s, bool capturing, bool measure)
{
    using T2 = typename Vec2<T>::type;
    const int H = band_height_of(p, sizeof(T) == 4);
    const WtGeom g = wt_geom(p, wk, nchan, H);
    const int n = g.n, S = g.S, NB = g.NB;
    const size_t nn = (size_t)n * n;
    const int type = sizeof(T) == 4 ? RG_C64 : RG_C128;
    const uint32_t K = p->vis_per_item > 0 ? (uint32_t)p->vis_per_item : 2048u;
    WtPlan &pl = ctx->wt_plan;
    auto *stats = ctx->counters.as<unsigned long long>();

    RG_TRY(zero_fill(grid, sizeof(double2) * (size_t)g.N * g.N * nchan, s));
    if (vis->n == 0)
        return RG_OK;

    // floor range (device); the plan fixes the floors per tower slot
    RG_TRY(ctx->scratch[8].ensure(RANGE_BYTES));
    int *d_range = ctx->scratch[8].as<int>();
    RG_TRY(floor_range(p, vis, g, d_range, s));
    int nwp_cap;
    if (measure) {
        int wp_min = 0, nwp = 0;
        bool empty = false;
        RG_TRY(read_range(ctx, d_range, wp_min, nwp, empty, s));
        if (empty) {
            unsigned long long cnt[CNT__N] = {0};
            cnt[CNT_DROPPED] = (unsigned long long)vis->n;
            RG_HIP(hipMemcpyAsync(ctx->counters.p, cnt, sizeof(cnt), hipMemcpyHostToDevice, s));
            RG_HIP(hipStreamSynchronize(s));
            return RG_OK;
        }
        nwp_cap = nwp + nwp / 8 + 4;
    } else {
        nwp_cap = pl.nwp_cap;
    }
    const int64_t nslots = g.ntw * nwp_cap;
    const int64_t nkeys = nslots * NB;
    // band records are addressed by 32-bit offsets: the expanded array holds
    // up to mk records per visibility
    if ((int64_t)g.mk * vis->n >= ((int64_t)1 << 32))
        return fail(RG_ERR_PARAM,
                    "W-towers call too large: %lld visibilities x %d bands per footprint exceed "
                    "2^32 band records; split the stream",
                    (long long)vis->n, g.mk);
    if (nkeys >= ((int64_t)1 << 31))
        return fail(RG_ERR_PARAM, "W-towers key space too large (%lld keys: %lld towers x %d "
                    "floors x %d bands)", (long long)nkeys, (long long)g.ntw, nwp_cap, NB);
    // the sort makes one item per key; band work items of at most K records
    // per band window are built after it
    const int64_t max_items_w = nkeys + (int64_t)g.mk * vis->n / K + 1;
    RG_TRY(ctx->scratch[4].ensure(sizeof(Item) * std::max<int64_t>(nkeys + 1, max_items_w)));
    const TowerKeyer kf = wt_keyer(p, vis, wk, g, d_range, nwp_cap, d_off);
    BinResult br;
    // (task, band) keys: 4 counter replicas at the reference's workload (the
    // scans read every replica; 16 were slower, 1 << 20 counters in all)
    RG_TRY((bin_visibilities<TowerKeyer, T>(ctx, kf, vis->n, nkeys, 0x40000000u, s, br,
                                            (int64_t)1 << 20)));

    // capacities: tables by the slot count, buffers by the plan (or, measuring,
    // by the counts read back below)
    const int ntw = (int)g.ntw;
    const int64_t nchunks_bound = nslots / TASKS_PER_CHUNK + g.ntw;
    int ntasks_cap, nchunks_cap, ntowers_cap, nitems_cap;
    if (measure) {
        ntasks_cap = (int)std::min<int64_t>(nslots, 0x7fffffff);
        nchunks_cap = (int)std::min<int64_t>(nchunks_bound, 0x7fffffff);
        ntowers_cap = ntw;
        nitems_cap = (int)std::min<int64_t>(max_items_w, 0x7fffffff);
    } else {
        ntasks_cap = pl.ntasks_cap;
        nchunks_cap = pl.nchunks_cap;
        ntowers_cap = pl.ntowers_cap;
        nitems_cap = pl.nitems_cap;
    }
    // table layout (int32): 7 per-tower arrays, 3 per-task, 2 per-chunk, one
    // per slot, totals
    const size_t o_tn = 0, o_to = o_tn + ntw, o_co = o_to + ntw, o_rot = o_co + ntw,
                 o_tor = o_rot + ntw, o_rc0 = o_tor + ntw, o_rnc = o_rc0 + ntw,
                 o_ts = o_rnc + ntw, o_twp = o_ts + ntasks_cap, o_trk = o_twp + ntasks_cap,
                 o_cf = o_trk + ntasks_cap, o_cn = o_cf + nchunks_cap, o_ros = o_cn + nchunks_cap,
                 o_tot = o_ros + nslots, o_end = o_tot + 16;
    RG_TRY(ctx->scratch[18].ensure(sizeof(int32_t) * o_end));
    int32_t *tab = ctx->scratch[18].as<int32_t>();
    TaskTables tt;
    tt.tower_ntasks = tab + o_tn;
    tt.tower_task_off = tab + o_to;
    tt.tower_chunk_off = tab + o_co;
    tt.rank_of_tower = tab + o_rot;
    tt.tower_of_rank = tab + o_tor;
    tt.rank_chunk0 = tab + o_rc0;
    tt.rank_nchunk = tab + o_rnc;
    tt.task_slot = tab + o_ts;
    tt.task_wp = tab + o_twp;
    tt.task_rank = tab + o_trk;
    tt.chunk_first = tab + o_cf;
    tt.chunk_count = tab + o_cn;
    tt.rank_of_slot = tab + o_ros;
    tt.totals = tab + o_tot;
    tt.ntw = ntw;
    tt.nwp = nwp_cap;
    tt.nk = NB;
    tt.ntasks_cap = ntasks_cap;
    tt.nchunks_cap = nchunks_cap;
    tt.ntowers_cap = ntowers_cap;
    hipLaunchKernelGGL(k_tower_count, dim3(ntw), dim3(256), 0, s, (const uint32_t *)br.vis_start,
                       tt);
    hipLaunchKernelGGL(k_tower_scan, dim3(1), dim3(1024), 0, s, tt, stats);
    hipLaunchKernelGGL(k_tower_tasks, dim3(ntw), dim3(256), 0, s, (const uint32_t *)br.vis_start,
                       tt, (const int *)d_range);
    RG_HIP(hipGetLastError());

    // band window work items (in the binning's item_start slot; window lengths
    // and prefix in the count / cursor slots)
    RG_TRY(ctx->scratch[9].ensure(sizeof(uint32_t) * (nkeys + 1)));
    uint32_t *d_wstart = ctx->scratch[9].as<uint32_t>();
    uint32_t *item_start_w = br.item_start;
    Item *items_w = br.items;
    uint32_t *prefix = ctx->scratch[3].as<uint32_t>();
    {
        KernelTimer kt(ctx, RG_TIMER_BINNING, s);
        uint32_t *len = ctx->scratch[0].as<uint32_t>();
        const unsigned wblocks = (unsigned)std::min<int64_t>((nkeys + 255) / 256, 1024);
        hipLaunchKernelGGL(k_window_len, dim3(wblocks), dim3(256), 0, s,
                           (const uint32_t *)br.vis_start, nkeys, NB, g.mk, len, d_wstart, stats);
        RG_HIP(hipGetLastError());
        RG_TRY(items_from_lengths(ctx, len, nkeys, K, prefix, item_start_w, items_w, s));
    }

    if (measure) {
        // the plan: counts of this call plus headroom for calls of the same shape
        unsigned long long c[CNT__N];
        RG_HIP(hipMemcpyAsync(c, ctx->counters.p, sizeof(c), hipMemcpyDeviceToHost, s));
        RG_HIP(hipStreamSynchronize(s));
        const int64_t nt = (int64_t)c[CNT_TASKS], ntow = (int64_t)c[CNT_TOWERS],
                      nch = (int64_t)c[CNT_CHUNKS], nit = (int64_t)c[CNT_ITEMS];
        pl.valid = false;
        pl.N = p->grid_size;
        pl.n = n;
        pl.M = g.M;
        pl.S = S;
        pl.R = wk->oversampling;
        pl.P = wk->plane_count;
        pl.H = H;
        pl.type = type;
        pl.nchan = nchan;
        pl.K = K;
        pl.floor_batch = p->floor_batch;
        pl.theta = p->theta;
        pl.winc = p->winc;
        pl.w_min = wk->w_min;
        pl.w_step = wk->w_step;
        pl.nvis = vis->n;
        pl.nwp_cap = nwp_cap;
        pl.ntasks_cap = (int)std::min<int64_t>(ntasks_cap, nt + nt / 8 + 16);
        pl.nchunks_cap = (int)std::min<int64_t>(nchunks_cap, nch + nch / 8 + 16);
        pl.ntowers_cap = (int)std::min<int64_t>(ntw, ntow + ntow / 8 + 4);
        pl.nitems_cap = (int)std::min<int64_t>(max_items_w, nit + nit / 8 + 64);
        if (nt == 0)
            return RG_OK;
        // this call's buffers use the plan's capacities (the tables above were
        // built with the slot-count bounds, which hold every index)
        ntasks_cap = pl.ntasks_cap;
        nchunks_cap = pl.nchunks_cap;
        ntowers_cap = pl.ntowers_cap;
        nitems_cap = pl.nitems_cap;
        pl.valid = true;
    }

    // subimages (one per active tower, by rank)
    RG_TRY(ctx->scratch[10].ensure(sizeof(T2) * nn * ntowers_cap));
    T2 *subimg = ctx->scratch[10].as<T2>();
    RG_TRY(zero_fill(subimg, sizeof(T2) * nn * ntowers_cap, s));
    const void *tw_inv = nullptr, *tw_stage = nullptr;
    RG_TRY(get_twiddles(ctx, n, type, +1, &tw_inv));
    RG_TRY(get_stage_twiddles(ctx, n, type, +1, &tw_stage));

    // batches of Q tasks (by default every task while the subgrids fit 16 GiB)
    constexpr size_t BATCH_BYTES = (size_t)16 << 30;
    const int q_fit = (int)std::max<size_t>(1, BATCH_BYTES / (sizeof(T2) * nn));
    const int Q = std::max(1, std::min(ntasks_cap, p->floor_batch > 0 ? p->floor_batch : q_fit));
    const int nbatch = (ntasks_cap + Q - 1) / Q;
    // spill slots (parts of split keys: at most 2 R / K for R records), per batch
    const uint64_t max_spill = 2 * ((uint64_t)g.mk * vis->n) / K + 1;
    const size_t band_bytes = sizeof(T2) * (size_t)H * n;
    const size_t g_bytes = sizeof(T2) * nn * Q;
    const size_t p_bytes = sizeof(T2) * nn * (size_t)nchunks_cap;
    const size_t sp_bytes = band_bytes * max_spill;
    const size_t is_bytes = (sizeof(uint32_t) * (size_t)max_items_w + 255) / 256 * 256;
    const size_t ctr_bytes = (sizeof(uint32_t) * nbatch + 255) / 256 * 256;
    const size_t si_bytes = sizeof(uint32_t) * (size_t)max_spill;
    RG_TRY(ctx->scratch[7].ensure(g_bytes + p_bytes + sp_bytes + is_bytes + ctr_bytes +
                                  si_bytes));
    char *base7 = ctx->scratch[7].as<char>();
    T2 *G = reinterpret_cast<T2 *>(base7);
    T2 *Pp = reinterpret_cast<T2 *>(base7 + g_bytes);
    T2 *spill = reinterpret_cast<T2 *>(base7 + g_bytes + p_bytes);
    uint32_t *item_spill = reinterpret_cast<uint32_t *>(base7 + g_bytes + p_bytes + sp_bytes);
    uint32_t *spill_ctr =
        reinterpret_cast<uint32_t *>(base7 + g_bytes + p_bytes + sp_bytes + is_bytes);
    uint32_t *spill_item = reinterpret_cast<uint32_t *>(base7 + g_bytes + p_bytes + sp_bytes +
                                                        is_bytes + ctr_bytes);
    RG_TRY(zero_fill(spill_ctr, sizeof(uint32_t) * nbatch, s));

    // every work item's records sorted by footprint column into the expanded
    // array (window records laid out item after item)
    const int64_t nexp = (int64_t)g.mk * vis->n;
    RG_TRY(ctx->scratch[11].ensure(sizeof(BandRec<T>) * (size_t)std::max<int64_t>(nexp, 1)));
    BandRec<T> *sorted = ctx->scratch[11].as<BandRec<T>>();
    {
        KernelTimer kt(ctx, RG_TIMER_BINNING, s);
        hipLaunchKernelGGL((k_sort_items<T>),
                           dim3((unsigned)std::min<int64_t>(max_items_w, 8192)), dim3(256), 0, s,
                           (const Rec<T> *)br.recs, sorted, (const Item *)items_w, stats,
                           (const uint32_t *)prefix, (const uint32_t *)d_wstart, n + S, S, H, NB);
        RG_HIP(hipGetLastError());
    }

    BandLaunch bl;
    bl.timer_ctx = ctx;
    bl.comp = p->compensated != 0;
    bl.recs = sorted;
    bl.items = items_w;
    bl.item_start = item_start_w;
    bl.bref.task_slot = tt.task_slot;
    bl.bref.totals = tt.totals;
    bl.bref.Q = Q;
    bl.rank_of_slot = tt.rank_of_slot;
    {
        // the FP32 band pass reads the table with zero rows after every
        // sub-kernel; taps are addressed by 32-bit byte offsets
        const int64_t nsub = (int64_t)wk->plane_count * wk->oversampling * wk->oversampling;
        const int64_t rows = padded_taps<T>() ? (int64_t)H + nsub * padded_rows(S, H) : nsub * S;
        const int64_t kb = rows * S * (int64_t)sizeof(T2);
        if (kb >= ((int64_t)1 << 31))
            return fail(RG_ERR_UNSUPPORTED, "batched W-towers kernel table too large (%lld bytes)",
                        (long long)kb);
        bl.kern = wk->data;
        if (padded_taps<T>()) {
            RG_TRY(ctx->scratch[12].ensure((size_t)kb));
            T2 *padded = ctx->scratch[12].as<T2>();
            hipLaunchKernelGGL(k_pad_kernels<T2>,
                               dim3((unsigned)std::min<int64_t>((rows * S + 255) / 256, 65536)),
                               dim3(256), 0, s, (const T2 *)wk->data, padded, nsub, S, H);
            RG_HIP(hipGetLastError());
            bl.kern = padded;
        }
        bl.kern_bytes = (uint32_t)kb;
    }
    bl.tw = tw_stage;
    bl.G = G;
    bl.spill = spill;
    bl.spill_cap = (uint32_t)std::min<uint64_t>(max_spill, 0xffffffffu);
    bl.item_spill = item_spill;
    bl.spill_item = spill_item;
    bl.stats = stats;
    // one band-pass workgroup per work item (the plan's item capacity; a
    // batch with more items reports an overrun)
    bl.blocks = (uint32_t)std::max(1, nitems_cap);
    // per task: which bands were written
    const int mw = (NB + 31) / 32;
    RG_TRY(ctx->scratch[19].ensure(sizeof(uint32_t) * (size_t)std::max(ntasks_cap, 1) * mw));
    uint32_t *band_masks = ctx->scratch[19].as<uint32_t>();
    {
        const int64_t ne = (int64_t)ntasks_cap * mw;
        hipLaunchKernelGGL(k_band_masks, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s,
                           (const uint32_t *)item_start_w, (const int32_t *)tt.task_slot,
                           (const int32_t *)tt.totals, NB, mw, band_masks);
        RG_HIP(hipGetLastError());
    }
    ColsLaunch cl;
    cl.G = G;
    cl.band_masks = band_masks;
    cl.bref = bl.bref;
    cl.task_wp = tt.task_wp;
    cl.chunk_first = tt.chunk_first;
    cl.chunk_count = tt.chunk_count;
    cl.P = Pp;
    cl.tw = tw_inv;
    cl.theta = p->theta;
    cl.winc = p->winc;
    cl.grid_y = (unsigned)std::min<int64_t>(std::max(nchunks_cap, 1), 65535);
    cl.radix = p->fft_radix == 4 ? 4 : 8;
    int64_t rgx = (int64_t)((nn + 255) / 256);
    if (rgx > 1024)
        rgx = 1024;
    const unsigned rgy = (unsigned)std::min<int64_t>(std::max(ntowers_cap, 1), 65535);
    for (int b = 0; b < nbatch; ++b) {
        bl.batch = b;
        bl.spill_ctr = spill_ctr + b;
        cl.batch = b;
        if (S == 16)
            RG_TRY((launch_bands<T, 16>(n, H, bl, s)));
        else
            RG_TRY((launch_bands<T, 8>(n, H, bl, s)));
        {
            KernelTimer kt(ctx, RG_TIMER_TOWER_COLS, s);
            RG_TRY(launch_cols<T>(n, H, cl, s));
        }
        {
            KernelTimer kt(ctx, RG_TIMER_TOWER_REDUCE, s);
            hipLaunchKernelGGL((k_tower_reduce<T2>), dim3((unsigned)rgx, rgy), dim3(256), 0, s,
                               (const T2 *)Pp, subimg, (int64_t)nn, bl.bref, b,
                               (const int32_t *)tt.rank_chunk0, (const int32_t *)tt.rank_nchunk,
                               (const int32_t *)tt.chunk_first, (const int32_t *)tt.chunk_count);
            RG_HIP(hipGetLastError());
        }
    }

    KernelTimer kt_finish(ctx, RG_TIMER_TOWER_FINISH, s);
    RG_TRY(fft2d(ctx, subimg, type, n, ntowers_cap, -1, s, nullptr, tt.totals + 1));
    const int64_t npx = (int64_t)g.N * g.N * nchan;
    hipLaunchKernelGGL((k_fold<T2>), dim3((unsigned)std::min<int64_t>((npx + 255) / 256, 1 << 20)),
                       dim3(256), 0, s, (double2 *)grid, g.N, npx, (const T2 *)subimg,
                       (const int32_t *)tt.rank_of_tower, n, g.M, g.chunk, g.cc);
    RG_HIP(hipGetLastError());
    (void)capturing;
    return RG_OK;
}

template <typename T>
static int wtowers_batched(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
                           const int64_t *d_off, int nchan, const rg_wkern *wk, void *grid,
                           hipStream_t s, bool capturing)
{
    const int H = band_height_of(p, sizeof(T) == 4);
    const int type = sizeof(T) == 4 ? RG_C64 : RG_C128;
    const bool fits = plan_fits(ctx->wt_plan, p, vis, wk, nchan, H, type);
    if (capturing) {
        if (!fits)
            return fail(RG_ERR_PARAM,
                        "a W-towers call on a capturing stream needs an earlier eager call of the "
                        "same shape (grid, subgrid, kernels, parameters, no more visibilities) on "
                        "this context");
        return wtowers_batched_t<T>(ctx, p, vis, d_off, nchan, wk, grid, s, true, false);
    }
    if (fits) {
        RG_TRY((wtowers_batched_t<T>(ctx, p, vis, d_off, nchan, wk, grid, s, false, false)));
        unsigned long long over = 0;
        RG_HIP(hipMemcpyAsync(&over, ctx->counters.as<unsigned long long>() + CNT_OVERFLOW,
                              sizeof(over), hipMemcpyDeviceToHost, s));
        RG_HIP(hipStreamSynchronize(s));
        if (over == 0)
            return RG_OK;
        // a capacity of the plan was too small for this call: measure again
        RG_TRY(zero_fill(ctx->counters.p, sizeof(unsigned long long) * CNT__N, s));
    }
    ctx->wt_plan.valid = false;
    return wtowers_batched_t<T>(ctx, p, vis, d_off, nchan, wk, grid, s, false, true);
}

int wtowers_run(rg_context *ctx, const rg_grid_params *p, const rg_flat_vis *vis,
                const int64_t *d_off, int nchan, const rg_wkern *wk, void *grid, int grid_type,
                hipStream_t s, bool capturing)
{
    if (grid_type != RG_C128)
        return fail(RG_ERR_PARAM, "W-towers folds into an FP64 (RG_C128) master grid");
    const int n = p->subgrid_size, S = wk->size;
    if (n + 2 * S >= 65536)
        return fail(RG_ERR_UNSUPPORTED, "subgrid too large");
    if (p->mode == RG_MODE_PLAIN) {
        if (vis->vis_type != RG_C128 || wk->type != RG_C128)
            return fail(RG_ERR_PARAM, "plain mode is FP64: visibilities and kernels must be RG_C128");
        if (nchan != 1)
            return fail(RG_ERR_UNSUPPORTED, "plain mode grids one channel per call");
        if (capturing)
            return fail(RG_ERR_UNSUPPORTED,
                        "plain W-towers (the reference's per-task launches with host read-backs) "
                        "cannot be captured");
        return wtowers_plain(ctx, p, vis, wk, grid, s);
    }
    if (S != 8 && S != 16)
        return fail(RG_ERR_UNSUPPORTED, "batched W-towers supports kernel sizes 8 and 16");
    if (n < 64 || n > 512)
        return fail(RG_ERR_UNSUPPORTED, "batched W-towers supports subgrid sizes 64..512 (got %d)", n);
    static_assert(SORT_BINS >= 512 + 16, "sort bins must cover n + S");
    if (wk->type == RG_C64)
        return wtowers_batched<float>(ctx, p, vis, d_off, nchan, wk, grid, s, capturing);
    return wtowers_batched<double>(ctx, p, vis, d_off, nchan, wk, grid, s, capturing);
}

} // namespace rg
