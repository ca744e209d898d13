// tiles.hpp -- tiled Romein accumulation kernel of batched W-projection
// (device code).
//
// One workgroup of S*S threads takes one work item: up to K sorted records of
// one key.  A key is (group, tile): the group selects the target array (a
// channel grid) and the tile a TILE x TILE block of footprint top-left
// positions, whose footprints cover (TILE + S - 1)^2 cells held in LDS (rows
// TILE + S - 1 rounded up to S apart, so that the binning can hand each record
// its cell and sub-kernel as byte offsets).
// Thread (myU, myV) handles, for every record, the footprint cell congruent to
// (myU, myV) modulo S (Romein); it keeps a register sum while that cell stays
// the same and adds it into LDS when the cell changes.  Two threads never
// share a cell, so the LDS updates need no atomics, and a thread reads a
// cell's LDS value when it starts summing into it: leaving the cell is then a
// store, and the read's latency is hidden behind the records in between.  At the end the tile is
// added to the target with global atomics (neighbouring tiles and items of
// the same tile overlap).  Storing per-item partials and summing them per cell
// instead measured slower at the C2 shape: the atomics overlap with the
// accumulation of other workgroups.
#pragma once

#include "binning.hpp"
#include "romein.hpp"

namespace rg {

template <typename T> __device__ inline void two_sum_into(T &hi, T &lo, T b)
{
#pragma clang fp contract(off)
    const T s = hi + b;
    const T bp = s - hi;
    const T e = (hi - (s - bp)) + (b - bp);
    hi = s;
    lo += e;
}

struct TileSink {
    void *base;              // target arrays, group g at base + g * gsize^2 complex elements
    int gsize;               // edge of a target array (N or n)
    int ntx;                 // tiles per row
    int ntiles;              // tiles per group
    const int32_t *group_map; // optional: group -> rank (slot of the target array)
    int32_t group_offset;     // rank - offset = target slot
};

template <typename T, int S, int TILE, bool COMP, typename GT>
__global__ __launch_bounds__(S *S) void k_tiles(const Rec<T> *__restrict__ recs,
                                                const Item *__restrict__ items, uint32_t item0,
                                                uint32_t item_end,
                                                const unsigned long long *__restrict__ stats,
                                                const typename Vec2<T>::type *__restrict__ kern,
                                                uint32_t kern_bytes, TileSink sink)
{
    using T2 = typename Vec2<T>::type;
    constexpr int NT = S * S;
    constexpr int RT = TILE + S - 1;
    // cell rows RTP apart (RT rounded up to S): a record's (lx, ly) residues
    // follow from its byte offsets
    constexpr int RTP = (RT + S - 1) / S * S;
    constexpr int CELLS = RT * RTP;
    constexpr int z = (int)sizeof(T2);
    static_assert(CELLS * z <= 65536, "cell byte offsets fit 16 bits");
    constexpr int LC = COMP ? CELLS : 1;
    __shared__ T2 s_cell[CELLS]; // (re, im) of every cell of the tile and its halo
    __shared__ T2 s_lo[LC];      // COMP: the two-sum low parts
    __shared__ Rec<T> s_rec[NT];

    const uint32_t end = stats ? (uint32_t)stats[CNT_ITEMS] : item_end;
    const uint32_t it = item0 + blockIdx.x;
    if (it >= end)
        return;
    const Item item = items[it];

    for (int c = threadIdx.x; c < CELLS; c += NT) {
        s_cell[c].x = T(0);
        s_cell[c].y = T(0);
        if (COMP) {
            s_lo[c].x = T(0);
            s_lo[c].y = T(0);
        }
    }

    const int myU = threadIdx.x % S, myV = threadIdx.x / S;
    const int myUz = myU * z, myVSz = myV * S * z;
    char *cell_b = reinterpret_cast<char *>(s_cell);
    char *lo_b = reinterpret_cast<char *>(s_lo);
    auto cellp = [&](int c) { return reinterpret_cast<T2 *>(cell_b + c); };
    auto lop = [&](int c) { return reinterpret_cast<T2 *>(lo_b + c); };
    const __amdgpu_buffer_rsrc_t ktab =
        __builtin_amdgcn_make_buffer_rsrc((void *)kern, 0, (int)kern_bytes, BUF_RSRC_W3);
    int cur = -1;
    T sr = T(0), si = T(0);
    T2 pre; // !COMP: the LDS value of cell cur when the thread started it
    pre.x = T(0);
    pre.y = T(0);
    // adds the register sum into cell cur
    auto flush = [&]() {
        if (COMP) {
            two_sum_into(cellp(cur)->x, lop(cur)->x, sr);
            two_sum_into(cellp(cur)->y, lop(cur)->y, si);
        } else {
            T2 o = pre;
            o.x += sr;
            o.y += si;
            *cellp(cur) = o;
        }
    };

    // Records are staged NT at a time and consumed U at a time: the U tap
    // loads are issued together, then the U updates run in stream order.  A
    // partial batch is padded with zero-weight records (tile corner 0).
    constexpr int U = 8;
    static_assert(NT % U == 0, "record batch must be a multiple of the unroll");
    for (uint32_t b = 0; b < item.count; b += NT) {
        const uint32_t nloc = min((uint32_t)NT, item.count - b);
        const uint32_t nlocU = (nloc + U - 1) / U * U;
        __syncthreads();
        if (threadIdx.x < nloc) {
            s_rec[threadIdx.x] = recs[item.start + b + threadIdx.x];
        } else if (threadIdx.x < nlocU) {
            Rec<T> z;
            z.pos = 0;
            z.kidx = 0;
            z.vis.x = T(0);
            z.vis.y = T(0);
            s_rec[threadIdx.x] = z;
        }
        __syncthreads();
        using B = RecBatch<T, U>;
        auto fetch = [&](uint32_t j, B &bt) {
#pragma unroll
            for (int q = 0; q < U; ++q) {
                // pos = (ly RTP + lx) z | (ly S z) << 16, kidx = sub-kernel byte offset
                const Rec<T> r = s_rec[j + q];
                const int cxz = (myUz - (int)r.pos) & ((S - 1) * z);
                const int cyS = (myVSz - (int)(r.pos >> 16)) & ((S - 1) * S * z);
                bt.cell[q] = (int)(r.pos & 0xffffu) + cyS * (RTP / S) + cxz;
                bt.v[q] = r.vis;
                bt.k[q] = tap_load(ktab, r.kidx + (uint32_t)(cyS + cxz), (T2 *)nullptr);
            }
        };
        auto update = [&](const B &bt) {
#pragma unroll
            for (int q = 0; q < U; ++q) {
                if (bt.cell[q] != cur) {
                    if (cur >= 0)
                        flush();
                    cur = bt.cell[q];
                    if (!COMP)
                        pre = *cellp(cur);
                    sr = T(0);
                    si = T(0);
                }
                // conj(k) * v
                sr = fma(bt.k[q].x, bt.v[q].x, sr);
                sr = fma(bt.k[q].y, bt.v[q].y, sr);
                si = fma(bt.k[q].x, bt.v[q].y, si);
                si = fma(-bt.k[q].y, bt.v[q].x, si);
            }
        };
        pipelined_records<U, B>(nlocU, fetch, update);
    }
    if (cur >= 0)
        flush();
    __syncthreads();

    const uint32_t group = item.key / (uint32_t)sink.ntiles;
    const int tile = (int)(item.key - group * (uint32_t)sink.ntiles);
    const int tx = tile % sink.ntx, ty = tile / sink.ntx;
    const int ox = tx * TILE - S, oy = ty * TILE - S;
    int64_t slot = (int64_t)group;
    if (sink.group_map)
        slot = (int64_t)sink.group_map[group] - sink.group_offset;
    const int G = sink.gsize;
    GT *g = reinterpret_cast<GT *>(sink.base) + (size_t)slot * G * G * 2;
    for (int c = threadIdx.x; c < CELLS; c += NT) {
        const int ry = c / RTP, rx = c - ry * RTP;
        const int gx = ox + rx, gy = oy + ry;
        if (rx < RT && (unsigned)gx < (unsigned)G && (unsigned)gy < (unsigned)G) {
            GT re = (GT)s_cell[c].x, im = (GT)s_cell[c].y;
            if (COMP) {
                re += (GT)s_lo[c].x;
                im += (GT)s_lo[c].y;
            }
            if (re != GT(0) || im != GT(0)) {
                GT *dst = g + 2 * ((size_t)gy * G + gx);
                atomicAdd(dst, re);
                atomicAdd(dst + 1, im);
            }
        }
    }
}

} // namespace rg
