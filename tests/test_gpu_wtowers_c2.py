"""W-towers at the 512-station workload's own shape against the oracle.

Geometry exactly as the headline workload: 2048^2 grid (theta 0.1, lambda
20480), subgrid 512, margin 16, winc 10, W-kernels support 16 x oversample 8
over the residual range [-5, 5] (11 planes, step 1).  The station count is
reduced (48 stations of the same Exp(3 km) radial family, 16 dumps over 4 h,
baselines < 10 km, 50 sources) so that the oracle's W-towers
(src/reference/grid.c:295-555 restated) finishes in seconds; the floor range
(about 400 floors) and the set of active towers are those of the full
workload.  Every subgrid-512 kernel instance of the batched path runs here:
the band pass at 4, 8 and 16 rows (FP32) and 4 and 8 rows (FP64), its
spill/combine path (small work items), several task batches, and the plain
FP64 path of the reference's launch structure (wtowers_CUDA_flat,
grid_cuda.cu:522-746)."""
import numpy as np
import pytest
import torch

import pyoracle as orc
import rgamd
import synth
from gpu_util import kern, rel_max, rel_rms, stage

pytestmark = pytest.mark.gpu

N, SUB, MARGIN, WINC = 2048, 512, 16, 10.0
W_MIN, W_STEP, PLANES = -5.0, 1.0, 11


@pytest.fixture(scope="module")
def c2():
    obs = synth.config_c2(n_st=48, n_dumps=16)
    assert obs.grid_size == N
    vis = synth.predict_numpy(obs.u, obs.v, obs.w, obs.l, obs.m, obs.flux)
    table = rgamd.make_wkernels(PLANES, 8, 16, obs.theta, W_MIN, W_STEP)
    ref, fl, bad = orc.grid_wtowers(N, obs.theta, obs.u, obs.v, obs.w, vis, table, W_MIN, W_STEP,
                                    SUB, MARGIN, WINC, nthreads=16)
    # every visibility gridded exactly once: 8 S^2 flop each (grid.c:519-530)
    assert bad == 0 and fl == 8 * 256 * obs.n_vis
    chunk = SUB - MARGIN
    tasks = {(orc.tower_index(u, obs.theta, N, chunk), orc.tower_index(v, obs.theta, N, chunk),
              orc.floor_index(w, WINC)) for u, v, w in zip(obs.u, obs.v, obs.w)}
    return obs, vis, table, ref, orc.grid_to_image(ref), tasks


def grid(ctx, obs, vis, table, mode, dt, **kw):
    _, fv = stage(obs.u, obs.v, obs.w, vis, dt)
    _, wk = kern(table, W_MIN, W_STEP, dt)
    g = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    ctx.grid(rgamd.params(N, obs.theta, mode=mode, subgrid=SUB, margin=MARGIN, winc=WINC, **kw),
             fv, wk, g)
    st = ctx.stats()
    torch.cuda.synchronize()
    return g.cpu().numpy(), st


def test_workload_shape(c2):
    obs, _, _, _, _, tasks = c2
    floors = np.floor(obs.w / WINC + 0.5)
    assert floors.max() - floors.min() > 300  # hundreds of floors, as the full workload
    assert len({(a, b) for a, b, _ in tasks}) >= 9  # the active towers of the full workload
    assert len(tasks) > 1000


def test_plain_fp64(ctx, c2):
    obs, vis, table, ref, _, tasks = c2
    g, st = grid(ctx, obs, vis, table, rgamd.RG_MODE_PLAIN, torch.complex128)
    assert rel_max(g, ref) < 1e-11
    assert st.n_tasks == len(tasks) and st.n_gridded == obs.n_vis


@pytest.mark.parametrize("band,per_item,floor_batch,radix", [(8, 0, 0, 0), (4, 0, 0, 0),
                                                             (8, 37, 97, 0), (4, 61, 500, 0),
                                                             (8, 0, 0, 4)])
def test_batched_fp64(ctx, c2, band, per_item, floor_batch, radix):
    """FP64 bands of 8 and 4 rows (n = 512 takes 81 KB of LDS per workgroup at
    8 rows); per_item 37 / 61 split most band keys into spilled parts, and
    floor_batch < tasks runs several batches; radix 4 runs the column pass's
    FFT in stages 4, 4, 4, 4, 2."""
    obs, vis, table, ref, _, tasks = c2
    g, st = grid(ctx, obs, vis, table, rgamd.RG_MODE_BATCHED, torch.complex128, tile=band,
                 vis_per_item=per_item, floor_batch=floor_batch, fft_radix=radix)
    assert rel_max(g, ref) < 1e-11, (band, per_item, floor_batch)
    assert st.n_tasks == len(tasks) and st.n_gridded == obs.n_vis and st.n_dropped == 0


@pytest.mark.parametrize("band,per_item,floor_batch,radix,comp", [
    (8, 0, 0, 0, 0), (16, 0, 0, 0, 0), (4, 0, 0, 0, 0), (8, 37, 0, 0, 0), (16, 37, 250, 0, 0),
    (4, 53, 97, 0, 0), (16, 0, 0, 4, 0), (8, 0, 0, 0, 1), (16, 37, 250, 0, 1), (4, 53, 97, 0, 1)])
def test_batched_fp32(ctx, c2, band, per_item, floor_batch, radix, comp):
    """FP32 bands of 8 (default), 16 and 4 rows: the image within 1e-5 of the
    FP64 oracle's (max and RMS relative), also with split keys, batches,
    radix-4 column FFTs and two-sum compensated register sums."""
    obs, vis, table, _, img_ref, tasks = c2
    g, st = grid(ctx, obs, vis, table, rgamd.RG_MODE_BATCHED, torch.complex64, tile=band,
                 vis_per_item=per_item, floor_batch=floor_batch, fft_radix=radix,
                 compensated=comp)
    img = orc.grid_to_image(g)
    assert rel_max(img, img_ref) < 1e-5, (band, per_item, floor_batch, comp)
    assert rel_rms(img, img_ref) < 1e-5, (band, per_item, floor_batch, comp)
    assert st.n_tasks == len(tasks) and st.n_gridded == obs.n_vis


def test_fp32_batched_vs_fp64_plain(ctx, c2):
    """The metric's own gate on this shape: FP32 batched against FP64 plain on
    identical inputs and kernels, max and RMS relative image error <= 1e-5."""
    obs, vis, table, _, _, _ = c2
    g64, _ = grid(ctx, obs, vis, table, rgamd.RG_MODE_PLAIN, torch.complex128)
    g32, _ = grid(ctx, obs, vis, table, rgamd.RG_MODE_BATCHED, torch.complex64)
    i64, i32 = orc.grid_to_image(g64), orc.grid_to_image(g32)
    assert rel_max(i32, i64) < 1e-5 and rel_rms(i32, i64) < 1e-5
