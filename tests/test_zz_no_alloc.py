"""rg_grid_noalloc (batched W-projection through librg's static scratch, no
allocation) against rg_grid on the same inputs: two calls in a row, the
second with a smaller stream (the static scratch keeps the first call's
contents), FP64 and FP32 kernels, and a stream split over two calls with
accumulate = 1."""
import numpy as np
import pytest
import torch

import rgamd
import synth
from gpu_util import kern, rel_max, rel_rms, stage

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    """27 stations on a random Y, 256 dumps, 2048^2 grid, S16 x R8 kernels at
    w-step 10 (about 66 000 visibilities after the 10 km cut)."""
    obs = synth.config_c1()
    vis = synth.predict_numpy(obs.u, obs.v, obs.w, obs.l, obs.m, obs.flux)
    w_min, P = synth.kernel_range(obs.w, 10.0)
    table = rgamd.make_wkernels(P, 8, 16, obs.theta, w_min, 10.0)
    return obs, vis, table, w_min


def sub(obs, vis, lo, hi):
    return obs.u[lo:hi], obs.v[lo:hi], obs.w[lo:hi], vis[lo:hi]


def both(ctx, obs, vis, table, w_min, lo, hi, dt, gdt, **kw):
    u, v, w, x = sub(obs, vis, lo, hi)
    _, fv = stage(u, v, w, x, dt)
    _, wk = kern(table, w_min, 10.0, dt)
    N = obs.grid_size
    p = rgamd.params(N, obs.theta, mode=rgamd.RG_MODE_BATCHED, **kw)
    g_ref = torch.empty((N, N), dtype=gdt, device="cuda")
    ctx.grid(p, fv, wk, g_ref)
    g_na = torch.full((N, N), 7.0 + 3.0j, dtype=gdt, device="cuda")  # zeroed by the call
    rgamd.grid_noalloc(p, fv, wk, g_na)
    torch.cuda.synchronize()
    return g_ref.cpu().numpy().astype(np.complex128), g_na.cpu().numpy().astype(np.complex128)


# the counting sort places records with atomics, so the order of the sums in a
# tile (and the FP32 rounding) can change from call to call
@pytest.mark.parametrize("dt,gdt,tol", [(torch.complex128, torch.complex128, 1e-12),
                                        (torch.complex64, torch.complex64, 5e-6)])
def test_two_calls_match_rg_grid(ctx, c1, dt, gdt, tol):
    obs, vis, table, w_min = c1
    n = obs.n_vis
    print(f"[no_alloc] {n} visibilities, {dt}", flush=True)
    ref1, na1 = both(ctx, obs, vis, table, w_min, 0, n, dt, gdt)
    assert np.abs(ref1).max() > 0
    assert rel_max(na1, ref1) < tol and rel_rms(na1, ref1) < tol
    # second call, a smaller stream: the scratch still holds the first call's
    # counters and records
    ref2, na2 = both(ctx, obs, vis, table, w_min, 0, n // 3, dt, gdt, tile=48, vis_per_item=512)
    assert rel_max(na2, ref2) < tol and rel_rms(na2, ref2) < tol
    assert rel_max(ref2, ref1) > 1e-3  # a different result, so the grid was re-zeroed


def test_split_stream_accumulates(ctx, c1):
    obs, vis, table, w_min = c1
    n = obs.n_vis
    N = obs.grid_size
    _, wk = kern(table, w_min, 10.0, torch.complex128)
    p = rgamd.params(N, obs.theta, mode=rgamd.RG_MODE_BATCHED)
    _, fv_all = stage(obs.u, obs.v, obs.w, vis)
    g_ref = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    ctx.grid(p, fv_all, wk, g_ref)
    h = n // 2 + 123
    _, fv_a = stage(*sub(obs, vis, 0, h))
    _, fv_b = stage(*sub(obs, vis, h, n))
    g = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    rgamd.grid_noalloc(p, fv_a, wk, g)
    rgamd.grid_noalloc(p, fv_b, wk, g, accumulate=True)
    torch.cuda.synchronize()
    assert rel_max(g.cpu().numpy(), g_ref.cpu().numpy()) < 1e-12
