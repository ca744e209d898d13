"""W-towers on the GPU (plain and batched) against the oracle."""
import numpy as np
import pytest
import torch

import pyoracle as orc
import rgamd
import synth
from gpu_util import kern, rel_max, rel_rms, small_obs, stage

pytestmark = pytest.mark.gpu

SUB, MARGIN, WINC = 64, 16, 10.0


@pytest.fixture(scope="module")
def case():
    obs, vis = small_obs(n_st=14, n_dumps=32, arm_m=700.0, lam=2560.0)  # N = 256
    w_step = 1.0
    table = rgamd.make_wkernels(11, 8, 16, obs.theta, -5.0, w_step)
    ref, fl, bad = orc.grid_wtowers(obs.grid_size, obs.theta, obs.u, obs.v, obs.w, vis, table,
                                    -5.0, w_step, SUB, MARGIN, WINC, nthreads=8)
    assert bad == 0 and fl == 8 * 256 * obs.n_vis
    return obs, vis, table, ref


def run(ctx, obs, vis, table, mode, dt=torch.complex128, **kw):
    _, fv = stage(obs.u, obs.v, obs.w, vis, dt)
    kt, wk = kern(table, -5.0, 1.0, dt)
    N = obs.grid_size
    g = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    ctx.grid(rgamd.params(N, obs.theta, mode=mode, subgrid=SUB, margin=MARGIN, winc=WINC, **kw),
             fv, wk, g)
    st = ctx.stats()
    torch.cuda.synchronize()
    return g.cpu().numpy(), st


def test_plain_fp64_matches_oracle(ctx, case):
    obs, vis, table, ref = case
    g, st = run(ctx, obs, vis, table, rgamd.RG_MODE_PLAIN)
    assert rel_max(g, ref) < 1e-11
    assert st.n_tasks > 1 and st.n_towers >= 1


@pytest.mark.parametrize("floor_batch,band,radix", [(1, 8, 0), (5, 8, 0), (128, 8, 0), (5, 4, 0),
                                                   (5, 8, 4), (0, 4, 4)])
def test_batched_fp64_matches_oracle(ctx, case, floor_batch, band, radix):
    """radix 4: the column pass's FFT in radix-4 stages (4, 4, 4 at n = 64)."""
    obs, vis, table, ref = case
    g, st = run(ctx, obs, vis, table, rgamd.RG_MODE_BATCHED, floor_batch=floor_batch,
                vis_per_item=50, tile=band, fft_radix=radix)
    assert rel_max(g, ref) < 1e-11
    assert st.n_gridded == obs.n_vis and st.n_dropped == 0


@pytest.mark.parametrize("band,per_item,comp", [(8, 0, 0), (8, 37, 0), (16, 0, 0), (16, 37, 0),
                                               (4, 0, 0), (4, 37, 0), (8, 37, 1), (16, 0, 1),
                                               (4, 37, 1)])
def test_batched_fp32_within_tolerance(ctx, case, band, per_item, comp):
    """FP32 bands of 8 / 16 / 4 rows; per_item 37 splits most band keys over
    several work items (spilled parts summed by the combine pass); comp: two-sum
    compensated register sums."""
    obs, vis, table, ref = case
    g, st = run(ctx, obs, vis, table, rgamd.RG_MODE_BATCHED, dt=torch.complex64, tile=band,
                vis_per_item=per_item, compensated=comp)
    img, img_ref = orc.grid_to_image(g), orc.grid_to_image(ref)
    assert rel_max(img, img_ref) < 1e-5 and rel_rms(img, img_ref) < 1e-5
    assert st.n_gridded == obs.n_vis


def test_task_count_matches_binning(ctx, case):
    obs, vis, table, ref = case
    _, st = run(ctx, obs, vis, table, rgamd.RG_MODE_BATCHED)
    N, chunk = obs.grid_size, SUB - MARGIN
    tasks = {(orc.tower_index(obs.u[i], obs.theta, N, chunk),
              orc.tower_index(obs.v[i], obs.theta, N, chunk),
              orc.floor_index(obs.w[i], WINC)) for i in range(obs.n_vis)}
    assert st.n_tasks == len(tasks)
    assert st.n_towers == len({(a, b) for a, b, _ in tasks})


def test_fp64_required_for_plain(ctx, case):
    obs, vis, table, ref = case
    with pytest.raises(rgamd.RGError) as e:
        run(ctx, obs, vis, table, rgamd.RG_MODE_PLAIN, dt=torch.complex64)
    assert e.value.code == rgamd.RG_ERR_PARAM


def test_stress_shape_matches_oracle(ctx):
    """The stress-workload shape at reduced size: 4096^2 grid (theta 0.2),
    subgrid 256, margin 16, winc 10, phase centre at zenith angle 60 deg
    (|w| up to thousands of wavelengths, hundreds of floors), three channels."""
    obs = synth.config_c5(n_st=20, n_chan=3, n_dumps=24)
    vis = synth.predict_numpy(obs.u, obs.v, obs.w, obs.l, obs.m, obs.flux)
    sub, margin, winc = 256, 16, 10.0
    table = rgamd.make_wkernels(11, 8, 16, obs.theta, -5.0, 1.0)
    ref, fl, bad = orc.grid_wtowers(obs.grid_size, obs.theta, obs.u, obs.v, obs.w, vis, table,
                                    -5.0, 1.0, sub, margin, winc, nthreads=8)
    assert bad == 0 and obs.grid_size == 4096
    N = obs.grid_size
    for dt, tol in ((torch.complex128, 1e-11), (torch.complex64, 1e-5)):
        _, fv = stage(obs.u, obs.v, obs.w, vis, dt)
        kt, wk = kern(table, -5.0, 1.0, dt)
        g = torch.empty((N, N), dtype=torch.complex128, device="cuda")
        ctx.grid(rgamd.params(N, obs.theta, mode=rgamd.RG_MODE_BATCHED, subgrid=sub,
                              margin=margin, winc=winc), fv, wk, g)
        st = ctx.stats()
        torch.cuda.synchronize()
        got = g.cpu().numpy()
        assert rel_max(got, ref) < tol, dt
        assert st.n_gridded == obs.n_vis and st.n_tasks > 50
        del g


@pytest.mark.parametrize("sub,margin,S", [(128, 16, 8), (128, 8, 8), (256, 16, 16)])
@pytest.mark.parametrize("radix", [8, 4])
def test_other_shapes_match_oracle(ctx, sub, margin, S, radix):
    """Kernel support 8 and subgrids 128 / 256 through the batched path, the
    column pass's FFT in radix-8 stages (8, 8, 2 / 8, 8, 4) or radix-4 stages
    (4, 4, 4, 2 / 4, 4, 4, 4)."""
    obs, vis = small_obs(n_st=14, n_dumps=24, arm_m=900.0, lam=5120.0)  # N = 512
    table = rgamd.make_wkernels(11, 4, S, obs.theta, -5.0, 1.0)
    ref, _, bad = orc.grid_wtowers(obs.grid_size, obs.theta, obs.u, obs.v, obs.w, vis, table,
                                   -5.0, 1.0, sub, margin, WINC, nthreads=8)
    assert bad == 0
    N = obs.grid_size
    for dt, tol in ((torch.complex128, 1e-11), (torch.complex64, 1e-5)):
        _, fv = stage(obs.u, obs.v, obs.w, vis, dt)
        kt, wk = kern(table, -5.0, 1.0, dt)
        g = torch.empty((N, N), dtype=torch.complex128, device="cuda")
        ctx.grid(rgamd.params(N, obs.theta, mode=rgamd.RG_MODE_BATCHED, subgrid=sub,
                              margin=margin, winc=WINC, vis_per_item=97, fft_radix=radix),
                 fv, wk, g)
        st = ctx.stats()
        torch.cuda.synchronize()
        assert rel_max(g.cpu().numpy(), ref) < tol, (dt, sub, S)
        assert st.n_gridded == obs.n_vis
