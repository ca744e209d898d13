"""Batched W-towers built on the device: a call captured in a HIP graph and
replayed (against the eager result and the oracle), calls whose data outgrow
the context's plan (re-planned, still equal to the oracle), the frozen
context of a captured call, and many channels in one call (the 256-channel
workload's shape at reduced size: per-channel 1024^2 grids, theta 0.05)
against per-channel oracle runs of src/reference/grid.c:295-555."""
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
    table = rgamd.make_wkernels(11, 8, 16, obs.theta, -5.0, 1.0)
    ref, _, bad = orc.grid_wtowers(obs.grid_size, obs.theta, obs.u, obs.v, obs.w, vis, table,
                                   -5.0, 1.0, SUB, MARGIN, WINC, nthreads=8)
    assert bad == 0
    return obs, vis, table, ref


@pytest.fixture()
def own_ctx():
    c = rgamd.Context(0)
    yield c
    c.close()


def params(obs, **kw):
    return rgamd.params(obs.grid_size, obs.theta, mode=rgamd.RG_MODE_BATCHED, subgrid=SUB,
                        margin=MARGIN, winc=WINC, **kw)


# replay against eager: the counting sort and the band pass's column sort
# place records with atomics, so the order of the sums inside a cell (and the
# FP32 rounding) changes from call to call; FP64 agrees to ~1e-15, FP32 to a
# few 1e-7 of the grid's peak
@pytest.mark.parametrize("dt,tol", [(torch.complex128, 1e-12), (torch.complex64, 5e-6)])
def test_graph_replay_matches_eager_and_oracle(own_ctx, case, dt, tol):
    obs, vis, table, ref = case
    _, fv = stage(obs.u, obs.v, obs.w, vis, dt)
    _, wk = kern(table, -5.0, 1.0, dt)
    N = obs.grid_size
    p = params(obs)
    eager = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    own_ctx.grid(p, fv, wk, eager)  # makes the plan and sizes the scratch memory
    own_ctx.grid(p, fv, wk, eager)  # plan path
    st = own_ctx.stats()
    torch.cuda.synchronize()
    assert st.n_gridded == obs.n_vis and st.n_tasks > 1
    e = eager.cpu().numpy()
    if dt == torch.complex128:
        assert rel_max(e, ref) < 1e-11
    else:
        assert rel_max(orc.grid_to_image(e), orc.grid_to_image(ref)) < 1e-5
    g = torch.zeros_like(eager)
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            own_ctx.grid(p, fv, wk, g, stream=s.cuda_stream)
    torch.cuda.current_stream().wait_stream(s)
    for rep in range(3):
        g.fill_(3.0)  # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = g.cpu().numpy()
        assert rel_max(got, e) < tol, rep
        st = own_ctx.stats()  # no capacity overrun in the replay
        assert st.n_tasks > 1
    del graph
    own_ctx.thaw()


def test_plan_regrows_for_larger_inputs(own_ctx, case):
    """A second call with twice the floors (w doubled) and a third with more
    visibilities outgrow the plan made by the first: each is redone with a
    fresh plan and equals the oracle."""
    obs, vis, table, ref = case
    _, wk = kern(table, -5.0, 1.0)
    N = obs.grid_size
    p = params(obs)
    g = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    _, fv = stage(obs.u, obs.v, obs.w, vis)
    own_ctx.grid(p, fv, wk, g)
    torch.cuda.synchronize()
    assert rel_max(g.cpu().numpy(), ref) < 1e-11
    w2 = obs.w * 2.0
    ref2, _, _ = orc.grid_wtowers(N, obs.theta, obs.u, obs.v, w2, vis, table, -5.0, 1.0, SUB,
                                  MARGIN, WINC, nthreads=8)
    _, fv2 = stage(obs.u, obs.v, w2, vis)
    own_ctx.grid(p, fv2, wk, g)
    st = own_ctx.stats()
    torch.cuda.synchronize()
    assert rel_max(g.cpu().numpy(), ref2) < 1e-11
    assert st.n_gridded == obs.n_vis
    u3, v3, w3 = (np.concatenate([a, a[::-1]]) for a in (obs.u, obs.v, obs.w))
    vis3 = np.concatenate([vis, 0.5 * vis[::-1]])
    ref3, _, _ = orc.grid_wtowers(N, obs.theta, u3, v3, w3, vis3, table, -5.0, 1.0, SUB, MARGIN,
                                  WINC, nthreads=8)
    _, fv3 = stage(u3, v3, w3, vis3)
    own_ctx.grid(p, fv3, wk, g)
    torch.cuda.synchronize()
    assert rel_max(g.cpu().numpy(), ref3) < 1e-11


def test_frozen_context_refuses_to_reallocate(own_ctx, case):
    obs, vis, table, ref = case
    _, wk = kern(table, -5.0, 1.0)
    N = obs.grid_size
    p = params(obs)
    half = obs.n_vis // 2
    _, fv_small = stage(obs.u[:half], obs.v[:half], obs.w[:half], vis[:half])
    g = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    own_ctx.grid(p, fv_small, wk, g)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            own_ctx.grid(p, fv_small, wk, g, stream=s.cuda_stream)
    torch.cuda.current_stream().wait_stream(s)
    # a longer stream needs more scratch memory than the graph's buffers
    u3, v3, w3 = (np.concatenate([a] * 4) for a in (obs.u, obs.v, obs.w))
    _, fv_big = stage(u3, v3, w3, np.concatenate([vis] * 4))
    with pytest.raises(rgamd.RGError) as err:
        own_ctx.grid(p, fv_big, wk, g)
    assert err.value.code == rgamd.RG_ERR_PARAM and "rg_context_thaw" in str(err.value)
    graph.replay()  # the graph's memory is intact
    torch.cuda.synchronize()
    ref_half, _, _ = orc.grid_wtowers(N, obs.theta, obs.u[:half], obs.v[:half], obs.w[:half],
                                      vis[:half], table, -5.0, 1.0, SUB, MARGIN, WINC, nthreads=8)
    assert rel_max(g.cpu().numpy(), ref_half) < 1e-11
    del graph
    own_ctx.thaw()
    own_ctx.grid(p, fv_big, wk, g)
    torch.cuda.synchronize()
    assert rel_max(g.cpu().numpy(), 4 * ref) < 1e-11


@pytest.fixture(scope="module")
def channels():
    """12 channels over 100-200 MHz with uneven counts, 1024^2 grids (theta
    0.05), W-towers subgrid 256, margin 16, winc 10."""
    obs = synth.config_c4(n_st=24, n_chan=12, n_dumps=16, median=300.0, sigma=1.0)
    vis = synth.predict_numpy(obs.u, obs.v, obs.w, obs.l, obs.m, obs.flux)
    table = rgamd.make_wkernels(11, 8, 16, obs.theta, -5.0, 1.0)
    off = obs.chan_offsets
    refs = []
    for c in range(len(off) - 1):
        sl = slice(off[c], off[c + 1])
        r, _, bad = orc.grid_wtowers(obs.grid_size, obs.theta, obs.u[sl], obs.v[sl], obs.w[sl],
                                     vis[sl], table, -5.0, 1.0, 256, MARGIN, WINC, nthreads=8)
        assert bad == 0
        refs.append(r)
    return obs, vis, table, np.stack(refs)


def run_channels(ctx, obs, vis, table, dt, **kw):
    _, fv = stage(obs.u, obs.v, obs.w, vis, dt)
    _, wk = kern(table, -5.0, 1.0, dt)
    nc, N = len(obs.chan_offsets) - 1, obs.grid_size
    g = torch.empty((nc, N, N), dtype=torch.complex128, device="cuda")
    p = rgamd.params(N, obs.theta, mode=rgamd.RG_MODE_BATCHED, subgrid=256, margin=MARGIN,
                     winc=WINC, **kw)
    ctx.grid_channels(p, fv, obs.chan_offsets, wk, g)
    st = ctx.stats()
    torch.cuda.synchronize()
    return g.cpu().numpy(), st, (fv, wk, p, g)


def test_channels_fp64_match_oracle(own_ctx, channels):
    obs, vis, table, refs = channels
    assert obs.grid_size == 1024 and len(refs) == 12
    g, st, _ = run_channels(own_ctx, obs, vis, table, torch.complex128)
    for c in range(len(refs)):
        assert rel_max(g[c], refs[c]) < 1e-11, c
    assert st.n_gridded == obs.n_vis and st.n_dropped == 0


@pytest.mark.parametrize("band,per_item,floor_batch", [(8, 0, 0), (16, 37, 0), (4, 0, 7)])
def test_channels_fp32_within_tolerance(own_ctx, channels, band, per_item, floor_batch):
    obs, vis, table, refs = channels
    g, _, _ = run_channels(own_ctx, obs, vis, table, torch.complex64, tile=band,
                           vis_per_item=per_item, floor_batch=floor_batch)
    for c in range(len(refs)):
        img, img_ref = orc.grid_to_image(g[c]), orc.grid_to_image(refs[c])
        assert rel_max(img, img_ref) < 1e-5 and rel_rms(img, img_ref) < 1e-5, c


def test_channels_graph_replay(own_ctx, channels):
    obs, vis, table, refs = channels
    eager, _, (fv, wk, p, g) = run_channels(own_ctx, obs, vis, table, torch.complex128)
    off = np.ascontiguousarray(obs.chan_offsets, dtype=np.int64)
    out = torch.zeros_like(g)
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            own_ctx.grid_channels(p, fv, off, wk, out, stream=s.cuda_stream)
    torch.cuda.current_stream().wait_stream(s)
    for rep in range(3):
        out.fill_(2.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert rel_max(got, eager) < 1e-12, rep
        assert rel_max(got[3], refs[3]) < 1e-11, rep
    del graph
    own_ctx.thaw()
