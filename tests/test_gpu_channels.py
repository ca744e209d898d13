"""Many channels with uneven counts in one batched call (the C4 shape at
reduced size): parity with per-channel oracle runs, replay from a captured
graph, and the autotuner."""
import numpy as np
import pytest
import torch

import pyoracle as orc
import rgamd
import synth
from gpu_util import kern, rel_max, stage

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    obs = synth.config_c4(n_st=24, n_chan=12, n_dumps=16, median=300.0, sigma=1.0)
    vis = synth.predict_numpy(obs.u, obs.v, obs.w, obs.l, obs.m, obs.flux)
    w_step = 10.0
    w_min, P = synth.kernel_range(obs.w, w_step)
    table = rgamd.make_wkernels(P, 8, 16, obs.theta, w_min, w_step)
    off = obs.chan_offsets
    refs = []
    for c in range(len(off) - 1):
        sl = slice(off[c], off[c + 1])
        r, _, bad = orc.grid_wprojection(obs.grid_size, obs.theta, obs.u[sl], obs.v[sl], obs.w[sl],
                                         vis[sl], table, w_min, w_step)
        assert bad == 0
        refs.append(r)
    return obs, vis, table, w_min, w_step, np.stack(refs)


def run_channels(ctx, obs, vis, table, w_min, w_step, vdt, kdt, gdt, **kw):
    _, fv = stage(obs.u, obs.v, obs.w, vis, vdt)
    _, wk = kern(table, w_min, w_step, kdt)
    nc, N = len(obs.chan_offsets) - 1, obs.grid_size
    g = torch.empty((nc, N, N), dtype=gdt, device="cuda")
    ctx.grid_channels(rgamd.params(N, obs.theta, **kw), fv, obs.chan_offsets, wk, g)
    torch.cuda.synchronize()
    return g.cpu().numpy().astype(np.complex128)


def test_counts_are_uneven(case):
    obs = case[0]
    counts = np.diff(obs.chan_offsets)
    assert counts.min() > 0 and counts.max() > 3 * counts.min()
    assert obs.grid_size == 1024


def test_channels_fp64_match_oracle(ctx, case):
    obs, vis, table, w_min, w_step, refs = case
    g = run_channels(ctx, obs, vis, table, w_min, w_step, torch.complex128, torch.complex128,
                     torch.complex128)
    for c in range(len(refs)):
        assert rel_max(g[c], refs[c]) < 1e-12, c


@pytest.mark.parametrize("tile,per_item", [(32, 0), (48, 300)])
def test_channels_fp32_within_tolerance(ctx, case, tile, per_item):
    obs, vis, table, w_min, w_step, refs = case
    g = run_channels(ctx, obs, vis, table, w_min, w_step, torch.complex64, torch.complex64,
                     torch.complex64, tile=tile, vis_per_item=per_item)
    for c in range(len(refs)):
        assert rel_max(g[c], refs[c]) < 1e-5, c


def test_graph_replay_matches_eager(ctx, case):
    """The batched multi-channel call captured once in a graph and replayed
    several times, with eager work between replays; every replay matches the
    eager result."""
    obs, vis, table, w_min, w_step, refs = case
    _, fv = stage(obs.u, obs.v, obs.w, vis, torch.complex128)
    _, wk = kern(table, w_min, w_step, torch.complex128)
    nc, N = len(obs.chan_offsets) - 1, obs.grid_size
    off = np.ascontiguousarray(obs.chan_offsets, dtype=np.int64)  # read when the graph runs
    p = rgamd.params(N, obs.theta)
    eager = torch.empty((nc, N, N), dtype=torch.complex128, device="cuda")
    ctx.grid_channels(p, fv, off, wk, eager)  # also sizes the context's scratch memory
    torch.cuda.synchronize()
    g = torch.zeros_like(eager)
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            ctx.grid_channels(p, fv, off, wk, g, stream=s.cuda_stream)
    torch.cuda.current_stream().wait_stream(s)
    ref = eager.cpu().numpy()
    for rep in range(3):
        g.fill_(1.0)  # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert rel_max(g.cpu().numpy(), ref) < 1e-12, rep
        assert rel_max(g.cpu().numpy()[0], refs[0]) < 1e-12, rep
    del graph
    ctx.thaw()  # the shared context may grow its scratch memory again


def test_autotune_picks_a_candidate(ctx, case):
    obs, vis, table, w_min, w_step, refs = case
    _, fv = stage(obs.u, obs.v, obs.w, vis, torch.complex64)
    _, wk = kern(table, w_min, w_step, torch.complex64)
    nc, N = len(obs.chan_offsets) - 1, obs.grid_size
    g = torch.empty((nc, N, N), dtype=torch.complex64, device="cuda")
    cands = [{"tile": 32, "vis_per_item": 256}, {"tile": 48, "vis_per_item": 1024}]
    best, timings = rgamd.autotune(ctx, rgamd.params(N, obs.theta), fv, wk, g, cands,
                                   offsets=obs.chan_offsets, reps=2)
    assert len(timings) == 2 and all(ms > 0 for _, ms in timings)
    assert (best.tile, best.vis_per_item) in {(c["tile"], c["vis_per_item"]) for c in cands}
    ctx.grid_channels(best, fv, obs.chan_offsets, wk, g)
    torch.cuda.synchronize()
    got = g.cpu().numpy()
    for c in range(nc):
        assert rel_max(got[c], refs[c]) < 1e-5, c
