"""Batched W-towers at the edges of its input: an empty stream, a stream whose
visibilities all fall outside the tower set, a single visibility, a smaller
stream after a larger one (the plan's path), and channels of which some are
empty -- each against the oracle (src/reference/grid.c:295-555 restated) or
the exact expectation."""
import numpy as np
import pytest
import torch

import pyoracle as orc
import rgamd
from gpu_util import kern, rel_max, small_obs, stage

pytestmark = pytest.mark.gpu

SUB, MARGIN, WINC = 64, 16, 10.0


@pytest.fixture(scope="module")
def case():
    obs, vis = small_obs(n_st=10, n_dumps=16, arm_m=600.0, lam=2560.0)  # N = 256
    table = rgamd.make_wkernels(11, 8, 16, obs.theta, -5.0, 1.0)
    return obs, vis, table


def params(obs):
    return rgamd.params(obs.grid_size, obs.theta, mode=rgamd.RG_MODE_BATCHED, subgrid=SUB,
                        margin=MARGIN, winc=WINC)


def oracle(obs, u, v, w, vis, table):
    ref, _, _ = orc.grid_wtowers(obs.grid_size, obs.theta, u, v, w, vis, table, -5.0, 1.0, SUB,
                                 MARGIN, WINC, nthreads=8)
    return ref


@pytest.mark.parametrize("dt", [torch.complex128, torch.complex64])
def test_empty_stream_zeroes_the_grid(ctx, case, dt):
    obs, vis, table = case
    N = obs.grid_size
    e = np.zeros(0)
    _, fv = stage(e, e, e, np.zeros(0, np.complex128), dt)
    _, wk = kern(table, -5.0, 1.0, dt)
    g = torch.full((N, N), 5.0 + 1.0j, dtype=torch.complex128, device="cuda")
    ctx.grid(params(obs), fv, wk, g)
    st = ctx.stats()
    torch.cuda.synchronize()
    assert float(g.abs().max()) == 0.0
    assert st.n_vis == 0 and st.n_gridded == 0 and st.n_tasks == 0


def test_all_visibilities_outside_the_towers(ctx, case):
    obs, vis, table = case
    N = obs.grid_size
    far = np.full(obs.n_vis, 1e7)  # beyond the grid in u and v
    _, fv = stage(far, far, obs.w, vis)
    _, wk = kern(table, -5.0, 1.0)
    g = torch.full((N, N), 2.0 + 0.0j, dtype=torch.complex128, device="cuda")
    ctx.grid(params(obs), fv, wk, g)
    st = ctx.stats()
    torch.cuda.synchronize()
    assert float(g.abs().max()) == 0.0
    assert st.n_dropped == obs.n_vis and st.n_gridded == 0


@pytest.mark.parametrize("dt,tol", [(torch.complex128, 1e-11), (torch.complex64, 1e-5)])
def test_single_visibility(ctx, case, dt, tol):
    obs, vis, table = case
    i = obs.n_vis // 3
    u, v, w, x = obs.u[i:i + 1], obs.v[i:i + 1], obs.w[i:i + 1], vis[i:i + 1]
    ref = oracle(obs, u, v, w, x, table)
    _, fv = stage(u, v, w, x, dt)
    _, wk = kern(table, -5.0, 1.0, dt)
    g = torch.empty((obs.grid_size, obs.grid_size), dtype=torch.complex128, device="cuda")
    ctx.grid(params(obs), fv, wk, g)
    st = ctx.stats()
    torch.cuda.synchronize()
    assert st.n_gridded == 1 and st.n_tasks == 1
    assert rel_max(g.cpu().numpy(), ref) < tol


def test_smaller_stream_after_larger_uses_the_plan(case):
    """The first call makes the plan; a call with half the visibilities fits
    it (no re-plan) and still equals the oracle."""
    obs, vis, table = case
    ctx = rgamd.Context(0)
    try:
        _, wk = kern(table, -5.0, 1.0)
        g = torch.empty((obs.grid_size, obs.grid_size), dtype=torch.complex128, device="cuda")
        _, fv = stage(obs.u, obs.v, obs.w, vis)
        ctx.grid(params(obs), fv, wk, g)
        h = obs.n_vis // 2
        _, fvh = stage(obs.u[h:], obs.v[h:], obs.w[h:], vis[h:])
        ctx.grid(params(obs), fvh, wk, g)
        torch.cuda.synchronize()
        ref = oracle(obs, obs.u[h:], obs.v[h:], obs.w[h:], vis[h:], table)
        assert rel_max(g.cpu().numpy(), ref) < 1e-11
    finally:
        ctx.close()


def test_channels_with_empty_ones(ctx, case):
    """Four channels over the stream, the second and the last empty: empty
    channels give zero grids, the others equal per-channel oracle runs."""
    obs, vis, table = case
    N, n = obs.grid_size, obs.n_vis
    off = np.array([0, n // 3, n // 3, n, n], dtype=np.int64)
    _, fv = stage(obs.u, obs.v, obs.w, vis)
    _, wk = kern(table, -5.0, 1.0)
    g = torch.full((4, N, N), 1.0 + 1.0j, dtype=torch.complex128, device="cuda")
    ctx.grid_channels(params(obs), fv, off, wk, g)
    st = ctx.stats()
    torch.cuda.synchronize()
    got = g.cpu().numpy()
    assert st.n_gridded == n
    for c in range(4):
        sl = slice(off[c], off[c + 1])
        if off[c] == off[c + 1]:
            assert np.abs(got[c]).max() == 0.0, c
        else:
            ref = oracle(obs, obs.u[sl], obs.v[sl], obs.w[sl], vis[sl], table)
            assert rel_max(got[c], ref) < 1e-11, c
