"""W-projection on the GPU (plain and batched) against the oracle."""
import numpy as np
import pytest
import torch

import pyoracle as orc
import rgamd
import synth
from gpu_util import kern, rel_max, rel_rms, small_obs, stage

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    obs, vis = small_obs()
    w_step = 10.0
    w_min, P = synth.kernel_range(obs.w, w_step)
    table = rgamd.make_wkernels(P, 8, 16, obs.theta, w_min, w_step)
    ref, fl, bad = orc.grid_wprojection(obs.grid_size, obs.theta, obs.u, obs.v, obs.w, vis, table,
                                        w_min, w_step)
    assert bad == 0 and fl == 8 * 256 * obs.n_vis
    return obs, vis, table, w_min, w_step, ref


def run(ctx, obs, vis, table, w_min, w_step, mode, vis_dtype=torch.complex128,
        k_dtype=torch.complex128, grid_dtype=torch.complex128, **kw):
    _, fv = stage(obs.u, obs.v, obs.w, vis, vis_dtype)
    kt, wk = kern(table, w_min, w_step, k_dtype)
    N = obs.grid_size
    g = torch.empty((N, N), dtype=grid_dtype, device="cuda")
    ctx.grid(rgamd.params(N, obs.theta, mode=mode, **kw), fv, wk, g)
    st = ctx.stats()
    torch.cuda.synchronize()
    return g.cpu().numpy().astype(np.complex128), st


def test_plain_fp64_matches_oracle(ctx, case):
    obs, vis, table, w_min, w_step, ref = case
    g, st = run(ctx, obs, vis, table, w_min, w_step, rgamd.RG_MODE_PLAIN)
    assert rel_max(g, ref) < 1e-12
    assert st.n_bad_plane == 0


@pytest.mark.parametrize("tile", [32, 48])
def test_batched_fp64_matches_oracle(ctx, case, tile):
    obs, vis, table, w_min, w_step, ref = case
    g, st = run(ctx, obs, vis, table, w_min, w_step, rgamd.RG_MODE_BATCHED, tile=tile,
                vis_per_item=64)
    assert rel_max(g, ref) < 1e-12
    assert st.n_gridded == obs.n_vis and st.n_dropped == 0 and st.n_bad_plane == 0
    assert st.n_items >= 1


@pytest.mark.parametrize("comp", [0, 1])
@pytest.mark.parametrize("grid_dtype", [torch.complex64, torch.complex128])
def test_batched_fp32_within_tolerance(ctx, case, comp, grid_dtype):
    """FP32 visibilities and taps, FP64 index math: uv grid and image within 1e-5."""
    obs, vis, table, w_min, w_step, ref = case
    g, _ = run(ctx, obs, vis, table, w_min, w_step, rgamd.RG_MODE_BATCHED,
               vis_dtype=torch.complex64, k_dtype=torch.complex64, grid_dtype=grid_dtype,
               compensated=comp)
    assert rel_max(g, ref) < 1e-5
    assert rel_max(orc.grid_to_image(g), orc.grid_to_image(ref)) < 1e-5


def test_items_split_and_order_independence(ctx, case):
    obs, vis, table, w_min, w_step, ref = case
    a, _ = run(ctx, obs, vis, table, w_min, w_step, rgamd.RG_MODE_BATCHED, vis_per_item=7)
    perm = np.random.default_rng(3).permutation(obs.n_vis)
    shuffled = synth.Observation(obs.name, obs.u[perm], obs.v[perm], obs.w[perm], obs.l, obs.m,
                                 obs.flux, obs.theta, obs.lam, obs.n_bl, obs.n_dumps, obs.freqs)
    b, _ = run(ctx, shuffled, vis[perm], table, w_min, w_step, rgamd.RG_MODE_BATCHED)
    assert rel_max(a, ref) < 1e-12 and rel_max(b, ref) < 1e-12


def test_edges_dropped_and_bad_planes(ctx):
    """Footprints hanging over the grid edge keep their in-grid taps; fully
    outside visibilities are dropped; W-planes outside the table are skipped."""
    N, theta, S, R = 64, 0.1, 8, 4
    rng = np.random.default_rng(11)
    table = rng.standard_normal((3, R, R, S, S)) + 1j * rng.standard_normal((3, R, R, S, S))
    u = np.array([-318.0, 0.0, 316.0, 400.0, 0.0, 37.2, -250.0])
    v = np.array([10.0, -318.0, 316.0, 0.0, 900.0, 12.5, 250.0])
    w = np.array([0.0, 5.0, -5.0, 0.0, 0.0, 100.0, 0.0])
    vis = rng.standard_normal(7) + 1j * rng.standard_normal(7)
    ref, _, bad = orc.grid_wprojection(N, theta, u, v, w, vis, table, -10.0, 10.0)
    assert bad == 1
    obs = synth.Observation("edge", u, v, w, None, None, None, theta, N / theta, 0, 0, None)
    for mode in (rgamd.RG_MODE_PLAIN, rgamd.RG_MODE_BATCHED):
        g, st = run(ctx, obs, vis, table, -10.0, 10.0, mode)
        assert rel_max(g, ref) < 1e-12
        assert st.n_bad_plane == 1
        if mode == rgamd.RG_MODE_BATCHED:
            assert st.n_dropped == 2 and st.n_gridded == 4


def test_empty_stream(ctx, case):
    obs, vis, table, w_min, w_step, ref = case
    e = synth.Observation("empty", obs.u[:0], obs.v[:0], obs.w[:0], None, None, None, obs.theta,
                          obs.lam, 0, 0, None)
    for mode in (rgamd.RG_MODE_PLAIN, rgamd.RG_MODE_BATCHED):
        g, _ = run(ctx, e, vis[:0], table, w_min, w_step, mode)
        assert not np.any(g)


def test_channels_match_single_calls(ctx, case):
    obs, vis, table, w_min, w_step, ref = case
    n = obs.n_vis
    off = np.array([0, n // 5, n // 5, n // 2, n], np.int64)  # includes an empty channel
    _, fv = stage(obs.u, obs.v, obs.w, vis, torch.complex64)
    kt, wk = kern(table, w_min, w_step, torch.complex64)
    N = obs.grid_size
    grids = torch.empty((4, N, N), dtype=torch.complex64, device="cuda")
    ctx.grid_channels(rgamd.params(N, obs.theta), fv, off, wk, grids)
    torch.cuda.synchronize()
    got = grids.cpu().numpy()
    for c in range(4):
        sl = slice(off[c], off[c + 1])
        r, _, _ = orc.grid_wprojection(N, obs.theta, obs.u[sl], obs.v[sl], obs.w[sl], vis[sl],
                                       table, w_min, w_step)
        if off[c + 1] == off[c]:
            assert not np.any(got[c])
        else:
            assert rel_max(got[c], r) < 1e-5


def test_c1_full_size_fp64(ctx):
    """C1 (27 stations, 256 dumps, 2048^2): plain and batched FP64 vs the oracle."""
    obs = synth.config_c1()
    vis = synth.predict_numpy(obs.u, obs.v, obs.w, obs.l, obs.m, obs.flux)
    w_step = 10.0
    w_min, P = synth.kernel_range(obs.w, w_step)
    table = rgamd.make_wkernels(P, 8, 16, obs.theta, w_min, w_step)
    ref, _, bad = orc.grid_wprojection(obs.grid_size, obs.theta, obs.u, obs.v, obs.w, vis, table,
                                       w_min, w_step, nthreads=8)
    assert bad == 0
    for mode in (rgamd.RG_MODE_PLAIN, rgamd.RG_MODE_BATCHED):
        g, st = run(ctx, obs, vis, table, w_min, w_step, mode)
        assert rel_max(g, ref) < 1e-12, mode
        assert rel_rms(g, ref) < 1e-13, mode
    g32, _ = run(ctx, obs, vis, table, w_min, w_step, rgamd.RG_MODE_BATCHED,
                 vis_dtype=torch.complex64, k_dtype=torch.complex64)
    img32, img64 = orc.grid_to_image(g32), orc.grid_to_image(ref)
    assert rel_max(img32, img64) < 1e-5 and rel_rms(img32, img64) < 1e-5


def test_hot_tile_across_replicas(ctx):
    """60 000 visibilities in one 32 x 32 tile (235 blocks of the sort, so every
    counter replica holds part of the key) plus a sparse background: batched
    FP64 equals the oracle."""
    N, theta, S, R = 256, 0.1, 8, 4
    rng = np.random.default_rng(21)
    P = 5
    table = rng.standard_normal((P, R, R, S, S)) + 1j * rng.standard_normal((P, R, R, S, S))
    n_hot, n_bg = 60000, 5000
    u = np.concatenate([rng.uniform(-100.0, 100.0, n_hot), rng.uniform(-1200.0, 1200.0, n_bg)])
    v = np.concatenate([rng.uniform(-100.0, 100.0, n_hot), rng.uniform(-1200.0, 1200.0, n_bg)])
    w = rng.uniform(-20.0, 20.0, n_hot + n_bg)
    vis = rng.standard_normal(n_hot + n_bg) + 1j * rng.standard_normal(n_hot + n_bg)
    ref, _, bad = orc.grid_wprojection(N, theta, u, v, w, vis, table, -20.0, 10.0)
    assert bad == 0
    obs = synth.Observation("hot", u, v, w, None, None, None, theta, N / theta, 0, 0, None)
    for per_item in (0, 333):
        g, st = run(ctx, obs, vis, table, -20.0, 10.0, rgamd.RG_MODE_BATCHED,
                    vis_per_item=per_item)
        assert rel_max(g, ref) < 1e-12
        assert st.n_gridded == n_hot + n_bg and st.n_dropped == 0
