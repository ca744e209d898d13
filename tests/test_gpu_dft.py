"""Direct DFT on the GPU against the oracle; prediction against numpy."""
import math

import numpy as np
import pytest
import torch

import pyoracle as orc
import rgamd
import synth
from gpu_util import dev, rel_max, small_obs, stage

pytestmark = pytest.mark.gpu


def test_dft_fp64_and_mixed(ctx):
    obs, vis = small_obs(n_st=10, n_dumps=20)
    N, lam = obs.grid_size, obs.lam
    rng = np.random.default_rng(4)
    px = rng.integers(0, N, 500).astype(np.int32)
    py = rng.integers(0, N, 500).astype(np.int32)
    ref = orc.dft(px, py, N, lam, obs.u, obs.v, obs.w, vis)
    _, fv = stage(obs.u, obs.v, obs.w, vis)
    dpx, dpy = dev(px), dev(py)
    out = torch.empty(500, dtype=torch.complex128, device="cuda")
    ctx.dft(fv, dpx, dpy, N, lam, out, rgamd.RG_DFT_FP64)
    torch.cuda.synchronize()
    assert rel_max(out.cpu().numpy(), ref) < 1e-11
    ctx.dft(fv, dpx, dpy, N, lam, out, rgamd.RG_DFT_MIXED)
    torch.cuda.synchronize()
    assert rel_max(out.cpu().numpy(), ref) < 1e-5


def test_dft_image_layout_and_contract(ctx):
    obs, vis = small_obs(n_st=6, n_dumps=8, lam=320.0)
    N, lam = obs.grid_size, obs.lam
    _, fv = stage(obs.u, obs.v, obs.w, vis)
    img = torch.empty(N * N, dtype=torch.complex128, device="cuda")
    ctx.dft_image(fv, N, lam, img, N * N // 64, 64)
    torch.cuda.synchronize()
    ys, xs = np.divmod(np.arange(N * N), N)
    ref = orc.dft(xs, ys, N, lam, obs.u, obs.v, obs.w, vis)
    assert rel_max(img.cpu().numpy(), ref) < 1e-11
    with pytest.raises(rgamd.RGError) as e:
        ctx.dft_image(fv, N, lam, img, N * N // 64 - 1, 64)
    assert e.value.code == rgamd.RG_ERR_PARAM


def test_predict_matches_numpy():
    obs = synth.config_small(n_st=8, n_dumps=10, n_src=7)
    ref = synth.predict_numpy(obs.u, obs.v, obs.w, obs.l, obs.m, obs.flux)
    for dt, tol in ((torch.complex128, 1e-12), (torch.complex64, 1e-6)):
        out = torch.empty(obs.n_vis, dtype=dt, device="cuda")
        rgamd.predict_points(dev(obs.u), dev(obs.v), dev(obs.w), dev(obs.l), dev(obs.m),
                             dev(obs.flux), out)
        torch.cuda.synchronize()
        assert rel_max(out.cpu().numpy(), ref) < tol


def test_dft_split_over_visibility_ranges(ctx):
    """Few pixels and many visibilities: the stream is split over several
    ranges whose partial sums are added in order; same result as the oracle,
    and identical on a second run."""
    rng = np.random.default_rng(8)
    n = 300_000
    u, v = rng.uniform(-5000, 5000, n), rng.uniform(-5000, 5000, n)
    w = rng.uniform(-300, 300, n)
    vis = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    N, lam = 2048, 20480.0
    px = rng.integers(0, N, 40).astype(np.int32)
    py = rng.integers(0, N, 40).astype(np.int32)
    ref = orc.dft(px, py, N, lam, u, v, w, vis)
    for vdt, mode, tol in ((torch.complex128, rgamd.RG_DFT_FP64, 1e-11),
                           (torch.complex64, rgamd.RG_DFT_MIXED, 2e-5)):
        _, fv = stage(u, v, w, vis, vdt)
        out = torch.empty(40, dtype=torch.complex128, device="cuda")
        ctx.dft(fv, dev(px), dev(py), N, lam, out, mode)
        torch.cuda.synchronize()
        a = out.cpu().numpy()
        ctx.dft(fv, dev(px), dev(py), N, lam, out, mode)
        torch.cuda.synchronize()
        assert np.array_equal(a, out.cpu().numpy())
        assert rel_max(a, ref) < tol, mode
