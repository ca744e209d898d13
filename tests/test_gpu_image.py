"""Image conventions of the CPU reference pipeline on the GPU (uniform
weighting, hermitian completion, grid correction) against the oracle, and the
whole weighted / hermitian / forward-FFT image against the oracle pipeline."""
import numpy as np
import pytest
import torch

import pyoracle as orc
import rgamd
import synth
from gpu_util import dev, kern, rel_max, small_obs, stage

pytestmark = pytest.mark.gpu


def weight_inputs(seed, N, theta, n):
    """Visibilities with exact cell-boundary ties (theta u + 1/2 integral),
    a long run of visibilities in one cell and a scatter over the grid."""
    rng = np.random.default_rng(seed)
    half = N // 2
    u = rng.uniform(-(half - 1) / theta, (half - 1) / theta, n)
    v = rng.uniform(-(half - 1) / theta, (half - 1) / theta, n)
    k = n // 8
    u[:k] = (rng.integers(-half + 1, half - 1, k) - 0.5) / theta  # ties
    v[k:3 * k] = 3.2 / theta                                     # one busy row
    u[k:3 * k] = -7.1 / theta                                    # ... and cell
    vis = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return u, v, vis


@pytest.mark.parametrize("N", [64, 65])
def test_weight_fp64_matches_oracle_bit_for_bit(ctx, N):
    theta = 0.1
    u, v, vis = weight_inputs(5, N, theta, 20000)
    ref_vis, ref_cnt = orc.weight(N, theta, u, v, vis)
    du, dv, dvis = dev(u), dev(v), dev(vis)
    cnt = torch.empty(N * N, dtype=torch.int32, device="cuda")
    ctx.weight(du, dv, dvis, N, theta, counts=cnt)
    torch.cuda.synchronize()
    assert np.array_equal(cnt.cpu().numpy().view(np.uint32), ref_cnt.ravel())
    assert np.array_equal(dvis.cpu().numpy(), ref_vis)


def test_weight_fp32_rounds_the_fp64_quotient(ctx):
    N, theta = 64, 0.1
    u, v, vis = weight_inputs(7, N, theta, 5000)
    vis32 = vis.astype(np.complex64)
    ref_vis, _ = orc.weight(N, theta, u, v, vis32.astype(np.complex128))
    dvis = dev(vis32)
    ctx.weight(dev(u), dev(v), dvis, N, theta)
    torch.cuda.synchronize()
    assert np.array_equal(dvis.cpu().numpy(), ref_vis.astype(np.complex64))


def test_weight_leaves_visibilities_outside_the_grid(ctx):
    """The reference indexes out of bounds for cells outside the grid; here
    such visibilities are neither counted nor changed."""
    N, theta = 32, 0.1
    u, v, vis = weight_inputs(8, N, theta, 3000)
    out = np.zeros(len(u), bool)
    out[::7] = True
    u[out] = np.where(np.arange(out.sum()) % 2 == 0, 1e6, -(N // 2 + 0.5) / theta - 1e-9)
    u = np.where(out & (np.arange(len(u)) % 3 == 0), np.nan, u)
    ref_vis, ref_cnt = orc.weight(N, theta, u[~out], v[~out], vis[~out])
    du, dv, dvis = dev(u), dev(v), dev(vis)
    cnt = torch.empty(N * N, dtype=torch.int32, device="cuda")
    ctx.weight(du, dv, dvis, N, theta, counts=cnt)
    torch.cuda.synchronize()
    got = dvis.cpu().numpy()
    assert np.array_equal(got[out], vis[out])
    assert np.array_equal(got[~out], ref_vis)
    assert np.array_equal(cnt.cpu().numpy().view(np.uint32), ref_cnt.ravel())


def test_weight_of_no_visibilities_zeroes_the_counts(ctx):
    cnt = torch.full((16 * 16,), 5, dtype=torch.int32, device="cuda")
    e = torch.empty(0, dtype=torch.float64, device="cuda")
    ctx.weight(e, e, torch.empty(0, dtype=torch.complex128, device="cuda"), 16, 0.1, counts=cnt)
    torch.cuda.synchronize()
    assert int(cnt.abs().sum()) == 0


@pytest.mark.parametrize("n", [1, 2, 33, 64, 1024])
def test_make_hermitian_matches_oracle_bit_for_bit(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    t = dev(a)
    rgamd.make_hermitian(t)
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy(), orc.make_hermitian(a))
    a32 = a.astype(np.complex64)
    t32 = dev(a32)
    rgamd.make_hermitian(t32)
    torch.cuda.synchronize()
    ref32 = orc.make_hermitian(a32.astype(np.complex128)).astype(np.complex64)
    assert np.array_equal(t32.cpu().numpy(), ref32)


def test_gcf_correct_divides_by_the_separable_taper():
    n = 48
    rng = np.random.default_rng(3)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    taper = rgamd.pswf_eval(np.pi * 16 / 4, 2.0 * (np.arange(n) - n // 2) / n)
    taper = np.where(np.abs(taper) < 1e-3, 1.0, taper)
    t = dev(a)
    rgamd.gcf_correct(t, dev(taper))
    torch.cuda.synchronize()
    tt = taper[:, None] * taper[None, :]
    assert np.array_equal(t.cpu().numpy(), a.real / tt + 1j * (a.imag / tt))


@pytest.mark.parametrize("mode", [rgamd.RG_MODE_PLAIN, rgamd.RG_MODE_BATCHED])
def test_cpu_convention_image_matches_oracle_pipeline(ctx, mode):
    """weight -> W-projection -> make_hermitian -> shift -> forward FFT -> shift
    (src/reference/main.c:151-281) against the same steps in the oracle."""
    obs, vis = small_obs(seed=4)
    N, theta = obs.grid_size, obs.theta
    w_min, P = synth.kernel_range(obs.w, 10.0)
    table = rgamd.make_wkernels(P, 8, 16, theta, w_min, 10.0)
    wvis, _ = orc.weight(N, theta, obs.u, obs.v, vis)
    ref, _, bad = orc.grid_wprojection(N, theta, obs.u, obs.v, obs.w, wvis, table, w_min, 10.0)
    assert bad == 0
    ref = orc.fft_shift(orc.fft2d(orc.fft_shift(orc.make_hermitian(ref)), -1))

    arrs, fv = stage(obs.u, obs.v, obs.w, vis)
    _, wk = kern(table, w_min, 10.0)
    ctx.weight(arrs[0], arrs[1], arrs[3], N, theta)
    g = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    ctx.grid(rgamd.params(N, theta, mode=mode), fv, wk, g)
    rgamd.make_hermitian(g)
    rgamd.fft_shift(g)
    ctx.fft2d(g, -1)
    rgamd.fft_shift(g)
    torch.cuda.synchronize()
    got = g.cpu().numpy()
    assert rel_max(got, ref) < 1e-12
    # hermitian completion makes the image real
    assert np.abs(got.imag).max() < 1e-9 * np.abs(got.real).max()
