"""Batched FFT, shift and imaging on the GPU against numpy and the oracle."""
import numpy as np
import pytest
import torch

import pyoracle as orc
import rgamd
from gpu_util import dev, rel_max

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
def test_fft2d_fp64_and_fp32(ctx, n):
    rng = np.random.default_rng(n)
    batch = 2 if n <= 512 else 1
    a = rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))
    for sign, ref in ((+1, np.fft.ifft2(a) * n * n), (-1, np.fft.fft2(a))):
        t = dev(a)
        ctx.fft2d(t, sign)
        torch.cuda.synchronize()
        assert rel_max(t.cpu().numpy(), ref) < 1e-13 * np.log2(n) * 4
        t32 = dev(a).to(torch.complex64)
        ctx.fft2d(t32, sign)
        torch.cuda.synchronize()
        assert rel_max(t32.cpu().numpy(), ref) < 2e-6 * np.log2(n)


def test_fft_shift_and_grid_to_image(ctx):
    rng = np.random.default_rng(5)
    n = 256
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    t = dev(a)
    rgamd.fft_shift(t)
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy(), np.fft.fftshift(a))
    t = dev(a)
    ctx.grid_to_image(t)
    torch.cuda.synchronize()
    assert rel_max(t.cpu().numpy(), orc.grid_to_image(a)) < 1e-12


def test_bad_sizes_rejected(ctx):
    t = torch.zeros((12, 12), dtype=torch.complex128, device="cuda")
    with pytest.raises(rgamd.RGError) as e:
        ctx.fft2d(t, 1)
    assert e.value.code == rgamd.RG_ERR_PARAM
