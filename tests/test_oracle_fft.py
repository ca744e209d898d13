"""The oracle's FFT, shift and hermitian completion against numpy / identities."""
import numpy as np
import pyoracle as orc


def test_fft_matches_numpy():
    rng = np.random.default_rng(1)
    for n in (2, 4, 8, 16, 64, 128):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        inv = orc.fft2d(a, +1)
        fwd = orc.fft2d(a, -1)
        assert np.allclose(inv, np.fft.ifft2(a) * n * n, rtol=0, atol=1e-12 * n * n)
        assert np.allclose(fwd, np.fft.fft2(a), rtol=0, atol=1e-12 * n * n)
        x = a[0]
        assert np.allclose(orc.fft1d(x, -1), np.fft.fft(x), atol=1e-12 * n)


def test_fft_shift_is_numpy_fftshift_and_involution():
    rng = np.random.default_rng(2)
    for n in (2, 4, 8, 10, 32):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        s = orc.fft_shift(a)
        assert np.array_equal(s, np.fft.fftshift(a))
        assert np.array_equal(orc.fft_shift(s), a)


def test_make_hermitian_symmetry_and_real_image():
    rng = np.random.default_rng(3)
    n = 16
    g = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    h = orc.make_hermitian(g)
    # for even n the completion pairs (r, c) with (n - r, n - c) over rows/cols >= 1
    for r in range(1, n):
        for c in range(1, n):
            assert abs(h[r, c] - np.conj(h[n - r, n - c])) < 1e-12
    # grid.c:639-662 leaves row 0 and column 0 untouched except the pairs above
    assert np.array_equal(h[0], g[0])
    # the imaging FFT of the completed inner grid is real
    inner = np.zeros_like(h)
    inner[1:, 1:] = h[1:, 1:]
    img = orc.grid_to_image(inner)
    assert np.abs(img.imag).max() < 1e-10 * np.abs(img.real).max()
