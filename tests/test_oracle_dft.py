"""The oracle's direct DFT: closed form for a point source and the exact
DFT <-> FFT relation of the image conventions (transpose + point flip)."""
import math

import numpy as np
import pyoracle as orc


def test_dft_point_source_closed_form():
    rng = np.random.default_rng(4)
    N, lam = 32, 320.0
    n = 200
    u = rng.uniform(-100, 100, n)
    v = rng.uniform(-100, 100, n)
    w = rng.uniform(-50, 50, n)
    # pick a pixel (col, row) and place the source there: l from the row, m from the column
    col, row = 20, 9
    l0, m0 = (row - N // 2) / lam, (col - N // 2) / lam
    n0 = math.sqrt(1 - l0 * l0 - m0 * m0)
    vis = np.exp(2j * math.pi * (u * l0 + v * m0 + (1 - n0) * w))
    out = orc.dft(np.array([col, 0]), np.array([row, 0]), N, lam, u, v, w, vis)
    assert abs(out[0] - n) < 1e-9 * n
    # a second pixel: direct evaluation of the closed-form sum
    l1, m1 = (0 - N // 2) / lam, (0 - N // 2) / lam
    n1 = math.sqrt(1 - l1 * l1 - m1 * m1)
    expect = np.sum(vis * np.exp(-2j * math.pi * (u * l1 + v * m1 + (1 - n1) * w)))
    assert abs(out[1] - expect) < 1e-9 * n


def test_dft_equals_fft_of_gridded_cells():
    """Visibilities exactly on cells with w = 0: nearest-cell gridding + imaging FFT
    equals the DFT, at gridder pixel (N - x, N - y) mod N for DFT pixel (x, y)."""
    rng = np.random.default_rng(5)
    N, theta = 16, 0.1
    lam = N / theta
    n = 40
    cu = rng.integers(-N // 2 + 1, N // 2, n)
    cv = rng.integers(-N // 2 + 1, N // 2, n)
    u, v, w = cu / theta, cv / theta, np.zeros(n)
    vis = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    g = np.zeros((N, N), np.complex128)
    np.add.at(g, (cv + N // 2, cu + N // 2), vis)
    img = orc.grid_to_image(g)
    ys, xs = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    d = orc.dft(xs.ravel(), ys.ravel(), N, lam, u, v, w, vis).reshape(N, N)
    for y in range(N):
        for x in range(N):
            assert abs(d[y, x] - img[(N - x) % N, (N - y) % N]) < 1e-10
