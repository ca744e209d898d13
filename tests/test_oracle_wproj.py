"""Pins of the oracle's W-projection scatter (grid.c:157-202): tap placement
from an index-encoding kernel table, the conj(kernel) convention, linearity,
the serial/OpenMP equality, and the single-visibility imaging invariant with
the generated W-kernels."""
import math

import numpy as np
import pyoracle as orc
import rgamd


def index_table(P, R, S):
    """Kernel whose taps encode their own (plane, yf, xf, y, x)."""
    p, yf, xf, y, x = np.meshgrid(*[np.arange(k) for k in (P, R, R, S, S)], indexing="ij")
    return (p * 100000 + yf * 10000 + xf * 1000 + y * 30 + x).astype(np.complex128)


def test_taps_land_where_frac_coord_says():
    N, S, R, P = 64, 8, 4, 3
    k = index_table(P, R, S)
    # theta = 1: u, v in cells.  u = 5.6 -> q = floor(5.6*4 + .5) = 22 -> flx 5, xf 2
    # v = -3.2 -> q = floor(-12.8 + .5) = -13 -> fly -4, yf 3; w = 11 -> plane 1 (w_min -10, step 20)
    g, fl, bad = orc.grid_wprojection(N, 1.0, [5.6], [-3.2], [11.0], [1.0], k, -10.0, 20.0)
    assert bad == 0 and fl == 8 * S * S
    gx, gy = 5 + N // 2 - S // 2, -4 + N // 2 - S // 2
    for y in range(S):
        for x in range(S):
            assert g[gy + y, gx + x] == 1 * 100000 + 3 * 10000 + 2 * 1000 + y * 30 + x
    assert np.count_nonzero(g) == S * S
    # everything outside the footprint is zero
    mask = np.zeros_like(g, bool)
    mask[gy:gy + S, gx:gx + S] = True
    assert not np.any(g[~mask])


def test_conjugate_kernel_convention():
    N, S, R = 32, 4, 2
    k = np.zeros((1, R, R, S, S), np.complex128)
    k[0, :, :, S // 2, S // 2] = 1j  # delta at the centre tap, value i
    vis = np.array([2.0 + 3.0j])
    g, _, _ = orc.grid_wprojection(N, 1.0, [0.0], [0.0], [0.0], vis, k, 0.0, 1.0)
    # V * conj(i) = -i V
    assert g[N // 2, N // 2] == vis[0] * -1j
    assert np.count_nonzero(g) == 1


def test_out_of_range_plane_and_edges():
    N, S, R = 32, 4, 2
    k = np.ones((2, R, R, S, S), np.complex128)
    g, fl, bad = orc.grid_wprojection(N, 1.0, [0.0, 0.0], [0.0, 0.0], [0.0, 100.0], [1, 1], k, 0.0, 1.0)
    assert bad == 1 and fl == 8 * S * S
    # a footprint hanging over the left edge keeps only its in-grid taps
    g, _, _ = orc.grid_wprojection(N, 1.0, [-15.0], [0.0], [0.0], [1.0], k, 0.0, 1.0)
    assert np.count_nonzero(g) == S * (S - 1)


def test_linearity_and_omp_equality():
    rng = np.random.default_rng(7)
    N, S, R, P = 64, 8, 4, 5
    k = rng.standard_normal((P, R, R, S, S)) + 1j * rng.standard_normal((P, R, R, S, S))
    n = 300
    u = rng.uniform(-200, 200, n)
    v = rng.uniform(-200, 200, n)
    w = rng.uniform(-40, 40, n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ga, _, _ = orc.grid_wprojection(N, 0.1, u, v, w, a, k, -40, 20)
    gb, _, _ = orc.grid_wprojection(N, 0.1, u, v, w, b, k, -40, 20)
    gab, _, _ = orc.grid_wprojection(N, 0.1, u, v, w, a + 2 * b, k, -40, 20)
    assert np.allclose(gab, ga + 2 * gb, atol=1e-12)
    go, _, _ = orc.grid_wprojection(N, 0.1, u, v, w, a, k, -40, 20, nthreads=4)
    assert np.array_equal(go, ga)


def test_single_visibility_images_to_its_fringe():
    """One visibility V = 1 gridded with the generated W-kernels images (after
    dividing by the taper A(l)A(m)) to exp(2 pi i (u l + v m + w (n - 1)))."""
    N, theta, S, R = 128, 0.1, 16, 8
    w_step = 5.0
    # on an oversample position (x = 12.375, y = -7.75 cells) and on plane 18 (w = 40)
    u, v, w = 12.375 / theta, -7.75 / theta, 40.0
    table = rgamd.make_wkernels(21, R, S, theta, -50.0, w_step)
    g, _, bad = orc.grid_wprojection(N, theta, [u], [v], [w], [1.0], table, -50.0, w_step)
    assert bad == 0
    img = orc.grid_to_image(g)
    c = np.pi * S / 4
    idx = np.arange(N) - N // 2
    lm = theta * idx / N  # l of column c, m of row r (gridder convention)
    A = np.array([rgamd.pswf(c, 2 * x / theta) for x in lm])
    L, M = np.meshgrid(lm, lm)  # L[r, c] = l_c, M[r, c] = m_r
    nn = np.sqrt(1 - L ** 2 - M ** 2)
    expect = np.exp(2j * np.pi * (u * L + v * M + w * (nn - 1)))
    inner = (np.abs(L) < 0.3 * theta) & (np.abs(M) < 0.3 * theta)
    got = img / np.outer(A, A)
    # no position or w quantisation left: only the S-tap truncation of the kernel
    err = np.abs(got - expect)[inner].max()
    assert err < 1e-3, err
    # the wrong sign of the w term is far off
    wrong = np.exp(2j * np.pi * (u * L + v * M - w * (nn - 1)))
    assert np.abs(got - wrong)[inner].max() > 0.1
