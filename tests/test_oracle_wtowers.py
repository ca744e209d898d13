"""Pins of the oracle's W-towers (grid.c:295-555): reduction to W-projection
when every w is zero, the flop accounting of grid.c:519-530, cipow against
the closed form, tower coverage, and the single-visibility fringe through
the w-screens."""
import math

import numpy as np
import pyoracle as orc
import rgamd


def _obs(rng, n, uv=250.0, wmax=0.0):
    u = rng.uniform(-uv, uv, n)
    v = rng.uniform(-uv, uv, n)
    w = rng.uniform(-wmax, wmax, n) if wmax else np.zeros(n)
    vis = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return u, v, w, vis


def test_w_zero_reduces_to_wprojection():
    rng = np.random.default_rng(8)
    N, theta, S, R = 64, 0.1, 8, 4
    k = rng.standard_normal((1, R, R, S, S)) + 1j * rng.standard_normal((1, R, R, S, S))
    u, v, w, vis = _obs(rng, 400)
    gp, flp, _ = orc.grid_wprojection(N, theta, u, v, w, vis, k, 0.0, 1.0)
    # margin >= S + 2 keeps every footprint inside its subgrid
    gt, flt, bad = orc.grid_wtowers(N, theta, u, v, w, vis, k, 0.0, 1.0, 32, 12, 10.0)
    assert bad == 0 and flt == flp == 8 * S * S * len(u)
    assert np.abs(gt - gp).max() < 1e-12 * np.abs(gp).max()


def test_margin_equal_to_support_drops_the_outer_row():
    """With margin == S a visibility just inside a chunk's lower edge has its
    footprint start one cell before the subgrid; those taps are dropped (the GPU
    reference's bounds check).  Chunk 2 of N=64, n=32, M=8 starts at cell 16: a
    visibility at y = 15.82 cells (tower cy = 2) loses footprint row 0."""
    N, theta, S, R = 64, 0.1, 8, 4
    k = np.ones((1, R, R, S, S), np.complex128)
    u, v = 1.41 / theta, 15.8169 / theta
    assert orc.tower_index(v, theta, N, 24) == 2
    gp, _, _ = orc.grid_wprojection(N, theta, [u], [v], [0.0], [1.0], k, 0.0, 1.0)
    gt, _, _ = orc.grid_wtowers(N, theta, [u], [v], [0.0], [1.0], k, 0.0, 1.0, 32, 8, 10.0)
    gy = orc.frac_coord(N, S, R, theta, u, v)[1]
    d = gp - gt
    assert np.allclose(d[gy], gp[gy]) and np.abs(np.delete(d, gy, axis=0)).max() < 1e-12


def test_every_visibility_gridded_once_and_parallel_equal():
    rng = np.random.default_rng(9)
    N, theta, S, R = 64, 0.1, 8, 4
    k = rng.standard_normal((11, R, R, S, S)) + 1j * rng.standard_normal((11, R, R, S, S))
    u, v, w, vis = _obs(rng, 500, wmax=60.0)
    g1, fl, bad = orc.grid_wtowers(N, theta, u, v, w, vis, k, -5.0, 1.0, 32, 8, 10.0)
    assert bad == 0 and fl == 8 * S * S * len(u)
    g4, fl4, _ = orc.grid_wtowers(N, theta, u, v, w, vis, k, -5.0, 1.0, 32, 8, 10.0, nthreads=4)
    assert fl4 == fl and np.array_equal(g1, g4)


def test_cipow_closed_form():
    n, theta, winc = 16, 0.2, 7.0
    F = orc.wtransfer(n, theta, winc)
    idx = np.fft.fftshift(np.arange(n) - n // 2)  # shifted layout: index 0 is l = 0
    L, M = np.meshgrid(theta * idx / n, theta * idx / n)
    base = 1 - np.sqrt(1 - L ** 2 - M ** 2)
    assert np.allclose(F, np.exp(2j * np.pi * winc * base), atol=1e-14)
    for k in (-5, -1, 0, 1, 2, 3, 17, 100):
        for y, x in ((0, 0), (3, 5), (9, 14)):
            got = orc.cipow(complex(F[y, x]), k)
            expect = np.exp(2j * np.pi * k * winc * base[y, x])
            assert abs(got - expect) < 1e-12


def test_tower_windows_cover_the_grid():
    N, n, M = 2048, 512, 16
    chunk = n - M
    cc = N // chunk + 1
    cover = np.zeros(N, int)
    for cx in range(cc):
        x0 = chunk * cx - N // 2 - M // 2
        lo, hi = max(x0, -N // 2), min(x0 + n, N // 2)
        cover[lo + N // 2:hi + N // 2] += 1
    assert cover.min() >= 1 and cover.max() == 2


def test_single_visibility_fringe_through_screens():
    N, theta, S, R, n, M, winc = 128, 0.1, 16, 8, 64, 16, 10.0
    w_step = 1.0
    table = rgamd.make_wkernels(11, R, S, theta, -5.0, w_step)
    # on an oversample position and exactly on floor 3 (w = 30)
    u, v, w = 2.375 / theta, -3.5 / theta, 30.0
    g, _, bad = orc.grid_wtowers(N, theta, [u], [v], [w], [1.0], table, -5.0, w_step, n, M, winc)
    assert bad == 0
    img = orc.grid_to_image(g)
    c = np.pi * S / 4
    lm = theta * (np.arange(N) - N // 2) / N
    A = np.array([rgamd.pswf(c, 2 * x / theta) for x in lm])
    L, M2 = np.meshgrid(lm, lm)
    nn = np.sqrt(1 - L ** 2 - M2 ** 2)
    expect = np.exp(2j * np.pi * (u * L + v * M2 + w * (nn - 1)))
    inner = (np.abs(L) < 0.3 * theta) & (np.abs(M2) < 0.3 * theta)
    got = img / np.outer(A, A)
    assert np.abs(got - expect)[inner].max() < 1e-3
    wrong = np.exp(2j * np.pi * (u * L + v * M2 - w * (nn - 1)))
    assert np.abs(got - wrong)[inner].max() > 1e-2


def test_fringes_on_several_floors_of_one_tower():
    """Three visibilities of one tower on floors -2, 3 and 7 (gaps of 5 and 4
    floors, off-floor residual w): the image is the closed-form sum of their
    fringes.  This pins the inter-floor step subimg = F^(wp - last_wp) subimg +
    IFFT(subgrid) (grid.c:473-480) and the w = 0 transfer F^-last_wp
    (grid.c:489-496): a power taken from the wrong floor puts a phase error of
    2 pi k winc (1 - n) on the earlier floors' fringes."""
    N, theta, S, R, n, M, winc = 128, 0.1, 16, 8, 64, 16, 10.0
    table = rgamd.make_wkernels(11, R, S, theta, -5.0, 1.0)
    u = np.array([2.375, 1.125, 3.0]) / theta
    v = np.array([-3.5, -2.25, -4.125]) / theta
    w = np.array([-21.0, 31.0, 68.5])  # floors -2, 3, 7 (residuals -1, 1, -1.5)
    vis = np.array([1.0, 0.7 - 0.4j, -0.5 + 0.9j])
    chunk = n - M
    assert len({(orc.tower_index(a, theta, N, chunk), orc.tower_index(b, theta, N, chunk))
                for a, b in zip(u, v)}) == 1
    assert [orc.floor_index(x, winc) for x in w] == [-2, 3, 7]
    g, _, bad = orc.grid_wtowers(N, theta, u, v, w, vis, table, -5.0, 1.0, n, M, winc)
    assert bad == 0
    img = orc.grid_to_image(g)
    c = np.pi * S / 4
    lm = theta * (np.arange(N) - N // 2) / N
    A = np.array([rgamd.pswf(c, 2 * x / theta) for x in lm])
    L, M2 = np.meshgrid(lm, lm)
    nn = np.sqrt(1 - L ** 2 - M2 ** 2)
    inner = (np.abs(L) < 0.3 * theta) & (np.abs(M2) < 0.3 * theta)
    got = img / np.outer(A, A)

    def fringes(wv):
        return sum(a * np.exp(2j * np.pi * (x * L + y * M2 + z * (nn - 1)))
                   for a, x, y, z in zip(vis, u, v, wv))

    # kernel interpolation error ~1e-3 per unit fringe (R = 8, 16 taps)
    assert np.abs(got - fringes(w))[inner].max() < 5e-3
    # each visibility's floor term moved by one floor is caught (>= 0.047)
    for k in range(3):
        ws = w.copy()
        ws[k] += winc
        assert np.abs(got - fringes(ws))[inner].max() > 2e-2
