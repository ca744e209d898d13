"""Pins of the oracle's coordinate and index rules (frac_coord, W-plane,
tower and floor indices) by known answers and closed forms."""
import math

import numpy as np
import pyoracle as orc


def test_frac_coord_known_answers():
    # N=64, S=8, R=8, theta=1: x = u cells.  Worked by hand from grid.c:49-56.
    assert orc.frac_coord(64, 8, 8, 1.0, 3.3, 0.0) == (3 + 32 - 4, 0 + 32 - 4, 2, 0)
    # exact half-oversample tie rounds up: 3 + 2.5/8
    assert orc.frac_coord(64, 8, 8, 1.0, 3.3125, 0.0)[2] == 3
    # 2.9375 = 3 - 0.5/8 belongs to cell 3, oversample 0
    gx, _, xf, _ = orc.frac_coord(64, 8, 8, 1.0, 2.9375, 0.0)
    assert (gx, xf) == (3 + 28, 0)
    # negative coordinate
    gx, _, xf, _ = orc.frac_coord(64, 8, 8, 1.0, -0.3, 0.0)
    assert (gx, xf) == (-1 + 28, 6)
    # offsets and theta scaling: x = theta (u - d_u) = 0.5 * (10 - 4) = 3
    gx, gy, xf, yf = orc.frac_coord(32, 4, 4, 0.5, 10.0, -2.0, 4.0, 2.0)
    assert (gx, gy, xf, yf) == (3 + 16 - 2, -2 + 16 - 2, 0, 0)


def test_frac_coord_is_nearest_oversample_position():
    """Closed form: with q = floor(x R + 1/2), the top-left is floor(q / R) and the
    oversample index q mod R (checked on dyadic x, where x R is exact)."""
    rng = np.random.default_rng(0)
    N, S = 128, 16
    for R in (1, 2, 4, 8, 16):
        for k in rng.integers(-4096, 4096, size=400):
            x = float(k) / 64.0  # dyadic, exact
            q = math.floor(x * R + 0.5)
            gx, _, xf, _ = orc.frac_coord(N, S, R, 1.0, x, 0.0)
            assert gx - N // 2 + S // 2 == q // R, (x, R)
            assert xf == q % R, (x, R)


def test_packed_offsets():
    go, so = orc.frac_coord_packed(64, 8, 8, 1.0, 3.3, -1.3)
    gx, gy, xf, yf = orc.frac_coord(64, 8, 8, 1.0, 3.3, -1.3)
    assert go == gy * 64 + gx and so == 64 * (yf * 8 + xf)


def test_w_plane_nearest():
    # plane i sits at w_min + i w_step; the index is the nearest plane, ties upward
    assert orc.w_plane(0.0, -50.0, 10.0) == 5
    assert orc.w_plane(4.999, -50.0, 10.0) == 5
    assert orc.w_plane(5.0, -50.0, 10.0) == 6
    assert orc.w_plane(-55.0, -50.0, 10.0) == 0
    assert orc.w_plane(-55.1, -50.0, 10.0) == -1


def test_tower_and_floor_index():
    # N=2048, chunk 496: cell floor(u*theta+.5) + 1024, divided by 496, truncated
    assert orc.tower_index(0.0, 0.1, 2048, 496) == 1024 // 496
    assert orc.tower_index(-10240.0, 0.1, 2048, 496) == 0
    assert orc.tower_index(10235.0, 0.1, 2048, 496) == (1024 + 1024) // 496  # 4
    # floor wp: (wp - 1/2) winc <= w < (wp + 1/2) winc
    assert orc.floor_index(4.999, 10.0) == 0
    assert orc.floor_index(5.0, 10.0) == 1
    assert orc.floor_index(-5.0, 10.0) == 0
    assert orc.floor_index(-5.001, 10.0) == -1


def test_oversample_index_stays_in_table_at_rounding_ties():
    """With a non-power-of-two oversampling the FP64 formula can land on xf = R:
    R = 3, x = 0.8333333333333333 gives floor(x + 1/6) = 0 and
    floor(3 x + .5) = 3.  The index is clamped into [0, R)."""
    R = 3
    x = 0.8333333333333333
    assert math.floor(x + 0.5 / R) == 0 and math.floor((x - 0) * R + 0.5) == R
    gx, _, xf, _ = orc.frac_coord(64, 6, R, 1.0, x, 0.0)
    assert xf == R - 1 and gx == 0 + 32 - 3
