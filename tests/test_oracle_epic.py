"""EPIC E-field gridding in the oracle (src/epic/epic_grid.cu): the footprint
of a single antenna, the total-sum identity and linearity."""
import numpy as np

import pyoracle as orc


def test_single_antenna_places_the_conjugate_illumination():
    g, S = 16, 3
    rng = np.random.default_rng(1)
    il = rng.standard_normal((S, S)) + 1j * rng.standard_normal((S, S))
    f = np.array([[2.0 - 1.0j]])
    out = orc.epic_grid(f, [5], [9], il, g)
    ref = np.zeros((g, g), complex)
    ref[9:9 + S, 5:5 + S] = np.conj(il) * f[0, 0]  # rows y, columns x
    assert np.allclose(out[0], ref, rtol=0, atol=1e-15)


def test_total_sum_and_edge_drops():
    g, S, nb, na = 32, 2, 5, 40
    rng = np.random.default_rng(2)
    f = rng.standard_normal((nb, na)) + 1j * rng.standard_normal((nb, na))
    x = rng.integers(0, g - S + 1, nb * na)
    y = rng.integers(0, g - S + 1, nb * na)
    il = np.ones((S, S))
    out = orc.epic_grid(f, x, y, il, g)
    assert np.allclose(out.sum(axis=(1, 2)), f.sum(axis=1) * S * S)
    # an antenna at the last column keeps only its first tap column
    out = orc.epic_grid(np.ones((1, 1)), [g - 1], [0], il, g)
    assert out[0].sum() == S and out[0][:S, g - 1].sum() == S


def test_linear_in_the_fields():
    g, S = 24, 2
    rng = np.random.default_rng(3)
    x, y = rng.integers(0, g - 1, 30), rng.integers(0, g - 1, 30)
    il = rng.standard_normal((S, S)) + 1j * rng.standard_normal((S, S))
    f1 = rng.standard_normal((1, 30)) + 1j * rng.standard_normal((1, 30))
    f2 = rng.standard_normal((1, 30)) + 1j * rng.standard_normal((1, 30))
    a = orc.epic_grid(f1, x, y, il, g) + 2.5 * orc.epic_grid(f2, x, y, il, g)
    assert np.allclose(orc.epic_grid(f1 + 2.5 * f2, x, y, il, g), a, atol=1e-13)
