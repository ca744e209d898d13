"""Uniform weighting: weights of the visibilities of one coarse cell sum to one."""
import numpy as np
import pyoracle as orc


def test_uniform_weights_sum_to_one_per_cell():
    rng = np.random.default_rng(6)
    N, theta = 32, 0.1
    u = rng.uniform(-100, 100, 500)
    v = rng.uniform(-100, 100, 500)
    wv, wg = orc.weight(N, theta, u, v, np.ones(500, np.complex128))
    cells = (np.floor(theta * v + .5).astype(int) + N // 2) * N + np.floor(theta * u + .5).astype(int) + N // 2
    sums = np.zeros(N * N)
    np.add.at(sums, cells, wv.real)
    occ = np.bincount(cells, minlength=N * N)
    assert np.allclose(sums[occ > 0], 1.0)
    assert np.array_equal(wg.ravel(), occ)
