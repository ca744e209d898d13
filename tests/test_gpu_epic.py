"""EPIC E-field gridding on the GPU (src/epic/epic_grid.cu) against the
oracle: the reference's workload shape (1000 timestamps x 256 antennas,
64^2 grids, 2x2 top hat), a complex illumination with antennas on the grid
edges, and a grid too large for LDS (global-atomic path)."""
import numpy as np
import pytest
import torch

import pyoracle as orc
import rgamd
from gpu_util import dev

pytestmark = pytest.mark.gpu


def run(ctx, f, x, y, il, g):
    nb = f.shape[0]
    grids = torch.full((nb, g, g), 3.0, dtype=torch.complex64, device="cuda")
    ctx.epic_grid(dev(f.astype(np.complex64)), dev(x.astype(np.int32)),
                  dev(y.astype(np.int32)), dev(il.astype(np.complex64)), grids)
    torch.cuda.synchronize()
    return grids.cpu().numpy()


def fields(rng, nb, na):
    return (rng.normal(0, 5, (nb, na)) + 1j * rng.normal(0, 5, (nb, na))).astype(np.complex64)


def test_reference_workload_matches_oracle(ctx):
    g, S, nb, na = 64, 2, 1000, 256
    rng = np.random.default_rng(1)
    f = fields(rng, nb, na)
    x = rng.integers(S, g - S + 1, nb * na)  # loc_distribution(ILLUM_X, GRID_SIZE - ILLUM_X)
    y = rng.integers(S, g - S + 1, nb * na)
    il = np.ones((S, S), np.complex64)
    got = run(ctx, f, x, y, il, g)
    ref = orc.epic_grid(f.astype(np.complex128), x, y, il, g)
    assert np.abs(got - ref).max() < 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("g,S", [(40, 3), (128, 4)])
def test_complex_illumination_and_edges_match_oracle(ctx, g, S):
    rng = np.random.default_rng(g)
    nb, na = 7, 300
    f = fields(rng, nb, na)
    x = rng.integers(-S, g + 1, nb * na)  # some footprints hang over every edge
    y = rng.integers(-S, g + 1, nb * na)
    il = (rng.standard_normal((S, S)) + 1j * rng.standard_normal((S, S))).astype(np.complex64)
    got = run(ctx, f, x, y, il, g)
    ref = orc.epic_grid(f.astype(np.complex128), x, y, il.astype(np.complex128), g)
    assert np.abs(got - ref).max() < 1e-5 * np.abs(ref).max()


def test_invalid_arguments_are_rejected(ctx):
    f = torch.zeros((2, 3), dtype=torch.complex64, device="cuda")
    xy = torch.zeros(6, dtype=torch.int32, device="cuda")
    il = torch.zeros((0, 0), dtype=torch.complex64, device="cuda")
    grids = torch.zeros((2, 8, 8), dtype=torch.complex64, device="cuda")
    with pytest.raises(rgamd.RGError):
        ctx.epic_grid(f, xy, xy, il, grids)  # support 0
