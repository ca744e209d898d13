"""The launch-parameter search on the device: W-towers and W-projection calls
tuned on their own data, the choice cached on disk and reused, and the tuned
call still equal to the oracle (src/reference/grid.c:295-555 restated)."""
import json

import pytest
import torch

import pyoracle as orc
import rgamd
import rgtune
from gpu_util import kern, rel_max, rel_rms, small_obs, stage

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    obs, vis = small_obs(n_st=14, n_dumps=32, arm_m=700.0, lam=2560.0)  # N = 256
    table = rgamd.make_wkernels(11, 8, 16, obs.theta, -5.0, 1.0)
    ref, _, bad = orc.grid_wtowers(obs.grid_size, obs.theta, obs.u, obs.v, obs.w, vis, table,
                                   -5.0, 1.0, 64, 16, 10.0, nthreads=8)
    assert bad == 0
    return obs, vis, table, ref


def test_wtowers_tuned_cached_and_correct(ctx, case, tmp_path):
    obs, vis, table, ref = case
    path = str(tmp_path / "tune.json")
    _, fv = stage(obs.u, obs.v, obs.w, vis, torch.complex64)
    _, wk = kern(table, -5.0, 1.0, torch.complex64)
    N = obs.grid_size
    g = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    base = rgamd.params(N, obs.theta, subgrid=64, margin=16, winc=10.0)
    p, info = rgtune.tuned_params(ctx, base, fv, wk, g, cache=path, reps=1, packaged=None)
    assert info["source"] == "measured"
    tried = {(e["tile"], e["fft_radix"]) for e in info["table"]}
    assert {(8, 8), (4, 8), (16, 8)} <= tried and any(r == 4 for _, r in tried)
    assert json.load(open(path))[info["key"]]["params"] == info["chosen"]
    p2, info2 = rgtune.tuned_params(ctx, base, fv, wk, g, cache=path, packaged=None)
    assert info2["source"] == "cache" and info2["chosen"] == info["chosen"]
    ctx.grid(p2, fv, wk, g)
    torch.cuda.synchronize()
    img, img_ref = orc.grid_to_image(g.cpu().numpy()), orc.grid_to_image(ref)
    assert rel_max(img, img_ref) < 1e-5 and rel_rms(img, img_ref) < 1e-5


def test_wproj_tuned_and_cached(ctx, case, tmp_path):
    obs, vis, table, _ = case
    path = str(tmp_path / "tune.json")
    _, fv = stage(obs.u, obs.v, obs.w, vis, torch.complex64)
    _, wk = kern(table, -5.0, 1.0, torch.complex64)
    N = obs.grid_size
    g = torch.empty((N, N), dtype=torch.complex64, device="cuda")
    base = rgamd.params(N, obs.theta)
    p, info = rgtune.tuned_params(ctx, base, fv, wk, g, cache=path, reps=1, packaged=None)
    assert info["source"] == "measured" and "n=0" in info["key"] and "grid=c64" in info["key"]
    assert p.tile in (32, 48) and p.vis_per_item in (1024, 2048, 4096)
    _, info2 = rgtune.tuned_params(ctx, base, fv, wk, g, cache=path, packaged=None)
    assert info2["source"] == "cache"
