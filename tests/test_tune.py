"""Launch-parameter search and its on-disk cache (host logic; timings are
supplied by a stand-in measure function, no GPU)."""
import json
import os

import pytest
import torch

import rgamd
import rgtune


def descs(fp32=True, S=16, R=8):
    t = rgamd.RG_C64 if fp32 else rgamd.RG_C128
    fv = rgamd.FlatVis(0, 0, 0, 0, 1000, t, 0)
    wk = rgamd.WKern(0, t, 11, R, S, -5.0, 1.0)
    return fv, wk


def base(subgrid=512):
    return rgamd.params(2048, 0.1, subgrid=subgrid, margin=16 if subgrid else 0,
                        winc=10.0 if subgrid else 0.0)


def cost(c):
    """A separable cost with its minimum at tile 16, 4096 per item, radix 4,
    floor_batch 0."""
    return (10.0 + {8: 1.0, 4: 2.0, 16: 0.0}[c["tile"]] +
            {2048: 0.5, 1024: 0.7, 4096: 0.0}[c["vis_per_item"]] +
            {8: 0.3, 4: 0.0}[c["fft_radix"]] + (0.0 if c.get("floor_batch", 0) == 0 else 0.2))


def test_coordinate_search_finds_separable_minimum():
    fv, _ = descs()
    space = rgtune.search_space(base(), fv, n_tasks=4000)
    assert space["tile"] == [8, 4, 16] and space["floor_batch"] == [0, 2000]
    calls = []
    best, ms, table = rgtune.coordinate_search(space, lambda c: calls.append(c) or cost(c))
    assert best == {"tile": 16, "vis_per_item": 4096, "fft_radix": 4, "floor_batch": 0}
    assert ms == pytest.approx(10.0)
    # every distinct candidate measured once
    assert len(calls) == len({tuple(sorted(c.items())) for c in calls}) == len(table)


def test_search_spaces():
    fv64, _ = descs(fp32=False)
    assert 16 not in rgtune.search_space(base(), fv64)["tile"]  # 16-row bands are FP32 only
    wp = rgtune.search_space(base(0), fv64)
    assert set(wp) == {"tile", "vis_per_item"} and wp["tile"] == [32, 48]


def test_cache_written_then_reused(tmp_path):
    path = str(tmp_path / "sub" / "tune.json")
    fv, wk = descs()
    grid = torch.empty((2, 2), dtype=torch.complex64)
    n = []

    def measure(c):
        n.append(c)
        return cost(c)

    p, info = rgtune.tuned_params(None, base(), fv, wk, grid, cache=path, device="devA",
                                  measure=measure)
    assert info["source"] == "measured" and n
    assert (p.tile, p.vis_per_item, p.fft_radix) == (16, 4096, 4)
    assert p.grid_size == 2048 and p.subgrid_size == 512 and p.winc == 10.0
    data = json.load(open(path))
    assert info["key"] in data and data[info["key"]]["params"]["tile"] == 16
    assert "dev=devA" in info["key"] and "n=512" in info["key"] and "prec=fp32" in info["key"]

    def never(c):
        raise AssertionError("a cached key is not measured again")

    p2, info2 = rgtune.tuned_params(None, base(), fv, wk, grid, cache=path, device="devA",
                                    measure=never)
    assert info2["source"] == "cache" and info2["key"] == info["key"]
    assert (p2.tile, p2.vis_per_item, p2.fft_radix, p2.floor_batch) == (16, 4096, 4, 0)

    # another precision, device or shape is a different key; both entries kept
    k0 = len(n)
    fv64, wk64 = descs(fp32=False)
    _, info3 = rgtune.tuned_params(None, base(), fv64, wk64, grid.to(torch.complex128),
                                   cache=path, device="devA", measure=measure)
    _, info4 = rgtune.tuned_params(None, base(), fv, wk, grid, cache=path, device="devB",
                                   measure=measure)
    _, info5 = rgtune.tuned_params(None, base(256), fv, wk, grid, cache=path, device="devA",
                                   measure=measure)
    assert {info3["source"], info4["source"], info5["source"]} == {"measured"} and len(n) > k0
    assert len({info["key"], info3["key"], info4["key"], info5["key"]}) == 4
    assert len(json.load(open(path))) == 4

    # retune measures again and overwrites the entry
    k1 = len(n)
    _, info6 = rgtune.tuned_params(None, base(), fv, wk, grid, cache=path, device="devA",
                                   measure=lambda c: 1.0 if c["tile"] == 4 else cost(c),
                                   retune=True)
    assert info6["source"] == "measured" and info6["chosen"]["tile"] == 4 and len(n) == k1
    assert json.load(open(path))[info["key"]]["params"]["tile"] == 4


def test_unreadable_cache_is_replaced(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    path.write_text("{not json")
    monkeypatch.setenv("RGAMD_TUNE_CACHE", str(path))
    assert rgtune.cache_path() == str(path)
    fv, wk = descs()
    grid = torch.empty((2, 2), dtype=torch.complex64)
    with pytest.warns(UserWarning, match="unreadable tuning cache"):
        _, info = rgtune.tuned_params(None, base(), fv, wk, grid, device="d", measure=cost)
    assert info["source"] == "measured" and info["path"] == str(path)
    assert list(json.load(open(path))) == [info["key"]]
    assert not [f for f in os.listdir(tmp_path) if f.startswith(".tune-")]


def test_shipped_table_after_user_cache(tmp_path):
    fv, wk = descs()
    grid = torch.empty((2, 2), dtype=torch.complex64)
    key = rgtune.tune_key(base(), fv, wk, rgamd.RG_C64, 1, "devA")
    shipped = tmp_path / "shipped.json"
    shipped.write_text(json.dumps({key: {"params": {"tile": 4, "vis_per_item": 1024,
                                                    "floor_batch": 0, "fft_radix": 8},
                                         "ms": 5.0}}))
    user = str(tmp_path / "user.json")

    def never(c):
        raise AssertionError("no search while a table holds the key")

    p, info = rgtune.tuned_params(None, base(), fv, wk, grid, cache=user, device="devA",
                                  measure=never, packaged=str(shipped))
    assert info["source"] == "packaged" and (p.tile, p.vis_per_item) == (4, 1024)
    assert not os.path.exists(user)  # a table hit writes nothing
    # a search (retune) goes to the user's file, which then wins over the table
    _, info2 = rgtune.tuned_params(None, base(), fv, wk, grid, cache=user, device="devA",
                                   measure=cost, packaged=str(shipped), retune=True)
    assert info2["source"] == "measured" and info2["chosen"]["tile"] == 16
    p3, info3 = rgtune.tuned_params(None, base(), fv, wk, grid, cache=user, device="devA",
                                    measure=never, packaged=str(shipped))
    assert info3["source"] == "cache" and p3.tile == 16
