"""Launch-parameter search for batched gridding calls, with an on-disk cache.

The batched paths have a few knobs that change speed but not results:
  W-projection (subgrid_size 0): tile edge (32, 48) and visibilities per work
      item (vis_per_item);
  W-towers (subgrid_size n > 0): band height (tile: 4, 8, 16 in FP32),
      vis_per_item, tasks per batch (floor_batch) and the radix of the column
      pass's FFT stages (fft_radix: 8 or 4).
tuned_params() times candidates on the caller's own data (a coordinate search:
one knob at a time, the others held at the best values so far, two sweeps) and
keeps the winner in a JSON file keyed by the call's shape, so that later calls
of the same shape -- in this process or another -- skip the search.

Key: grid size N, subgrid size n (0 for W-projection), kernel support S and
oversampling R, visibility precision (fp32 / fp64), grid type, channel count
and the device (ISA, compute units).  Lookup order: the user's cache
file ($RGAMD_TUNE_CACHE, else ~/.cache/rgamd/tune.json), then the table
shipped beside this module (tuned/mi355x.json: searches made on an MI355X for
the shapes of the benchmark workloads), then a search whose result goes to
the user's file.  Entries that no longer parse are ignored (and replaced on
the next write); writes go through a temporary file and an atomic rename,
merged with what another process may have written meanwhile.
"""
from __future__ import annotations

import json
import os
import tempfile
import time
import warnings

import rgamd

KNOBS = ("tile", "vis_per_item", "floor_batch", "fft_radix")
PACKAGED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned", "mi355x.json")


def cache_path() -> str:
    return os.environ.get("RGAMD_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "rgamd", "tune.json")


def device_name(device: int = 0) -> str:
    """ISA and compute-unit count (the marketing name that ROCm reports can be
    generic or empty, depending on the environment)."""
    import torch
    pr = torch.cuda.get_device_properties(device)
    arch = getattr(pr, "gcnArchName", "") or pr.name
    return f"{arch} {pr.multi_processor_count}CU"


def tune_key(p: rgamd.GridParams, fv: rgamd.FlatVis, wk: rgamd.WKern, grid_type: int,
             n_chan: int = 1, device: str = "") -> str:
    prec = "fp32" if fv.vis_type == rgamd.RG_C64 else "fp64"
    gt = "c64" if grid_type == rgamd.RG_C64 else "c128"
    return (f"N={p.grid_size} n={p.subgrid_size} S={wk.size} R={wk.oversampling} prec={prec} "
            f"grid={gt} chan={n_chan} mode={p.mode} comp={p.compensated} dev={device}")


def search_space(p: rgamd.GridParams, fv: rgamd.FlatVis, n_tasks: int = 0) -> dict:
    """Candidate values per knob (the first value of each is the default)."""
    fp32 = fv.vis_type == rgamd.RG_C64
    if p.subgrid_size == 0:
        return {"tile": [32, 48], "vis_per_item": [2048, 1024, 4096]}
    space = {"tile": [8, 4] + ([16] if fp32 else []),
             "vis_per_item": [2048, 1024, 4096],
             "fft_radix": [8, 4]}
    if n_tasks > 1:
        space["floor_batch"] = [0, (n_tasks + 1) // 2]
    return space


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except FileNotFoundError:
        return {}
    except (OSError, ValueError) as e:
        warnings.warn(f"ignoring unreadable tuning cache {path}: {e}")
        return {}


def _store(path: str, key: str, entry: dict) -> None:
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    data = _load(path)
    data[key] = entry
    fd, tmp = tempfile.mkstemp(prefix=".tune-", dir=d)
    try:
        with os.fdopen(fd, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise


def _with(p: rgamd.GridParams, cand: dict) -> rgamd.GridParams:
    q = rgamd.GridParams()
    for name, _ in rgamd.GridParams._fields_:
        setattr(q, name, getattr(p, name))
    for k, v in cand.items():
        setattr(q, k, int(v))
    return q


def coordinate_search(space: dict, measure, sweeps: int = 2):
    """Minimises measure(candidate dict) -> ms one knob at a time.  Returns
    (best candidate, best ms, [(candidate, ms)] of every distinct candidate
    measured)."""
    best = {k: v[0] for k, v in space.items()}
    seen = {}

    def ms_of(c):
        t = tuple(sorted(c.items()))
        if t not in seen:
            seen[t] = float(measure(dict(c)))
        return seen[t]

    best_ms = ms_of(best)
    for _ in range(sweeps):
        changed = False
        for knob, values in space.items():
            for v in values:
                c = dict(best, **{knob: v})
                ms = ms_of(c)
                if ms < best_ms:
                    best, best_ms, changed = c, ms, True
        if not changed:
            break
    table = [(dict(t), ms) for t, ms in seen.items()]
    return best, best_ms, table


def tuned_params(ctx: rgamd.Context, base: rgamd.GridParams, fv: rgamd.FlatVis,
                 wk: rgamd.WKern, grid, offsets=None, cache: str | None = None,
                 space: dict | None = None, reps: int = 3, retune: bool = False,
                 device: str | None = None, measure=None, packaged: str | None = PACKAGED):
    """GridParams for `base`'s call on this data: the cached choice for the
    call's key (user cache, then the shipped table `packaged`), else the
    result of a search (which overwrites `grid`) that is then cached.
    retune=True searches whatever the caches hold.  Returns (params, info)
    with info = {"key", "source" ("cache" | "packaged" | "measured"),
    "chosen", "ms", "table", "path"}.  `measure` replaces the device timing
    (a function of a candidate dict returning ms); `space` replaces
    search_space()."""
    path = cache or cache_path()
    n_chan = 1 if offsets is None else len(offsets) - 1
    if device is None:
        device = device_name(getattr(ctx, "device", 0))
    key = tune_key(base, fv, wk, rgamd.grid_type_of(grid), n_chan, device)
    if not retune:
        for source, f in (("cache", path), ("packaged", packaged)):
            hit = _load(f).get(key) if f else None
            if isinstance(hit, dict) and isinstance(hit.get("params"), dict):
                chosen = {k: int(v) for k, v in hit["params"].items() if k in KNOBS}
                return _with(base, chosen), {"key": key, "source": source, "chosen": chosen,
                                             "ms": hit.get("ms"),
                                             "table": hit.get("table", []), "path": f}

    def run(p):
        if offsets is None:
            ctx.grid(p, fv, wk, grid)
        else:
            ctx.grid_channels(p, fv, offsets, wk, grid)

    if measure is None:
        import torch

        def measure(cand):
            p = _with(base, cand)
            run(p)  # warm-up: plan, scratch memory
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                run(p)
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps

        if space is None:
            run(base)
            st = ctx.stats()
            space = search_space(base, fv, int(st.n_tasks))
    elif space is None:
        space = search_space(base, fv)
    best, best_ms, table = coordinate_search(space, measure)
    entry = {"params": best, "ms": best_ms,
             "table": [dict(c, ms=ms) for c, ms in table],
             "when": time.strftime("%Y-%m-%dT%H:%M:%S")}
    _store(path, key, entry)
    return _with(base, best), {"key": key, "source": "measured", "chosen": best, "ms": best_ms,
                               "table": entry["table"], "path": path}
