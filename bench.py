#!/usr/bin/env python3
"""Gridding benchmark on one MI355X: Mvis/s on a 2048^2 grid, FP32 batched vs
FP64 plain, gated at 1e-5 relative image error.

Workload: 512 stations (radius ~ Exp(3 km)), 64 dumps over 4 h at dec -30 deg,
150 MHz, baselines cut at 10 km (first-sample length), 50 point sources,
theta 0.1, lambda 20480 (2048^2 grid), W-kernels support 16 x oversample 8.
W-towers: subgrid 512, margin 16, winc 10.  Inputs are resident in HBM before
the timed region; a step is one full gridding pass (sort, scatter, subgrid
FFTs and w-screens, fold) over all visibilities.  The final N^2 image FFT is
not part of a step.

Prints ONE JSON line.  value = FP32 batched W-towers Mvis/s (the workload's
own gridder), compared with the FP64 plain W-towers gridder; the W-projection
numbers of the same stream are reported alongside under results.
"""
import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "radio-gridders_amd"))

import numpy as np  # noqa: E402

METRIC = "Mvis/s gridded (2048² grid): FP32 batched vs FP64 plain at rel err ≤1e-5"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (published)
PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")


def profiled_traffic(name, launch_params):
    """HBM bytes per launch of a kernel, from the FETCH_SIZE / WRITE_SIZE
    summary under profiles/ (counters scaled by a 1 GiB copy), or None when
    the summary is missing or was taken with other launch parameters."""
    path = os.path.join(PROFILES, name)
    traffic, prof_params = None, None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_dispatch:"):
                    traffic = float(line.split(":", 1)[1])
                elif line.startswith("launch params:"):
                    prof_params = json.loads(line.split(":", 1)[1])
    except (OSError, ValueError):
        return None
    return traffic if prof_params == launch_params else None


def launch_params(p):
    """The launch parameters a batched call runs with (defaults filled in)."""
    towers = p.subgrid_size > 0
    fp32_band = p.tile if p.tile in (4, 16) else 8
    return {"tile": fp32_band if towers else (p.tile or 32),
            "vis_per_item": p.vis_per_item or 2048, "floor_batch": int(p.floor_batch),
            "fft_radix": (p.fft_radix or 8) if towers else 0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--stations", type=int, default=512)
    ap.add_argument("--dumps", type=int, default=64)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true", help="skip the FP64 plain runs")
    ap.add_argument("--only", choices=["all", "wproj", "wtowers", "channels", "wstack", "epic",
                                       "c1"],
                    default="all")
    ap.add_argument("--skip-epic", action="store_true", help="skip the EPIC gridding workload")
    ap.add_argument("--skip-c1", action="store_true",
                    help="skip the 27-station W-projection precision-form workload")
    ap.add_argument("--skip-wstack", action="store_true",
                    help="skip the W-stacking predict workload")
    ap.add_argument("--skip-channels", action="store_true",
                    help="skip the 256-channel batched W-projection workload")
    ap.add_argument("--c5", action="store_true",
                    help="also run the stress workload (~260 M visibilities, 4096^2, W-towers "
                         "subgrid 256); adds about a minute")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--calibrate", action="store_true", help="run a 1 GiB fill and a 1 GiB sum for counter scaling")
    ap.add_argument("--dft-pixels", type=int, default=65536,
                    help="seeded pixels of the FP64 DFT oracle comparison (0 = skip)")
    ap.add_argument("--tower-band", type=int, default=0, help="W-towers band height (4, 8 or 16)")
    ap.add_argument("--floor-batch", type=int, default=0, help="W-towers tasks per batch")
    ap.add_argument("--vis-per-item", type=int, default=0,
                    help="visibilities per work item (0 = default)")
    ap.add_argument("--fft-radix", type=int, default=0,
                    help="W-towers column-pass FFT stage radix (8 or 4, 0 = 8)")
    ap.add_argument("--no-tune", action="store_true",
                    help="FP32 batched calls with the default (or given) launch parameters "
                         "instead of the tuned ones")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the graph-replay runs (headline from the eager calls)")
    ap.add_argument("--retune", action="store_true",
                    help="search the launch parameters even when a cache holds them")
    ap.add_argument("--tune-cache", default=None,
                    help="tuning cache file (default $RGAMD_TUNE_CACHE or "
                         "~/.cache/rgamd/tune.json; radio-gridders_amd/tuned/mi355x.json is "
                         "read after it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the gridding path is "
              "single-GPU in this build", file=sys.stderr)
        sys.exit(2)

    import torch

    import rgamd
    import rgtune
    import synth

    t0 = time.time()
    obs = synth.config_c2(n_st=args.stations, n_dumps=args.dumps)
    N, theta = obs.grid_size, obs.theta
    log(f"[bench] {obs.n_vis} visibilities, {obs.n_bl} baselines, N={N} ({time.time()-t0:.1f}s)")
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    u, v, w = dev(obs.u), dev(obs.v), dev(obs.w)
    vis128 = torch.empty(obs.n_vis, dtype=torch.complex128, device="cuda")
    rgamd.predict_points(u, v, w, dev(obs.l), dev(obs.m), dev(obs.flux), vis128)
    vis64 = vis128.to(torch.complex64)
    fv128 = rgamd.flat_vis(u, v, w, vis128)
    fv64 = rgamd.flat_vis(u, v, w, vis64)

    # W-projection table over the data's w range (step 10), W-towers table over
    # the residual range [-winc/2, winc/2] (step 1)
    S, R = 16, 8
    wp_step = 10.0
    wp_min, wp_planes = synth.kernel_range(obs.w, wp_step)
    t1 = time.time()
    tab_p = dev(rgamd.make_wkernels(wp_planes, R, S, theta, wp_min, wp_step))
    tab_t = dev(rgamd.make_wkernels(11, R, S, theta, -5.0, 1.0))
    log(f"[bench] kernels: {wp_planes} W-projection planes, 11 W-towers planes "
        f"({time.time()-t1:.1f}s)")
    k128_p, k64_p = rgamd.wkern_desc(tab_p, wp_min, wp_step), None
    tab_p32 = tab_p.to(torch.complex64)
    k64_p = rgamd.wkern_desc(tab_p32, wp_min, wp_step)
    tab_t32 = tab_t.to(torch.complex64)
    k128_t, k64_t = rgamd.wkern_desc(tab_t, -5.0, 1.0), rgamd.wkern_desc(tab_t32, -5.0, 1.0)
    SUB, MARGIN, WINC = 512, 16, 10.0

    ctx = rgamd.Context(0)
    stream = torch.cuda.current_stream()
    if args.calibrate:
        # dispatches of known traffic for scaling the FETCH_SIZE / WRITE_SIZE
        # counters of the gridding kernels: a fill that writes 1 GiB and a
        # sum that reads it
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        total = float(src.sum())
        assert total == float(1 << 28)
        del src

    def timed(label, p, fv, wk, grid, steps, warmup, graph=False):
        """ms per step of `steps` calls (kernel timers off), then the per-kernel
        breakdown from a second pass with the timers on.  graph=True: the call
        is captured once in a HIP graph (after an eager call of its shape) and
        the graph is replayed `steps` times instead (no host work per step)."""
        ctx.set_timing(False)
        for _ in range(max(warmup, 1 if graph else 0)):
            ctx.grid(p, fv, wk, grid)
        torch.cuda.synchronize()
        run = lambda: ctx.grid(p, fv, wk, grid)  # noqa: E731
        g = None
        if graph:
            g = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                with torch.cuda.graph(g, stream=side):
                    ctx.grid(p, fv, wk, grid, stream=side.cuda_stream)
            torch.cuda.current_stream().wait_stream(side)
            g.replay()
            torch.cuda.synchronize()
            run = g.replay
        t_start = time.perf_counter()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(steps):
            run()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_start
        ms = ev0.elapsed_time(ev1) / steps
        st = ctx.stats()  # raises when a replay outgrew the capacities it was captured with
        timers = {}
        if g is not None:
            del g
            ctx.thaw()
        else:
            ctx.set_timing(True)
            for name in rgamd.TIMERS:
                ctx.timer(name)  # reset
            for _ in range(steps):
                ctx.grid(p, fv, wk, grid)
            torch.cuda.synchronize()
            for name in rgamd.TIMERS:
                tm, cnt = ctx.timer(name)
                if cnt:
                    timers[name] = {"ms_per_step": tm / steps, "launches_per_step": cnt / steps}
            ctx.set_timing(False)
        log(f"[bench] {label}{' (graph replay)' if graph else ''}: {ms:.3f} ms/step "
            f"({obs.n_vis / ms / 1e3:.1f} Mvis/s), wall {wall / steps * 1e3:.1f} ms/step, "
            f"timers {timers}")
        return {"ms_per_step": ms, "mvis_s": obs.n_vis / ms / 1e3, "timers": timers,
                "gridded": st.n_gridded, "dropped": st.n_dropped, "bad_plane": st.n_bad_plane,
                "items": st.n_items, "tasks": st.n_tasks, "towers": st.n_towers,
                "records": st.n_records, "bands": st.n_bands, "chunks": st.n_chunks,
                "_stats": st}

    def image(grid):
        img = grid.to(torch.complex128).clone()
        ctx.grid_to_image(img)
        return img

    def errs(a, b):
        d = (a - b).abs()
        return {"max_rel": float(d.max() / b.abs().max()),
                "rms_rel": float(torch.sqrt((d ** 2).mean()) / torch.sqrt((b.abs() ** 2).mean()))}

    def tune(label, r, base, fv, wk, grid):
        """Launch parameters of the FP32 batched call from the tuning cache, or
        searched on this data (outside the timed regions) and cached."""
        if args.no_tune:
            return base
        t0 = time.time()
        p, info = rgtune.tuned_params(ctx, base, fv, wk, grid, cache=args.tune_cache,
                                      retune=args.retune, reps=5)
        r["tuning"] = {"source": info["source"], "key": info["key"], "chosen": info["chosen"],
                       "ms": info["ms"], "candidates": info["table"]}
        log(f"[bench] {label} FP32 launch parameters ({info['source']}, "
            f"{time.time() - t0:.1f}s): {info['chosen']}")
        return p

    res = {}
    dft_images = {}
    K, W = args.steps, args.warmup
    if args.only in ("all", "wproj"):
        g64 = torch.empty((N, N), dtype=torch.complex128, device="cuda")
        g32 = torch.empty((N, N), dtype=torch.complex64, device="cuda")
        r = {}
        if not args.skip_plain:
            r["fp64_plain"] = timed("W-proj FP64 plain", rgamd.params(N, theta, rgamd.RG_MODE_PLAIN),
                                    fv128, k128_p, g64, K, W)
        else:
            ctx.grid(rgamd.params(N, theta, rgamd.RG_MODE_BATCHED), fv128, k128_p, g64)
        r["fp64_batched"] = timed("W-proj FP64 batched", rgamd.params(N, theta), fv128, k128_p,
                                  g64 if args.skip_plain else torch.empty_like(g64), K, W)
        p32 = tune("W-proj", r, rgamd.params(N, theta, vis_per_item=args.vis_per_item), fv64,
                   k64_p, g32)
        r["launch_params"] = launch_params(p32)
        r["fp32_batched"] = timed("W-proj FP32 batched", p32, fv64, k64_p, g32, K, W)
        if not args.no_graph:
            r["fp32_batched_graph"] = timed("W-proj FP32 batched", p32, fv64, k64_p, g32, K, W,
                                            graph=True)
        if not args.skip_plain:
            ctx.grid(rgamd.params(N, theta, rgamd.RG_MODE_PLAIN), fv128, k128_p, g64)
        i32, i64 = image(g32), image(g64)
        r["error_fp32_batched_vs_fp64"] = errs(i32, i64)
        dft_images["wproj_fp64_plain" if not args.skip_plain else "wproj_fp64_batched"] = i64
        dft_images["wproj_fp32_batched"] = i32
        torch.cuda.synchronize()
        res["wproj"] = r
        del g32
    if args.only in ("all", "wtowers"):
        g64 = torch.empty((N, N), dtype=torch.complex128, device="cuda")
        g32 = torch.empty((N, N), dtype=torch.complex128, device="cuda")
        r = {}
        pp = rgamd.params(N, theta, rgamd.RG_MODE_PLAIN, subgrid=SUB, margin=MARGIN, winc=WINC)
        pb = rgamd.params(N, theta, rgamd.RG_MODE_BATCHED, subgrid=SUB, margin=MARGIN, winc=WINC,
                          tile=args.tower_band, floor_batch=args.floor_batch,
                          vis_per_item=args.vis_per_item, fft_radix=args.fft_radix)
        gb = torch.empty_like(g64)
        r["fp64_batched"] = timed("W-towers FP64 batched", pb, fv128, k128_t, gb, K, W)
        if not args.skip_plain:
            r["fp64_plain"] = timed("W-towers FP64 plain", pp, fv128, k128_t, g64, 1, 0)
        else:
            g64.copy_(gb)
        p32 = tune("W-towers", r, pb, fv64, k64_t, g32)
        r["launch_params"] = launch_params(p32)
        r["fp32_batched"] = timed("W-towers FP32 batched", p32, fv64, k64_t, g32, K, W)
        if not args.no_graph:
            r["fp32_batched_graph"] = timed("W-towers FP32 batched", p32, fv64, k64_t, g32, K, W,
                                            graph=True)
        # the precision form with two-sum compensated register sums in the band lanes
        pc = rgamd.GridParams.from_buffer_copy(p32)
        pc.compensated = 1
        gc = torch.empty_like(g64)
        r["fp32_batched_compensated"] = timed("W-towers FP32 batched compensated", pc, fv64,
                                              k64_t, gc, K, W)
        i32, i64, ib, ic = image(g32), image(g64), image(gb), image(gc)
        del gc
        r["error_fp32_batched_vs_fp64"] = errs(i32, i64)
        r["error_fp32_compensated_vs_fp64"] = errs(ic, i64)
        r["error_fp64_batched_vs_fp64"] = errs(ib, i64)
        r["speedup_fp32_vs_fp64_batched"] = (r["fp64_batched"]["ms_per_step"] /
                                            r["fp32_batched"]["ms_per_step"])
        dft_images["wtowers_fp64" + ("_plain" if not args.skip_plain else "_batched")] = i64
        dft_images["wtowers_fp32_batched"] = i32
        dft_images["wtowers_fp32_batched_compensated"] = ic
        res["wtowers"] = r
        del gb, ib

    if args.only in ("all", "c1") and not args.skip_c1:
        try:
            res["c1_c3"] = bench_c1(ctx, K, W, args.skip_cpu, args.dft_pixels)
        except Exception as e:  # reported, never hides the headline
            log(f"[bench] 27-station workload failed: {e!r}")
            res["c1_c3"] = {"error": repr(e)}

    if args.only in ("all", "channels") and not args.skip_channels:
        try:
            res["channels"] = bench_channels(ctx, K, W, args.tune_cache, args.retune)
        except Exception as e:  # reported, never hides the headline
            log(f"[bench] channels workload failed: {e!r}")
            res["channels"] = {"error": repr(e)}

    if args.only in ("all", "wstack") and not args.skip_wstack:
        try:
            res["wstack"] = bench_wstack(ctx, K, W, args.skip_cpu, args.cpu_seconds)
        except Exception as e:  # reported, never hides the headline
            log(f"[bench] W-stack predict failed: {e!r}")
            res["wstack"] = {"error": repr(e)}

    if args.only in ("all", "epic") and not args.skip_epic:
        try:
            res["epic"] = bench_epic(ctx, K, W, args.skip_cpu)
        except Exception as e:  # reported, never hides the headline
            log(f"[bench] EPIC gridding failed: {e!r}")
            res["epic"] = {"error": repr(e)}

    if args.c5:
        try:
            res["stress"] = bench_stress(ctx, K, W, tune_cache=args.tune_cache,
                                         retune=args.retune, no_tune=args.no_tune)
        except Exception as e:  # reported, never hides the headline
            log(f"[bench] stress workload failed: {e!r}")
            res["stress"] = {"error": repr(e)}

    dft = None
    if args.dft_pixels > 0 and dft_images:
        dft = dft_errors(ctx, fv128, N, obs.lam, theta, S, dft_images, args.dft_pixels)

    # headline: FP32 batched W-towers (W-projection when only that ran)
    if "wtowers" not in res and "wproj" not in res:
        print(json.dumps({"metric": METRIC, "value": None, "results": res}), flush=True)
        return
    head_kind = "wtowers" if "wtowers" in res else "wproj"
    # the step replayed from a captured graph (the batched path's serving mode:
    # no host work per step); kernel times come from the eager run's timers
    eager = res[head_kind]["fp32_batched"]
    head = dict(res[head_kind].get("fp32_batched_graph", eager))
    head["timers"] = eager["timers"]
    plain = res[head_kind].get("fp64_plain")
    nv = obs.n_vis
    # the counter summaries were taken on the default 512-station, 64-dump
    # workload (and are used only with the launch parameters they name)
    default_workload = (args.stations, args.dumps) == (512, 64)
    lp = res[head_kind]["launch_params"]
    H = lp["tile"] if lp["tile"] in (4, 16) else 8
    # FP32 tables: W-projection as given; W-towers band pass: the padded copy
    # (H zero rows after every S x S sub-kernel and before the first)
    tab_bytes = (wp_planes * 64 * S * S * 8 if head_kind == "wproj"
                 else (H + 11 * 64 * (S + H)) * S * 8)
    roof, job = roofline(head_kind, head, head["_stats"], nv, S, N, SUB, H,
                         lp if default_workload else None, tab_bytes)

    cpu = None
    if not args.skip_cpu:
        if head_kind == "wtowers":
            cpu, cpu_grid = cpu_baseline_wtowers(obs, vis128, tab_t, SUB, MARGIN, WINC,
                                                 head["towers"])
            # the CPU restatement's image against the GPU FP64 image of the same inputs
            gi = torch.from_numpy(cpu_grid).cuda()
            del cpu_grid
            ctx.grid_to_image(gi)
            ref = dft_images.get("wtowers_fp64_plain", dft_images.get("wtowers_fp64_batched"))
            cpu["error_gpu_fp64_vs_cpu"] = errs(ref, gi)
            cpu["error_gpu_fp32_batched_vs_cpu"] = errs(dft_images["wtowers_fp32_batched"], gi)
            del gi
            log(f"[bench] CPU vs GPU images: {cpu['error_gpu_fp64_vs_cpu']}")
            cpu["wprojection"] = cpu_baseline(obs, vis128, tab_p, wp_min, wp_step,
                                              args.cpu_seconds)
        else:
            cpu = cpu_baseline(obs, vis128, tab_p, wp_min, wp_step, args.cpu_seconds)
    for kind in res.values():
        if isinstance(kind, dict):
            for v in kind.values():
                if isinstance(v, dict):
                    v.pop("_stats", None)

    out = {
        "metric": METRIC,
        "value": head["mvis_s"],
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (FP64 index math, FP64 fold)",
        "data": "synthetic (seeded layout, Earth-rotation tracks, 50 point sources)",
        "config": {"workload": ("512 stations Exp(3 km), 64 dumps over 4 h, dec -30, 150 MHz, "
                                f"baselines < 10 km: {nv} visibilities gridded on 2048^2 "
                                "(theta 0.1, lambda 20480), W-kernels S16 x R8 "
                                + ("W-projection (w-step 10)" if head_kind == "wproj"
                                   else "W-towers subgrid 512, margin 16, winc 10")),
                   "n_vis": nv, "grid": N, "path": head_kind, "launch_params": lp,
                   "mode": ("one batched call per step, replayed from a captured HIP graph"
                            if "fp32_batched_graph" in res[head_kind] else
                            "one batched call per step")},
        "eager_mvis_s": eager["mvis_s"],
        "speedup_vs_fp64_plain": (head["mvis_s"] / plain["mvis_s"]) if plain else None,
        "error_vs_fp64": res[head_kind]["error_fp32_batched_vs_fp64"],
        "gate": ("max and RMS relative image error of FP32 batched against FP64 plain on "
                 "identical inputs and kernels <= 1e-5 (the metric's gate); the error against "
                 "the FP64 direct DFT is reported, not gated: it is set by the W-kernel's "
                 "oversampling and support (about 7e-4 in the inner field for every path and "
                 "precision alike), so no gridder with these kernels reaches 1e-5 against it"),
        "gate_met": (res[head_kind]["error_fp32_batched_vs_fp64"]["max_rel"] <= 1e-5 and
                     res[head_kind]["error_fp32_batched_vs_fp64"]["rms_rel"] <= 1e-5),
        "error_vs_dft": dft,
        "roofline": roof,
        "job_roofline": job,
        "cpu_baseline": cpu,
        "results": res,
    }
    print(json.dumps(out), flush=True)


def bench_c1(ctx, steps, warmup, skip_cpu, npix):
    """27 stations on a random Y (10 km arms), 256 dumps over 4 h at dec -30
    deg, 150 MHz, baselines < 10 km (about 90 000 visibilities), 10 sources,
    2048^2 grid (theta 0.1, lambda 20480), W-projection with S16 x R8 kernels
    at w-step 10.  Plain FP64 W-projection against the precision form (FP32
    visibilities and tables, FP64 uvw-to-cell and plane indexing, FP32 tile
    sums, plain and two-sum compensated), each image against the FP64 plain
    image, the FP64 direct DFT on `npix` seeded pixels, and the full 2048^2
    image of the oracle's FP64 CPU W-projection (src/reference/grid.c:187-202
    restated), whose serial run is also timed."""
    import torch

    import rgamd
    import synth
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc  # the checker: reference image and CPU timing only

    obs = synth.config_c1()
    N, theta, nv = obs.grid_size, obs.theta, obs.n_vis
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    u, v, w = dev(obs.u), dev(obs.v), dev(obs.w)
    vis128 = torch.empty(nv, dtype=torch.complex128, device="cuda")
    rgamd.predict_points(u, v, w, dev(obs.l), dev(obs.m), dev(obs.flux), vis128)
    fv128 = rgamd.flat_vis(u, v, w, vis128)
    vis64 = vis128.to(torch.complex64)
    fv64 = rgamd.flat_vis(u, v, w, vis64)
    S, R, w_step = 16, 8, 10.0
    w_min, P = synth.kernel_range(obs.w, w_step)
    tab_h = rgamd.make_wkernels(P, R, S, theta, w_min, w_step)
    tab = dev(tab_h)
    tab32 = tab.to(torch.complex64)
    k128, k64 = rgamd.wkern_desc(tab, w_min, w_step), rgamd.wkern_desc(tab32, w_min, w_step)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def run(p, fv, wk, dt, reps):
        g = torch.empty((N, N), dtype=dt, device="cuda")
        for _ in range(warmup):
            ctx.grid(p, fv, wk, g)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            ctx.grid(p, fv, wk, g)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        img = g.to(torch.complex128)
        ctx.grid_to_image(img)
        return {"ms_per_step": ms, "mvis_s": nv / ms / 1e3}, img

    out = {"n_vis": nv, "grid": N, "w_planes": P}
    images = {}
    out["fp64_plain"], images["fp64_plain"] = run(rgamd.params(N, theta, rgamd.RG_MODE_PLAIN),
                                                  fv128, k128, torch.complex128, steps)
    out["fp32_batched"], images["fp32_batched"] = run(rgamd.params(N, theta), fv64, k64,
                                                      torch.complex64, steps)
    out["fp32_batched_compensated"], images["fp32_batched_compensated"] = run(
        rgamd.params(N, theta, compensated=1), fv64, k64, torch.complex128, steps)

    def errs(a, b):
        d = (a - b).abs()
        return {"max_rel": float(d.max() / b.abs().max()),
                "rms_rel": float(torch.sqrt((d ** 2).mean()) / torch.sqrt((b.abs() ** 2).mean()))}

    # the full image of the oracle's FP64 CPU W-projection (serial, as grid.c)
    t = time.perf_counter()
    gref, _, bad = orc.grid_wprojection(N, theta, obs.u, obs.v, obs.w, vis128.cpu().numpy(),
                                        tab_h, w_min, w_step, fast=True)
    cpu_s = time.perf_counter() - t
    iref = torch.from_numpy(gref).cuda()
    ctx.grid_to_image(iref)
    for name, img in images.items():
        out[name]["error_vs_fp64_plain"] = errs(img, images["fp64_plain"])
        out[name]["error_vs_cpu_image"] = errs(img, iref)
    if npix > 0:
        out["error_vs_dft"] = dft_errors(ctx, fv128, N, obs.lam, theta, S, images, npix)
    if not skip_cpu:
        _, host = host_threads()
        out["cpu_baseline"] = {"value": nv / cpu_s / 1e6, "unit": "Mvis/s", "cores": 1,
                               "kind": "port", "host": host,
                               "sample": f"all {nv} visibilities, serial FP64 W-projection "
                                         "(grid.c:187-202 has no OpenMP), -O3 -ffast-math"}
    log(f"[bench] C1: plain FP64 {out['fp64_plain']['ms_per_step']:.3f} ms, FP32 batched "
        f"{out['fp32_batched']['ms_per_step']:.3f} ms; FP32 vs CPU image "
        f"{out['fp32_batched']['error_vs_cpu_image']}")
    del tab, tab32, iref
    return out


def bench_wstack(ctx, steps, warmup, skip_cpu, cpu_seconds):
    """W-stacking predict at the reference test's workload (test/test_wstack.cpp:
    theta 0.1, lambda 20480, test-card sky, 40 M random visibilities with u, v
    in [-theta lambda/2, theta lambda/2) and w in [-dw, dw), the Sze-Tan
    kernels of test/kernels/sze): stack build (8 sources: their rows
    transformed, columns summed directly into the 4096^2 planes) and
    degridding (visibility sort included) timed apart; degridding Mvis/s
    against the oracle's OpenMP degridding of the same stack on a sample."""
    import torch

    import rgamd
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc  # the checker: accuracy sample and CPU baseline only

    sze = os.path.join(REPO, "tests", "golden", "sze")
    ku = rgamd.load_sep_kern(os.path.join(sze, "sepkern_uv_transpose.hdf5"), True)
    kw = rgamd.load_sep_kern(os.path.join(sze, "sepkern_w_transpose.hdf5"), True)
    klm = rgamd.load_sep_kern(os.path.join(sze, "sepkern_lm.hdf5"), False)
    kn = rgamd.load_sep_kern(os.path.join(sze, "sepkern_n.hdf5"), False)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dk = {n: dev(k.data) for n, k in dict(uv=ku, w=kw, lm=klm, n=kn).items()}
    desc = {n: rgamd.sep_kern_desc(dk[n], k.size, k.oversampling)
            for n, k in dict(uv=ku, w=kw, lm=klm, n=kn).items()}
    theta, lam = 0.1, 20480.0
    p = rgamd.wstack_params(theta, lam, ku.du, kw.dw, klm.x0, ku.size, kw.size)
    card = np.array([0.95, 0.95, -0.95, -0.95, 0.95, -0.95, -0.95, 0.95, 0.0, 0.5, 0.0, -0.5, 0.5,
                     0.0, -0.5, 0.0]) * (theta / 2)
    l, m = card[0::2].copy(), card[1::2].copy()
    dl, dm = dev(l), dev(m)
    nv = 40_000_000
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    u = (torch.rand(nv, dtype=torch.float64, device="cuda", generator=gen) - 0.5) * (theta * lam)
    v = (torch.rand(nv, dtype=torch.float64, device="cuda", generator=gen) - 0.5) * (theta * lam)
    w = (torch.rand(nv, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1) * kw.dw
    out = torch.empty(nv, dtype=torch.complex128, device="cuda")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    og, wp, skipped = ctx.wstack_build(p, dl, dm, None, desc["lm"], desc["n"])
    for _ in range(warmup):
        ctx.wstack_build(p, dl, dm, None, desc["lm"], desc["n"])
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        ctx.wstack_build(p, dl, dm, None, desc["lm"], desc["n"])
    e1.record()
    torch.cuda.synchronize()
    build_ms = e0.elapsed_time(e1) / steps
    for _ in range(warmup):
        ctx.wstack_degrid(p, u, v, w, desc["uv"], desc["w"], out)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        ctx.wstack_degrid(p, u, v, w, desc["uv"], desc["w"], out)
    e1.record()
    torch.cuda.synchronize()
    degrid_ms = e0.elapsed_time(e1) / steps
    st = ctx.stats()
    # accuracy against the direct sum over the quantised sky on a sample
    sel = torch.randperm(nv, device="cuda", generator=gen)[:20000]
    us, vs, ws = u[sel].cpu().numpy(), v[sel].cpu().numpy(), w[sel].cpu().numpy()
    ref = orc.predict_quantized(l, m, None, theta, lam, us, vs, ws)
    err = np.abs(out[sel].cpu().numpy() - ref)
    taps = kw.size * ku.size * ku.size
    # compulsory bytes of the call (sort + degridding) per visibility: uvw in,
    # the 16-byte record written and read at both sort levels and read by the
    # degridding, the predicted visibility out; the stack regions each work
    # item stages are not charged.  The per-tap count (every tap at 16 B, as
    # the visibility's 4136 algorithmic bytes) is kept as `effective`: the
    # taps come from LDS, so it is no floor on HBM traffic.
    comp_vis = 24 + 4 * 16 + 16
    per_vis = taps * 16 + 24 + 16
    ach = nv * comp_vis / (degrid_ms * 1e-3) / 1e9
    ach_eff = nv * per_vis / (degrid_ms * 1e-3) / 1e9
    r = {"n_vis": nv, "planes": wp, "plane": og, "sources": len(l), "skipped_sources": skipped,
         "not_predicted": int(st.n_dropped), "build_ms": build_ms, "degrid_ms": degrid_ms,
         "degrid_mvis_s": nv / (degrid_ms * 1e-3) / 1e6,
         "total_mvis_s": nv / ((build_ms + degrid_ms) * 1e-3) / 1e6,
         "error_vs_quantised_dft": {"max_abs": float(err.max()), "mean_abs": float(err.mean()),
                                    "per_source_mean": float(err.mean() / len(l))},
         "roofline": {"kernel": "k_ws_degrid_tile with the visibility sort (one call)",
                      "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": ach / HBM_PEAK_GBS, "bytes_per_unit": {"visibility": comp_vis},
                      "traffic": None,
                      "effective": {"GBps": ach_eff, "bytes_per_unit": {"visibility": per_vis},
                                    "note": "every tap charged at 16 B; taps are read from "
                                            "LDS"}}}
    log(f"[bench] W-stack predict: {og}^2 x {wp} planes, build {build_ms:.2f} ms, degrid "
        f"{degrid_ms:.2f} ms ({r['degrid_mvis_s']:.0f} Mvis/s), mean error per source "
        f"{r['error_vs_quantised_dft']['per_source_mean']:.2e}")
    if not skip_cpu:
        # the oracle's OpenMP degridding of the same stack on a sample
        stack = ctx.wstack_stack(og, wp)
        threads, host = host_threads()
        ns = 200_000
        su, sv, sw = u[:ns].cpu().numpy(), v[:ns].cpu().numpy(), w[:ns].cpu().numpy()
        t = time.perf_counter()
        orc.wstack_degrid(stack, int(np.floor(theta * lam)), su, sv, sw, ku.du, kw.dw, ku.data,
                          ku.size, ku.oversampling, kw.data, kw.size, kw.oversampling,
                          nthreads=threads, fast=True)
        dt = time.perf_counter() - t
        reps = max(1, int(cpu_seconds / max(dt, 1e-3)))
        reps = min(reps, 40)
        t = time.perf_counter()
        for _ in range(reps):
            orc.wstack_degrid(stack, int(np.floor(theta * lam)), su, sv, sw, ku.du, kw.dw, ku.data,
                              ku.size, ku.oversampling, kw.data, kw.size, kw.oversampling,
                              nthreads=threads, fast=True)
        dt = (time.perf_counter() - t) / reps
        r["cpu_baseline"] = {"value": ns / dt / 1e6, "unit": "Mvis/s", "cores": threads,
                             "kind": "port", "host": host,
                             "sample": f"{ns} of the visibilities, degridding only, "
                                       f"{reps} repetitions"}
        log(f"[bench] CPU W-stack degrid: {r['cpu_baseline']['value']:.1f} Mvis/s on {threads} "
            f"threads")
        del stack
    del u, v, w, out
    torch.cuda.empty_cache()
    return r


def bench_epic(ctx, steps, warmup, skip_cpu):
    """EPIC E-field gridding at the reference's workload (src/epic/epic_grid.cu:
    1000 timestamps x 256 antennas, 64^2 grids, 2x2 top-hat illumination,
    positions in [2, 62]); one launch grids every timestamp."""
    import torch

    import rgamd
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc  # the checker: accuracy and CPU baseline only

    g, S, nb, na = 64, 2, 1000, 256
    rng = np.random.default_rng(9)
    f = (rng.normal(0, 5, (nb, na)) + 1j * rng.normal(0, 5, (nb, na))).astype(np.complex64)
    x = rng.integers(S, g - S + 1, nb * na).astype(np.int32)
    y = rng.integers(S, g - S + 1, nb * na).astype(np.int32)
    il = np.ones((S, S), np.complex64)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    df, dx, dy, dil = dev(f), dev(x), dev(y), dev(il)
    grids = torch.empty((nb, g, g), dtype=torch.complex64, device="cuda")
    for _ in range(max(warmup, 1)):
        ctx.epic_grid(df, dx, dy, dil, grids)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(steps, 20)
    e0.record()
    for _ in range(reps):
        ctx.epic_grid(df, dx, dy, dil, grids)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    t = time.perf_counter()
    ref = orc.epic_grid(f.astype(np.complex128), x, y, il.astype(np.complex128), g)
    cpu_s = time.perf_counter() - t
    got = grids.cpu().numpy()
    per_ts = g * g * 8 + na * (8 + 4 + 4)
    ach = nb * per_ts / (ms * 1e-3) / 1e9
    r = {"timestamps": nb, "antennas": na, "grid": g, "support": S, "ms": ms,
         "timestamps_per_s": nb / (ms * 1e-3), "samples_per_s": nb * na / (ms * 1e-3),
         "max_rel_err_vs_oracle": float(np.abs(got - ref).max() / np.abs(ref).max()),
         "roofline": {"kernel": "k_epic_grid", "bound": "hbm", "achieved": ach,
                      "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                      "bytes_per_unit": {"timestamp": per_ts}}}
    if not skip_cpu:
        r["cpu_baseline"] = {"value": nb / cpu_s, "unit": "timestamps/s", "cores": 1,
                             "kind": "port", "sample": "the whole workload, FP64 naive scatter"}
    log(f"[bench] EPIC: {nb} timestamps x {na} antennas on {g}^2: {ms * 1e3:.1f} us "
        f"({r['timestamps_per_s'] / 1e6:.2f} M timestamps/s, {ach:.0f} GB/s)")
    return r


def bench_channels(ctx, steps, warmup, tune_cache=None, retune=False):
    """256 channels (100-200 MHz) with lognormal counts (median 40 000, sigma 1)
    from 256 stations, theta 0.05 (1024^2 FP32 grid per channel): one batched
    FP32 call over all channels (eager and replayed from a captured graph)
    against 256 plain FP64 per-channel calls; launch parameters from the
    tuner."""
    import torch

    import rgamd
    import rgtune
    import synth

    t0 = time.time()
    obs = synth.config_c4()
    off = np.ascontiguousarray(obs.chan_offsets, dtype=np.int64)
    nc, N, nv = len(off) - 1, obs.grid_size, obs.n_vis
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    u, v, w = dev(obs.u), dev(obs.v), dev(obs.w)
    vis128 = torch.empty(nv, dtype=torch.complex128, device="cuda")
    rgamd.predict_points(u, v, w, dev(obs.l), dev(obs.m), dev(obs.flux), vis128)
    vis64 = vis128.to(torch.complex64)
    fv128, fv64 = rgamd.flat_vis(u, v, w, vis128), rgamd.flat_vis(u, v, w, vis64)
    w_step = 10.0
    w_min, P = synth.kernel_range(obs.w, w_step)
    tab = dev(rgamd.make_wkernels(P, 8, 16, obs.theta, w_min, w_step))
    tab32 = tab.to(torch.complex64)
    k128, k64 = rgamd.wkern_desc(tab, w_min, w_step), rgamd.wkern_desc(tab32, w_min, w_step)
    log(f"[bench] channels: {nv} visibilities in {nc} channels, {N}^2 grids, {P} W-planes "
        f"({time.time() - t0:.1f}s)")
    grids = torch.empty((nc, N, N), dtype=torch.complex64, device="cuda")
    base = rgamd.params(N, obs.theta)
    best, info = rgtune.tuned_params(ctx, base, fv64, k64, grids, offsets=off, cache=tune_cache,
                                     retune=retune, reps=5)
    for _ in range(warmup):
        ctx.grid_channels(best, fv64, off, k64, grids)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        ctx.grid_channels(best, fv64, off, k64, grids)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    # the same call captured once and replayed
    graph = torch.cuda.CUDAGraph()
    gs = torch.cuda.Stream()
    gs.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(gs):
        with torch.cuda.graph(graph, stream=gs):
            ctx.grid_channels(best, fv64, off, k64, grids, stream=gs.cuda_stream)
    torch.cuda.current_stream().wait_stream(gs)
    graph.replay()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    ms_graph = e0.elapsed_time(e1) / steps
    # plain FP64, one call per channel
    g64 = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    plain = rgamd.params(N, obs.theta, rgamd.RG_MODE_PLAIN)
    errs = []
    check = set(np.linspace(0, nc - 1, 4).astype(int).tolist())
    torch.cuda.synchronize()
    t_plain = 0.0
    for c in range(nc):
        sub = rgamd.flat_vis(u[off[c]:off[c + 1]], v[off[c]:off[c + 1]], w[off[c]:off[c + 1]],
                             vis128[off[c]:off[c + 1]])
        e0.record()
        ctx.grid(plain, sub, k128, g64)
        e1.record()
        torch.cuda.synchronize()
        t_plain += e0.elapsed_time(e1)
        if c in check:
            a = grids[c].to(torch.complex128).clone()
            b = g64.clone()
            ctx.grid_to_image(a)
            ctx.grid_to_image(b)
            d = (a - b).abs()
            errs.append({"channel": c,
                         "max_rel": float(d.max() / b.abs().max()),
                         "rms_rel": float(torch.sqrt((d ** 2).mean()) /
                                          torch.sqrt((b.abs() ** 2).mean()))})
    del graph, grids
    ctx.thaw()  # the capture froze the context's scratch memory
    out = {"n_vis": nv, "channels": nc, "grid": N, "w_planes": P,
           "counts": {"min": int(np.diff(off).min()), "median": float(np.median(np.diff(off))),
                      "max": int(np.diff(off).max())},
           "tuning": {"source": info["source"], "key": info["key"], "chosen": info["chosen"],
                      "ms": info["ms"], "candidates": info["table"]},
           "fp32_batched": {"ms_per_step": ms, "mvis_s": nv / ms / 1e3},
           "fp32_batched_graph": {"ms_per_step": ms_graph, "mvis_s": nv / ms_graph / 1e3},
           "fp64_plain_per_channel": {"ms_per_step": t_plain, "mvis_s": nv / t_plain / 1e3,
                                      "launches": nc},
           "speedup_vs_fp64_plain": t_plain / ms,
           "error_fp32_batched_vs_fp64_plain": errs}
    log(f"[bench] channels: batched {ms:.3f} ms ({nv / ms / 1e3:.1f} Mvis/s), graph replay "
        f"{ms_graph:.3f} ms, plain per channel {t_plain:.2f} ms; launch parameters "
        f"({info['source']}) {info['chosen']}")
    return out


def bench_stress(ctx, steps, warmup, npix=4096, tune_cache=None, retune=False, no_tune=False):
    """Stress workload: 512 stations x 128 dumps x 16 channels (~260 M complex64
    visibilities), phase centre at zenith angle 60 deg, 4096^2 master grid
    (theta 0.2), W-towers subgrid 256, margin 16, winc 10.  FP32 batched timed;
    FP64 batched once for the FP32 error gate; the FP32 image against the FP64
    direct DFT on `npix` seeded pixels of the inner field.  W-kernels: 41
    residual-w planes, step 1/4 wavelength."""
    import torch

    import rgamd
    import rgtune
    import synth

    t0 = time.time()
    obs = synth.config_c5()
    nv, N = obs.n_vis, obs.grid_size
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    u, v, w = dev(obs.u), dev(obs.v), dev(obs.w)
    del obs.u, obs.v
    vis = torch.empty(nv, dtype=torch.complex64, device="cuda")
    chunk = 1 << 25
    for i in range(0, nv, chunk):  # predict in FP64, store complex64
        j = min(nv, i + chunk)
        tmp = torch.empty(j - i, dtype=torch.complex128, device="cuda")
        rgamd.predict_points(u[i:j], v[i:j], w[i:j], dev(obs.l), dev(obs.m), dev(obs.flux), tmp)
        vis[i:j] = tmp.to(torch.complex64)
        del tmp
    fv64 = rgamd.flat_vis(u, v, w, vis)
    # residual-w table over [-winc/2, winc/2] in steps of 1/4 wavelength: at
    # theta 0.2 the field edge has |n - 1| ~ 5e-3, so the plane quantisation
    # (half a step) is kept to ~1e-3 turns
    tab = dev(rgamd.make_wkernels(41, 8, 16, obs.theta, -5.0, 0.25))
    k128 = rgamd.wkern_desc(tab, -5.0, 0.25)
    tab32 = tab.to(torch.complex64)
    k64 = rgamd.wkern_desc(tab32, -5.0, 0.25)
    log(f"[bench] stress: {nv} visibilities, {N}^2 ({time.time() - t0:.1f}s)")
    p = rgamd.params(N, obs.theta, rgamd.RG_MODE_BATCHED, subgrid=256, margin=16, winc=10.0)
    g32 = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    tuning = None
    if not no_tune:
        p, info = rgtune.tuned_params(ctx, p, fv64, k64, g32, cache=tune_cache, retune=retune,
                                      reps=2)
        tuning = {"source": info["source"], "chosen": info["chosen"], "ms": info["ms"],
                  "candidates": info["table"]}
        log(f"[bench] stress FP32 launch parameters ({info['source']}): {info['chosen']}")
    for _ in range(warmup):
        ctx.grid(p, fv64, k64, g32)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        ctx.grid(p, fv64, k64, g32)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    st = ctx.stats()
    # FP64 batched reference on the same (complex64-rounded) visibilities
    vis128 = vis.to(torch.complex128)
    fv128 = rgamd.flat_vis(u, v, w, vis128)
    g64 = torch.empty_like(g32)
    e0.record()
    ctx.grid(p, fv128, k128, g64)
    e1.record()
    torch.cuda.synchronize()
    ms64 = e0.elapsed_time(e1)
    i32, i64 = g32.clone(), g64.clone()
    del g32, g64, vis128
    ctx.grid_to_image(i32)
    ctx.grid_to_image(i64)
    d = (i32 - i64).abs()
    err = {"max_rel": float(d.max() / i64.abs().max()),
           "rms_rel": float(torch.sqrt((d ** 2).mean()) / torch.sqrt((i64.abs() ** 2).mean()))}
    del d, i64
    dft = dft_errors(ctx, fv64, N, obs.lam, obs.theta, 16, {"wtowers_fp32_batched": i32}, npix,
                     inner_only=True)
    out = {"n_vis": nv, "grid": N, "subgrid": 256, "tasks": st.n_tasks, "towers": st.n_towers,
           "tuning": tuning,
           "fp32_batched": {"ms_per_step": ms, "mvis_s": nv / ms / 1e3},
           "fp64_batched": {"ms_per_step": ms64, "mvis_s": nv / ms64 / 1e3},
           "error_fp32_vs_fp64": err, "error_vs_dft": dft}
    log(f"[bench] stress: FP32 batched {ms:.1f} ms ({nv / ms / 1e3:.1f} Mvis/s), FP64 batched "
        f"{ms64:.1f} ms, {st.n_tasks} tasks, {st.n_towers} towers, err {err}")
    return out


def tower_work(st, nv, N, n, S, H, table_bytes, s_v=8, s_sub=8):
    """Compulsory HBM bytes and useful flop of one batched W-towers step, per
    kernel, from the call's own counts (st: rg_grid_stats of the step).  The
    byte count charges what this design must move through HBM at least once:
    the visibility stream and the sort's records, the band records, the
    subgrid bands written by the band pass and read by the column pass, the
    partial and tower subimages, the subimage FFTs, the fold and the kernel
    table once.  Taps gathered from the (L2-resident) table are not charged."""
    rec = 16  # Rec<float> / BandRec<float>
    nn = n * n
    lg = math.log2(n)
    band = H * n * s_sub
    b = {
        # count: uvw in, (key, rank, pos, kidx) out; scatter: that + vis in, records out
        "sort": nv * (24 + 16) + nv * (16 + s_v + rec),
        # binned records of every band window in, band records out
        "window_sort": st.n_records * rec * 2,
        "band_pass": st.n_records * rec + st.n_bands * band + table_bytes,
        "column_pass": st.n_bands * band + st.n_chunks * nn * s_sub,
        "reduce": st.n_chunks * nn * s_sub + 2 * st.n_towers * nn * s_sub,
        # forward FFT of every subimage (rows and columns, read + write), then
        # the fold reads them and updates the FP64 master grid
        "finish": 4 * st.n_towers * nn * s_sub + st.n_towers * nn * s_sub + 2 * N * N * 16,
    }
    f = {
        "band_pass": nv * 8 * S * S + st.n_bands * H * 5 * n * lg,
        "column_pass": st.n_tasks * (5 * nn * lg + 8 * nn),
        "finish": st.n_towers * 5 * nn * math.log2(nn),
    }
    return b, f


def wproj_work(st, nv, N, S, table_bytes, s_v=8, s_grid=8):
    """Compulsory HBM bytes and useful flop of one batched W-projection step:
    the visibility stream, the sort's records, the tile flushes (the grid,
    once) and the kernel table once."""
    rec = 16
    b = {"sort": nv * (24 + 16) + nv * (16 + s_v + rec),
         "tiles": st.n_records * rec + N * N * s_grid + table_bytes}
    f = {"tiles": nv * 8 * S * S}
    return b, f


def roofline(kind, head, st, nv, S, N, n, H, launch, table_bytes):
    """Dominant-kernel and whole-step rooflines.  `frac` (HBM) and `valu.frac`
    use the compulsory bytes and useful flop of tower_work / wproj_work over
    the kernel's average launch time; `effective` keeps the per-visibility
    count 24 + 8 + S^2 x 8 bytes (every tap gather charged, also those
    served from L2 / LDS) plus 3 n^2 x 8 bytes per (tower, floor) task, which
    is not a floor on HBM traffic and can exceed the HBM peak."""
    timers = head["timers"]
    if kind == "wproj":
        name, tkey, prof, part = ("k_tiles (W-projection)", "wproj_tiles",
                                  "wproj_tiles_fp32_c2.txt", "tiles")
        b, f = wproj_work(st, nv, N, S, table_bytes)
        eff = nv * (24 + 8 + S * S * 8)
    else:
        name, tkey, prof, part = ("k_tower_band_items (W-towers scatter + row IFFT)",
                                  "tower_tiles", "tower_band_items_fp32_c2.txt", "band_pass")
        b, f = tower_work(st, nv, N, n, S, H, table_bytes)
        eff = nv * (24 + 8 + S * S * 8) + st.n_tasks * 3 * n * n * 8
    t = timers.get(tkey)
    roof = None
    if t:
        launches = t["launches_per_step"]
        avg_ms = t["ms_per_step"] / launches
        byt = b[part] / launches
        ach = byt / (avg_ms * 1e-3) / 1e9
        traffic = profiled_traffic(prof, launch) if launch else None
        tf = f[part] / launches / (avg_ms * 1e-3) / 1e12
        roof = {"kernel": name, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                "bytes_per_launch": byt, "launches_per_step": launches, "avg_launch_ms": avg_ms,
                "measured_GBps": (traffic / (avg_ms * 1e-3) / 1e9) if traffic else None,
                "valu": {"achieved": tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": tf / FP32_PEAK_TFLOPS, "flop_per_launch": f[part] / launches,
                         # the published peak counts packed FP32 (two FMAs per lane);
                         # the build keeps packed FP32 instructions off, so its FP32
                         # issue limit is half of it
                         "peak_unpacked": FP32_PEAK_TFLOPS / 2,
                         "frac_unpacked": tf / (FP32_PEAK_TFLOPS / 2)},
                "effective": {"GBps": eff / launches / (avg_ms * 1e-3) / 1e9,
                              "bytes_per_launch": eff / launches,
                              "bytes_per_unit": {"visibility": 24 + 8 + S * S * 8}
                              | ({} if kind == "wproj" else {"task": 3 * n * n * 8})}}
    job_b, job_f = sum(b.values()), sum(f.values())
    t_roof = max(job_b / (HBM_PEAK_GBS * 1e9), job_f / (FP32_PEAK_TFLOPS * 1e12))
    job = {"bytes": job_b, "flop": job_f, "bytes_by_kernel": b, "flop_by_kernel": f,
           "t_roof_ms": t_roof * 1e3, "frac": t_roof * 1e3 / head["ms_per_step"],
           "counts": {k: getattr(st, k) for k in ("n_records", "n_bands", "n_chunks", "n_tasks",
                                                  "n_towers", "n_items")}}
    return roof, job


def dft_errors(ctx, fv, N, lam, theta, S, images, npix, seed=4, inner_only=False):
    """Images against the FP64 direct DFT on `npix` seeded pixels (of the whole
    image, or of the inner field only), after the layout map (DFT pixel (x, y)
    is gridder pixel ((N - x) % N, (N - y) % N)) and division by the taper
    A(l) A(m).  Reported over the inner field |l|, |m| < theta / 4 and over
    the valid field, the sampled pixels where A(l) A(m) >= 1e-2 (the taper
    division amplifies rounding by up to 100 there; beyond it, where A -> 0
    at the field edge, the quotient says nothing about the gridder)."""
    import torch

    import rgamd

    rng = np.random.default_rng(seed)
    if inner_only:
        side = N // 4 - 1  # pixels with |x - N/2| < N/8
        idx = rng.choice(side * side, size=min(npix, side * side), replace=False)
        ys, xs = np.divmod(idx, side)
        ys, xs = ys + N // 2 - N // 8 + 1, xs + N // 2 - N // 8 + 1
    else:
        idx = rng.choice(N * N, size=min(npix, N * N), replace=False)
        ys, xs = np.divmod(idx, N)
    px = torch.from_numpy(xs.astype(np.int32)).cuda()
    py = torch.from_numpy(ys.astype(np.int32)).cuda()
    out = torch.empty(len(idx), dtype=torch.complex128, device="cuda")
    t = time.perf_counter()
    ctx.dft(fv, px, py, N, lam, out, rgamd.RG_DFT_FP64)
    torch.cuda.synchronize()
    t_dft = time.perf_counter() - t
    d = out.cpu().numpy()
    taper = rgamd.pswf_eval(math.pi * S / 4, 2.0 * (np.arange(N) - N // 2) / N)
    r, c = (N - xs) % N, (N - ys) % N
    inner = (np.abs(r - N // 2) < N // 8) & (np.abs(c - N // 2) < N // 8)
    valid = taper[c] * taper[r] >= 1e-2
    res = {"pixels": int(len(idx)), "pixels_valid_field": int(valid.sum()),
           "pixels_inner_field": int(inner.sum()), "dft_seconds": t_dft}

    def rel(e, sel):
        return (float(e[sel].max() / np.abs(d[sel]).max()),
                float(np.sqrt(np.mean(e[sel] ** 2)) / np.sqrt(np.mean(np.abs(d[sel]) ** 2))))

    for name, img in images.items():
        g = img.cpu().numpy()[r, c] / (taper[c] * taper[r])
        e = np.abs(g - d)
        vm, vr = rel(e, valid)
        im, ir = rel(e, inner)
        res[name] = {"valid_max_rel": vm, "valid_rms_rel": vr, "inner_max_rel": im,
                     "inner_rms_rel": ir}
    log(f"[bench] DFT oracle ({len(idx)} pixels, {t_dft:.1f}s): {res}")
    return res


def host_threads():
    """Host threads for the CPU baselines: every core this process may run on
    (its CPU affinity), capped by OMP_NUM_THREADS when that is set (a batch
    system's per-job core share), and the facts reported with each CPU
    number."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = nproc
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = aff
    if omp and omp.strip().isdigit() and int(omp) > 0:
        threads = min(aff, int(omp))
    return threads, {"nproc": nproc, "affinity_cpus": aff, "OMP_NUM_THREADS": omp,
                     "threads": threads, "cpu": _cpu_model()}


def cpu_baseline(obs, vis128, tab_p, w_min, w_step, seconds):
    """The oracle's OpenMP W-projection (fast build; grid rows partitioned
    over the threads, race-free) on all available host threads, on a bounded
    prefix of the same visibility stream.  (The reference's CPU W-projection,
    grid.c:187-202, is serial.)"""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc

    threads, host = host_threads()
    vis = vis128.cpu().numpy()
    table = tab_p.cpu().numpy()
    N = obs.grid_size
    n0 = min(obs.n_vis, 20000)
    t = time.perf_counter()
    orc.grid_wprojection(N, obs.theta, obs.u[:n0], obs.v[:n0], obs.w[:n0], vis[:n0], table, w_min,
                         w_step, nthreads=threads, fast=True)
    dt = time.perf_counter() - t
    ns = int(min(obs.n_vis, max(n0, n0 * seconds / max(dt, 1e-6))))
    t = time.perf_counter()
    orc.grid_wprojection(N, obs.theta, obs.u[:ns], obs.v[:ns], obs.w[:ns], vis[:ns], table, w_min,
                         w_step, nthreads=threads, fast=True)
    dt = time.perf_counter() - t
    log(f"[bench] CPU W-projection: {ns} vis in {dt:.2f}s on {threads} threads")
    return {"value": ns / dt / 1e6, "unit": "Mvis/s", "cores": threads, "kind": "port",
            "sample": f"first {ns} visibilities of the stream, W-projection (OpenMP, grid rows "
                      f"partitioned over {threads} threads, all busy; -O3 -ffast-math), 2048^2 grid",
            "host": host}


def cpu_baseline_wtowers(obs, vis128, tab_t, n, margin, winc, ntowers):
    """The oracle's OpenMP W-towers (fast build, OpenMP over towers with
    schedule(dynamic), as grid.c:413-416) on all available host threads over
    the whole visibility stream (W-towers time is dominated by the
    per-(tower, floor) FFTs, so a prefix of the stream is not representative).
    Only the active towers can use threads: at most min(threads, towers) are
    busy.  Returns the record and the CPU grid."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc

    threads, host = host_threads()
    ns = obs.n_vis
    vis = vis128[:ns].cpu().numpy()
    table = tab_t.cpu().numpy()
    t = time.perf_counter()
    g, _, _ = orc.grid_wtowers(obs.grid_size, obs.theta, obs.u[:ns], obs.v[:ns], obs.w[:ns], vis,
                               table, -5.0, 1.0, n, margin, winc, nthreads=threads, fast=True)
    dt = time.perf_counter() - t
    busy = max(1, min(threads, ntowers))
    log(f"[bench] CPU W-towers: {ns} vis in {dt:.2f}s on {threads} threads ({busy} busy)")
    return ({"value": ns / dt / 1e6, "unit": "Mvis/s", "cores": busy, "kind": "port",
             "sample": (f"all {ns} visibilities of the workload, W-towers (OpenMP over towers: "
                        f"{threads} threads started, {ntowers} active towers, so {busy} busy; "
                        "-O3 -ffast-math, the oracle's own radix-2 FFT: FFTW is absent), 2048^2 grid"),
             "host": host, "seconds": dt}, g)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
