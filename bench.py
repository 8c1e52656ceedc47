#!/usr/bin/env python3
"""Benchmark of the IMEXLBM hot path on one MI355X.

Workload: the periodic D3Q27 Taylor-Green vortex at 512^3, Re=1600, u0=0.05,
rho0=1 (analytic initial field, synthetic: no input data), advanced by the
fused pull-stream + BGK-collide kernel with in-kernel periodic wrap.  One
"step" is one lattice update of every cell.  Each storage format (FP64 plain,
FP32, FP16 and BF16 shifted populations with FP32 arithmetic) runs in its own
context; the headline value is the FP16-stored path, the others are reported
beside it.  Before the warm-up each context runs the library's launch-shape
autotuner on its live arrays (``--no-autotune`` keeps the per-format
defaults); the chosen shape is reported under ``launch``.

MLUPS = cells * steps / wall seconds between two device synchronisations,
the reference's formula (src/main.cpp:77-78) with the conserved-quantity
reduction outside the timed region.  Where the domain allows it the library
advances two steps per pass (temporal blocking, bit-identical to two single
steps).  The roofline figure divides the algorithmic bytes of the updates one
launch performs (2 * 27 * sizeof(storage) per cell per update) by the mean
duration of the launch, measured with HIP events on the stream the kernel
runs on, against the 8 TB/s HBM peak; ``hbm_gbs`` is the measured HBM traffic
per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE) over the same duration.

The CPU baseline is the FP64 oracle (the reference's collide / 26-slab ghost
copy / pull-stream structure in C with OpenMP) timed on this host's cores on
a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PI_LIT = 3.1415926


def tau0(u0, n, re):
    # System::System, src/System.cpp:28
    return u0 * n / re / (1.0 / 3.0) / 2.0 / PI_LIT


def run_mode(capi, prec, n, re, u0, steps, warmup, autotune, use_graph, mode="wrap", tiles=8, halo=1,
             two_step=-1, whole_x=0):
    dom = {"nx": n, "ny": n, "nz": n, "tau0": tau0(u0, n, re), "rho0": 1.0, "u0": u0}
    kw = {"two_step": two_step}
    if mode == "tiled":
        kw.update({"mode": capi.TILED, "ntiles": tiles, "halo_depth": halo, "whole_x": whole_x})
    with capi.Context(dom, precision=prec, use_graph=int(use_graph), **kw) as ctx:
        ctx.initialize()
        ctx.step(1)                       # first collision: the state becomes post-collision
        tuned_ms = ctx.autotune(reps=3) if autotune else None
        ctx.step(warmup)
        ctx.synchronize()
        wall_ms, kern_ms = ctx.timed_steps(steps)
        cons = ctx.conserved()
        info = ctx.info()
    cells = info["cells"]
    bpu = info["bytes_per_update"]
    step_s = wall_ms / 1e3 / steps
    kstep_s = kern_ms / 1e3 / steps
    mlups = cells * steps / (wall_ms / 1e3) / 1e6
    achieved = bpu * cells / kstep_s / 1e9
    return {
        "precision": prec,
        "mlups": round(mlups, 1),
        "ms_per_step": round(step_s * 1e3, 4),
        "kernel_ms_per_step": round(kstep_s * 1e3, 4),
        "bytes_per_update": bpu,
        "achieved_gbs": round(achieved, 1),
        "roofline_frac": round(achieved / HBM_PEAK_GBS, 4),
        "roofline_mlups": round(HBM_PEAK_GBS * 1e9 / bpu / 1e6, 1),
        "launch": {k: info[k] for k in ("block", "cells_per_thread", "persistent", "nontemporal",
                                        "dir_pad", "use_graph", "xcd_groups", "xcd_run", "io16",
                                        "two_step", "two_step_tile", "two_step_zseg", "two_step_xcds",
                                        "two_step_prefetch", "two_step_io")},
        "updates_per_launch": 2 if info["two_step"] else 1,
        "autotune_best_ms": None if tuned_ms is None else round(tuned_ms, 4),
        "mass": cons[0], "ke": cons[4],
    }


def cpu_baseline(target_s=12.0):
    """The FP64 oracle on this host: 128^3 Taylor-Green, reference structure."""
    from oracle import oracle
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    oracle.set_threads(cores)
    n, re, u0 = 128, 1600, 0.05
    with oracle.Oracle(n, n, n, tau0(u0, n, re), 1.0, u0) as o:
        o.initialize()
        t0 = time.perf_counter()
        o.step(2)                         # first interval, discarded
        t1 = time.perf_counter()
        per = (t1 - t0) / 2
        k = max(2, min(200, int(target_s / max(per, 1e-6))))
        t2 = time.perf_counter()
        o.step(k)
        t3 = time.perf_counter()
    mlups = n ** 3 * k / (t3 - t2) / 1e6
    return {"value": round(mlups, 2), "unit": "MLUPS", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle (collide + 26-slab periodic ghost copy + pull stream, "
                      f"OpenMP), TGV {n}^3 Re={re}, {k} timed steps after 2 discarded, "
                      f"{t3 - t2:.1f} s"}


def traffic_from_profiles(prec, n, updates):
    """HBM bytes per launch of the step kernel (one step, or two steps per
    pass) from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary, if one
    exists for this workload."""
    path = os.path.join(ROOT, "profiles", f"pmc_fused_step_tgv{n}.md")
    if not os.path.exists(path):
        return None
    key = f"traffic{updates}_bytes="
    pat = re.compile(r"^\|\s*" + re.escape(prec) + r"\s*\|.*?\|\s*" + re.escape(key) + r"(\d+)", re.M)
    m = pat.search(open(path).read())
    return int(m.group(1)) if m else None


def roofline(h, traffic, copy_gbs):
    """Roofline of the dominant kernel.  achieved = algorithmic bytes per
    lattice update (2 * 27 * sizeof) x updates per launch / mean launch time.
    With two steps per pass the kernel reads and writes each population once
    per two updates, so HBM moves about half the algorithmic bytes: hbm_gbs
    (measured traffic / launch time) is what the memory system delivers."""
    ups = h["updates_per_launch"]
    launch_s = h["kernel_ms_per_step"] * ups / 1e3
    out = {
        "bound": "hbm",
        "achieved": h["achieved_gbs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": h["roofline_frac"],
        "traffic": traffic,
        "updates_per_launch": ups,
        "launch_ms": round(launch_s * 1e3, 4),
        "copy_ceiling": round(copy_gbs, 1),
        "frac_of_copy": round(h["achieved_gbs"] / copy_gbs, 4),
    }
    if traffic:
        out["hbm_gbs"] = round(traffic / launch_s / 1e9, 1)
        out["hbm_frac"] = round(traffic / launch_s / 1e9 / HBM_PEAK_GBS, 4)
    if ups == 2:
        out["note"] = ("two lattice updates per launch (temporal blocking): achieved counts "
                       "2*27*sizeof bytes per cell per update, HBM moves traffic bytes per launch "
                       "(hbm_gbs)")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512, help="cube edge (default 512)")
    ap.add_argument("--re", type=int, default=1600)
    ap.add_argument("--u0", type=float, default=0.05)
    ap.add_argument("--modes", default="fp64,fp32,fp16,bf16")
    ap.add_argument("--headline", default="fp16")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the launch-shape autotuner (per-format defaults)")
    ap.add_argument("--graph", action="store_true", help="replay steps from a hipGraph")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mode", choices=("wrap", "tiled"), default="wrap",
                    help="periodic wrap in the kernel, or subdomain tiles with device pack/unpack")
    ap.add_argument("--tiles", type=int, default=8)
    ap.add_argument("--halo", type=int, default=1, help="ghost layers filled per exchange (tiled)")
    ap.add_argument("--whole-x", action="store_true",
                    help="tiled: cut y and z only and wrap x in the kernels (no x ghosts)")
    ap.add_argument("--two-step", choices=("auto", "on", "off"), default="auto",
                    help="two steps per pass (temporal blocking): auto lets the autotuner decide")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(run with --gpus 1).", file=sys.stderr)
        return 2

    from imexlbm_amd import capi
    if capi.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1

    modes = [m for m in args.modes.split(",") if m]
    if args.headline not in modes:
        modes.append(args.headline)
    results = {}
    for m in modes:
        print(f"bench: {m} {args.n}^3 steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)
        results[m] = run_mode(capi, m, args.n, args.re, args.u0, args.steps, args.warmup,
                              not args.no_autotune, args.graph, args.mode, args.tiles, args.halo,
                              {"auto": -1, "on": 1, "off": 0}[args.two_step], int(args.whole_x))
        print(f"bench: {m} {results[m]['mlups']} MLUPS, {results[m]['achieved_gbs']} GB/s",
              file=sys.stderr, flush=True)

    h = results[args.headline]
    print("bench: copy probe", file=sys.stderr, flush=True)
    copy_gbs = capi.copy_probe(8 << 30, 10)
    cpu = None
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline()
    line = {
        "metric": "MLUPS + % HBM roofline, D3Q27 TGV 512^3: FP32/FP16-store vs FP64 plain",
        "value": h["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.headline == "fp64" else "f32",
        "data": "synthetic (analytic Taylor-Green initial field)",
        "config": {
            "workload": f"D3Q27 BGK Taylor-Green vortex {args.n}^3 periodic, Re={args.re}, "
                        f"u0={args.u0}, rho0=1, fused pull-stream+collide, "
                        + ("in-kernel wrap" if args.mode == "wrap" else
                           f"{args.tiles} tiles{' (x whole)' if args.whole_x else ''}, "
                           f"{args.halo}-layer device pack/unpack"),
            "storage": args.headline,
            "cells": args.n ** 3,
            "parallelism": "single GPU",
            "modes": results,
            "fp16_over_fp64": (round(results["fp16"]["mlups"] / results["fp64"]["mlups"], 3)
                               if "fp16" in results and "fp64" in results else None),
        },
        "roofline": roofline(h, traffic_from_profiles(args.headline, args.n, h["updates_per_launch"]),
                             copy_gbs),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
