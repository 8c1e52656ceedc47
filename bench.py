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
defaults); the chosen shape is reported under ``launch``.  Two sub-records at
256^3 FP64 follow: periodic wrap, and 2x2x2 subdomain tiles with a 3-layer
ghost fill through pack + unpack buffers.

MLUPS = cells * steps / wall seconds between two device synchronisations,
the reference's formula (src/main.cpp:77-78) with the conserved-quantity
reduction outside the timed region.  Where the domain allows it the library
advances two steps per pass (temporal blocking, bit-identical to two single
steps).

Roofline of the step kernel: a launch must read and write every population
once, 2 * 27 * sizeof(storage) * cells bytes, whether it performs one
lattice update or two.  ``roofline.achieved`` is those bytes over the mean
launch duration (HIP events on the stream the kernel runs on) and ``frac``
its share of the 8 TB/s HBM peak.  ``mlups_over_roofline`` is the MLUPS over
the MLUPS that 8 TB/s would give at 2 * 27 * sizeof bytes per update (the
form in which the 80%-of-roofline target is stated); with two updates per
launch it can pass 1.  ``traffic`` is the HBM bytes per launch that rocprofv3
FETCH_SIZE + WRITE_SIZE measured for the same kernel and launch shape, from
the committed summary under profiles/, or null when that shape has not been
profiled.

The CPU baseline is the FP64 oracle (the reference's collide / 26-slab ghost
copy / pull-stream structure in C with OpenMP) timed in a child process on
this host's cores: the 32^3 deck (Re=100, 200 steps, reports every 50) and a
shortened 256^3 Re=1600 deck, the first interval discarded.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PI_LIT = 3.1415926
LAUNCH_KEYS = ("block", "cells_per_thread", "persistent", "nontemporal", "dir_pad", "use_graph",
               "xcd_groups", "xcd_run", "io16", "two_step", "two_step_tile", "two_step_zseg",
               "two_step_xcds", "two_step_prefetch", "two_step_io", "halo_direct", "halo_depth")


def tau0(u0, n, re):
    # System::System, src/System.cpp:28
    return u0 * n / re / (1.0 / 3.0) / 2.0 / PI_LIT


def shape_key(prec, n, info, mode="wrap"):
    """Key of a launch shape in the committed HBM-traffic summaries (with the
    padding between direction planes when there is one)."""
    if info["two_step"]:
        s = f"ts{info['two_step_tile']}:z{info['two_step_zseg']}:x{info['two_step_xcds']}"
    else:
        s = f"step:b{info['block']}:c{info['cells_per_thread']}"
    pad = f":p{info['dir_pad']}" if info.get("dir_pad") else ""
    return f"{prec}:{n}:{mode}:{s}{pad}"


def traffic_from_profiles(key, n):
    """HBM bytes per launch for this kernel and launch shape, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary, or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_fused_step_tgv{n}.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"key=" + re.escape(key) + r" traffic=(\d+)", open(path).read())
    return int(m.group(1)) if m else None


def run_mode(capi, prec, n, re_, u0, steps, warmup, autotune, use_graph, mode="wrap", tiles=8, halo=1,
             two_step=-1, whole_x=0, halo_direct=-1, pin=None):
    dom = {"nx": n, "ny": n, "nz": n, "tau0": tau0(u0, n, re_), "rho0": 1.0, "u0": u0}
    kw = {"two_step": two_step}
    if mode == "tiled":
        kw.update({"mode": capi.TILED, "ntiles": tiles, "halo_depth": halo, "whole_x": whole_x,
                   "halo_direct": halo_direct})
    if pin:
        kw.update(pin)
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
    ups = 2 if info["two_step"] else 1
    mlups = cells * steps / (wall_ms / 1e3) / 1e6
    launch_ms = kern_ms / steps * ups           # mean duration of one step-kernel launch
    launch_bytes = bpu * cells                  # every population read once and written once
    kernel_gbs = launch_bytes / (launch_ms / 1e3) / 1e9
    roof_mlups = HBM_PEAK_GBS * 1e9 / bpu / 1e6
    return {
        "precision": prec,
        "mode": mode,
        "mlups": round(mlups, 1),
        "ms_per_step": round(wall_ms / steps, 4),
        "kernel_ms_per_step": round(kern_ms / steps, 4),
        "bytes_per_update": bpu,
        "updates_per_launch": ups,
        "launch_ms": round(launch_ms, 4),
        "launch_bytes": launch_bytes,
        "kernel_gbs": round(kernel_gbs, 1),
        "kernel_frac": round(kernel_gbs / HBM_PEAK_GBS, 4),
        "roofline_mlups": round(roof_mlups, 1),
        "mlups_over_roofline": round(mlups / roof_mlups, 4),
        "launch": {k: info[k] for k in LAUNCH_KEYS},
        "shape_key": shape_key(prec, n, info, mode),
        "autotune_best_ms": None if tuned_ms is None else round(tuned_ms, 4),
        "mass": cons[0], "ke": cons[4],
    }


def roofline(h, n, copy_gbs):
    """Roofline object of the dominant (step) kernel."""
    traffic = traffic_from_profiles(h["shape_key"], n)
    out = {
        "bound": "hbm",
        "achieved": h["kernel_gbs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": h["kernel_frac"],
        "traffic": traffic,
        "launch_bytes": h["launch_bytes"],
        "launch_ms": h["launch_ms"],
        "updates_per_launch": h["updates_per_launch"],
        "mlups_over_roofline": h["mlups_over_roofline"],
        "copy_ceiling": round(copy_gbs, 1),
        "frac_of_copy": round(h["kernel_gbs"] / copy_gbs, 4),
        "shape_key": h["shape_key"],
    }
    if traffic:
        out["hbm_gbs"] = round(traffic / (h["launch_ms"] / 1e3) / 1e9, 1)
        out["traffic_over_launch_bytes"] = round(traffic / h["launch_bytes"], 4)
    else:
        out["traffic_note"] = "no rocprofv3 FETCH_SIZE/WRITE_SIZE summary for this launch shape"
    return out


# ---------------------------------------------------------------------------
# CPU baseline (child process: the oracle with OpenMP, no GPU)
# ---------------------------------------------------------------------------
def _physical_cores():
    pairs = set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("physical id"):
                    phys = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    core = line.split(":")[1].strip()
                elif not line.strip():
                    if core is not None:
                        pairs.add((phys, core))
                    phys = core = None
    except OSError:
        return None
    return len(pairs) or None


def _cpu_child(target_s):
    """Runs in a child process: the 32^3 deck and a shortened 256^3 deck."""
    from oracle import oracle
    threads = oracle.max_threads()

    def deck(n, re_, u0, inter, intervals):
        # the driver's cadence (src/main.cpp:31-98): MLUPS per interval of
        # `inter` steps, conserved sums outside the timed intervals; the
        # first interval is discarded (tools/parse_scaling.sh:60)
        with oracle.Oracle(n, n, n, tau0(u0, n, re_), 1.0, u0) as o:
            o.initialize()
            rates = []
            for k in range(intervals):
                t0 = time.perf_counter()
                o.step(inter)
                dt = time.perf_counter() - t0
                rates.append(n ** 3 * inter / dt / 1e6)
                o.conserved()
            return rates

    r32 = deck(32, 100, 0.05, 50, 4)
    # 256^3: size the intervals from a one-step probe to about target_s
    with oracle.Oracle(256, 256, 256, tau0(0.05, 256, 1600), 1.0, 0.05) as o:
        o.initialize()
        t0 = time.perf_counter()
        o.step(1)
        per = time.perf_counter() - t0
    inter = max(2, min(50, int(target_s / 3 / max(per, 1e-6))))
    r256 = deck(256, 1600, 0.05, inter, 4)
    mean = lambda v: sum(v) / len(v)
    return {"threads": threads, "mlups32": mean(r32[1:]), "mlups256": mean(r256[1:]),
            "inter256": inter, "steps256": 4 * inter}


def cpu_baseline(target_s=15.0):
    env = dict(os.environ)
    env.setdefault("OMP_PROC_BIND", "close")
    env.setdefault("OMP_PLACES", "cores")
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    if "OMP_NUM_THREADS" not in env:
        env["OMP_NUM_THREADS"] = str(affinity)
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", str(target_s)],
                       env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=600)
    if r.returncode != 0:
        print(r.stderr[-2000:], file=sys.stderr)
        return None
    c = json.loads(r.stdout.strip().splitlines()[-1])
    return {"value": round(c["mlups256"], 2), "unit": "MLUPS", "cores": c["threads"], "kind": "port",
            "sample": (f"FP64 oracle (collide + 26-slab periodic ghost copy + pull stream, OpenMP "
                       f"{c['threads']} threads, OMP_PROC_BIND={env['OMP_PROC_BIND']}), TGV 256^3 "
                       f"Re=1600, 4 intervals of {c['inter256']} steps, first discarded"),
            "deck32_mlups": round(c["mlups32"], 2),
            "deck32_sample": "TGV 32^3 Re=100, 200 steps reported every 50, first interval discarded",
            "host_physical_cores": _physical_cores(),
            "host_affinity_cpus": affinity}


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
    ap.add_argument("--no-sub", action="store_true", help="skip the 256^3 FP64 sub-records")
    ap.add_argument("--mode", choices=("wrap", "tiled"), default="wrap",
                    help="periodic wrap in the kernel, or subdomain tiles with device pack/unpack")
    ap.add_argument("--tiles", type=int, default=8)
    ap.add_argument("--halo", type=int, default=1, help="ghost layers filled per exchange (tiled)")
    ap.add_argument("--whole-x", action="store_true",
                    help="tiled: cut y and z only and wrap x in the kernels (no x ghosts)")
    ap.add_argument("--two-step", choices=("auto", "on", "off"), default="auto",
                    help="two steps per pass (temporal blocking): auto lets the autotuner decide")
    ap.add_argument("--pin", default="",
                    help="fix launch-shape fields, e.g. two_step_tile=4,two_step_zseg=256 "
                         "(implies --no-autotune)")
    ap.add_argument("--cpu-child", type=float, default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_child is not None:
        print(json.dumps(_cpu_child(args.cpu_child)), flush=True)
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(run with --gpus 1).", file=sys.stderr)
        return 2

    from imexlbm_amd import capi
    if capi.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1

    pin = {}
    for kv in filter(None, args.pin.split(",")):
        k, v = kv.split("=")
        pin[k.strip()] = int(v)
    autotune = not (args.no_autotune or pin)
    two = {"auto": -1, "on": 1, "off": 0}[args.two_step]

    modes = [m for m in args.modes.split(",") if m]
    if args.headline not in modes:
        modes.append(args.headline)
    results = {}
    for m in modes:
        print(f"bench: {m} {args.n}^3 steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)
        results[m] = run_mode(capi, m, args.n, args.re, args.u0, args.steps, args.warmup, autotune,
                              args.graph, args.mode, args.tiles, args.halo, two, int(args.whole_x), pin=pin)
        r = results[m]
        print(f"bench: {m} {r['mlups']} MLUPS, launch {r['launch_ms']} ms, kernel {r['kernel_gbs']} GB/s "
              f"({r['kernel_frac']:.3f} of peak), {r['shape_key']}", file=sys.stderr, flush=True)

    sub = {}
    if not args.no_sub and args.mode == "wrap":
        # the 256^3 FP64 deck: periodic wrap and 2x2x2 tiles with a 3-layer
        # ghost fill through pack + unpack buffers
        for name, kw in (("fp64_256_wrap", {}),
                         ("fp64_256_tiled_2x2x2", {"mode": "tiled", "tiles": 8, "halo": 3,
                                                   "halo_direct": 0})):
            print(f"bench: {name} steps={args.steps}", file=sys.stderr, flush=True)
            r = run_mode(capi, "fp64", 256, args.re, args.u0, args.steps, args.warmup, True, args.graph,
                         two_step=-1, **kw)
            r["roofline"] = roofline(r, 256, float("nan"))
            r["roofline"].pop("copy_ceiling"), r["roofline"].pop("frac_of_copy")
            if kw:
                r["roofline"]["note"] = ("launch_ms brackets each pass together with the pack and unpack "
                                         "ghost-fill launches before it; traffic is the pass alone")
            sub[name] = r
            print(f"bench: {name} {r['mlups']} MLUPS ({r['mlups_over_roofline']:.3f} of the roofline "
                  f"MLUPS), kernel {r['kernel_frac']:.3f} of peak", file=sys.stderr, flush=True)
        sub["tiled_over_wrap"] = round(sub["fp64_256_tiled_2x2x2"]["mlups"] / sub["fp64_256_wrap"]["mlups"], 4)

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
            "accuracy_note": ("FP16-stored meets velocity L2 <= 1e-3 and KE error <= 0.5% against FP64 "
                              "over 2000 steps; BF16 does not (velocity L2 ~1e-2, KE ~1%)"),
            "deck256": sub or None,
        },
        "roofline": roofline(h, args.n, copy_gbs),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
