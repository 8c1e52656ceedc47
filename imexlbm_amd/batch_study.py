"""Batched Taylor-Green instances: one launch per step for all of them.

64 independent periodic TGV instances whose cube edge is drawn from
{32, 48, 64, 96, 128}, Re log-uniform in [100, 3200] and u0 uniform in
[0.02, 0.08] (numpy default_rng, fixed seed) advance together: one work-tile
table (or, two steps per pass, one column-tile table) covers every instance,
each step or pair of steps is one launch, replayed from a captured hipGraph,
and the launch shape comes from the library's autotuner.  Reported beside it:

  sequential_launch   the same context, one launch per instance per step
  one_after_another   each instance in its own context, run to completion in turn
  fp64 check          every FP32 instance against the FP64 batch (itself
                      bit-identical per instance to the single-instance FP64
                      path): velocity L2 and final KE error

Usage: python -m imexlbm_amd.batch_study [--instances 64] [--steps 1000]
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import capi
from .precision_study import tau0

EDGES = (32, 48, 64, 96, 128)


def instances(count=64, seed=20261203):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(EDGES))
        re = float(math.exp(rng.uniform(math.log(100.0), math.log(3200.0))))
        u0 = float(rng.uniform(0.02, 0.08))
        out.append({"nx": n, "ny": n, "nz": n, "tau0": tau0(u0, n, re), "rho0": 1.0, "u0": u0,
                    "re": re})
    return out


def _timed(ctx, steps):
    ctx.step(2)
    ctx.synchronize()
    wall, kern = ctx.timed_steps(steps)
    return wall, kern


def run(count=64, steps=1000, prec="fp32", log=sys.stderr):
    doms = instances(count)
    cells = sum(d["nx"] * d["ny"] * d["nz"] for d in doms)
    res = {"instances": count, "steps": steps, "precision": prec, "cells": cells,
           "edges": [d["nx"] for d in doms]}

    # batched: one launch per step, graph replay, autotuned launch shape
    with capi.Context(doms, precision=prec, use_graph=1) as c:
        c.initialize()
        c.step(1)
        res["autotune_ms"] = c.autotune(reps=5)
        info = c.info()
        res["launch"] = {k: info[k] for k in ("block", "cells_per_thread", "persistent", "nontemporal",
                                             "xcd_groups", "work_tiles", "two_step", "two_step_tile",
                                             "two_step_zseg", "two_step_xcds", "two_step_prefetch")}
        wall, kern = _timed(c, steps)
        res["batched_mlups"] = cells * steps / (wall / 1e3) / 1e6
        res["batched_ms_per_step"] = wall / steps
        print(f"batch: batched {res['batched_mlups']:.0f} MLUPS", file=log, flush=True)
        ke_b = [c.conserved(k)[4] for k in range(count)]
        macro_b = [c.get_macro(k) for k in range(count)]

    # same context, one launch per instance per step
    with capi.Context(doms, precision=prec, sequential=1) as c:
        c.initialize()
        c.step(1)
        wall, _ = _timed(c, steps)
        res["sequential_launch_mlups"] = cells * steps / (wall / 1e3) / 1e6
        print(f"batch: sequential launches {res['sequential_launch_mlups']:.0f} MLUPS", file=log, flush=True)

    # each instance on its own, one after another
    total = 0.0
    for d in doms:
        with capi.Context(d, precision=prec) as c:
            c.initialize()
            c.step(1)
            wall, _ = _timed(c, steps)
            total += wall
    res["one_after_another_mlups"] = cells * steps / (total / 1e3) / 1e6
    print(f"batch: one after another {res['one_after_another_mlups']:.0f} MLUPS", file=log, flush=True)

    # FP64 reference batch: same steps (1 + 2 + steps)
    with capi.Context(doms, precision="fp64", use_graph=1) as c:
        c.initialize()
        c.step(steps + 3)
        ke64 = [c.conserved(k)[4] for k in range(count)]
        errs = []
        for k in range(count):
            ref = c.get_macro(k)
            num = sum(float(np.sum((b - a) ** 2)) for a, b in zip(ref[1:], macro_b[k][1:]))
            den = sum(float(np.sum(a ** 2)) for a in ref[1:])
            errs.append(math.sqrt(num / den))
    ke_err = [abs(a - b) / b for a, b in zip(ke_b, ke64)]
    # an instance whose FP64 run is itself non-finite is an unstable deck
    # (BGK at omega -> 2 on a coarse grid), not a precision effect
    ok64 = [math.isfinite(k) for k in ke64]
    fin = [e for e, ok in zip(errs, ok64) if ok and math.isfinite(e)]
    res["unstable_fp64"] = sum(1 for ok in ok64 if not ok)
    res["unstable_only_reduced"] = sum(1 for e, ok in zip(errs, ok64) if ok and not math.isfinite(e))
    res["vel_l2_max"] = max(fin) if fin else float("nan")
    res["vel_l2_median"] = float(np.median(fin)) if fin else float("nan")
    res["ke_err_max"] = max((k for k, ok in zip(ke_err, ok64) if ok and math.isfinite(k)), default=float("nan"))
    res["per_instance"] = [{"n": d["nx"], "re": round(d["re"], 1), "u0": round(d["u0"], 4),
                            "omega": round(1.0 / (d["tau0"] + 0.5), 5), "fp64_stable": ok,
                            "vel_l2": e, "ke_err": k}
                           for d, e, k, ok in zip(doms, errs, ke_err, ok64)]
    res["speedup_vs_one_after_another"] = res["batched_mlups"] / res["one_after_another_mlups"]
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--prec", default="fp32")
    a = ap.parse_args(argv)
    print(json.dumps(run(a.instances, a.steps, a.prec)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
