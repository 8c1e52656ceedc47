"""Time fixed launch shapes of the step kernels on the Taylor-Green vortex.

Each shape runs in its own context: initialise, one collision, two warm-up
steps, then `--passes` timed launches of the chosen kernel (HIP events on the
context's stream).  Used to profile one launch shape at a time (rocprofv3
counters are per dispatch), and to compare shapes outside the autotuner.

A shape is prec:mode:tile:zseg, where mode is wrap or tiled (2x2x2 tiles with
a 3-layer ghost fill through pack + unpack buffers), tile the two-step column
tile (-1: one pass per step) and zseg the planes per sweep; the tile order
over the XCDs is the default one (each XCD its own range of columns).  The key printed
with each result is the one bench.py reports as roofline.shape_key.

Usage: python -m imexlbm_amd.shape_probe --n 512 --shapes fp16:wrap:4:512,fp64:wrap:3:128
"""
from __future__ import annotations

import argparse
import sys

from . import capi
from .precision_study import tau0


def probe(n, prec, mode, tile, zseg, passes=8, re=1600, u0=0.05):
    dom = {"nx": n, "ny": n, "nz": n, "tau0": tau0(u0, n, re), "rho0": 1.0, "u0": u0}
    kw = {"two_step": 0} if tile < 0 else {"two_step": 1, "two_step_tile": tile, "two_step_zseg": zseg}
    if mode == "tiled":
        kw.update({"mode": capi.TILED, "ntiles": 8, "halo_depth": 3, "halo_direct": 0})
    with capi.Context(dom, precision=prec, **kw) as c:
        c.initialize()
        c.step(1)
        c.step(2)
        c.synchronize()
        ups = 1 if tile < 0 else 2
        wall, kern = c.timed_steps(passes * ups)
        info = c.info()
    if tile < 0:
        key = f"{prec}:{n}:{mode}:step:b{info['block']}:c{info['cells_per_thread']}"
    else:
        key = f"{prec}:{n}:{mode}:ts{info['two_step_tile']}:z{info['two_step_zseg']}:x{info['two_step_xcds']}"
    return key, kern / passes


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--shapes", required=True)
    ap.add_argument("--passes", type=int, default=8)
    ap.add_argument("--copy-gib", type=int, default=0,
                    help="also run the copy probe over this many GiB (a dispatch of known size)")
    a = ap.parse_args(argv)
    if a.copy_gib:
        gbs = capi.copy_probe(a.copy_gib << 30, 2)
        print(f"copy {a.copy_gib} GiB: {gbs:.1f} GB/s", flush=True)
    for s in a.shapes.split(","):
        prec, mode, tile, zseg = s.split(":")
        key, ms = probe(a.n, prec, mode, int(tile), int(zseg), a.passes)
        print(f"{key} {ms:.4f} ms per launch", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
