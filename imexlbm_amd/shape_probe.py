"""Time fixed launch shapes of the step kernels on the Taylor-Green vortex.

Each shape runs in its own context: initialise, one collision, two warm-up
steps, then `--passes` timed launches of the chosen kernel (HIP events on the
context's stream).  Used to profile one launch shape at a time (rocprofv3
counters are per dispatch), and to compare shapes outside the autotuner.

A shape is prec:mode:tile:zseg[:pad[:xcds]], where mode is wrap or tiled
(2x2x2 tiles with a 3-layer ghost fill through pack + unpack buffers), tile
the two-step column tile (-1: one pass per step), zseg the planes per sweep,
pad the elements between direction planes (-1 or absent: the library's
default) and xcds the tile order over the XCDs (8, the default: each XCD its
own range of columns; 1: dispatch order).  The key printed with each result
is the one bench.py reports as roofline.shape_key.

Usage: python -m imexlbm_amd.shape_probe --n 512 --shapes fp16:wrap:4:512,fp64:wrap:3:128
"""
from __future__ import annotations

import argparse
import sys

from . import capi
from .precision_study import tau0


def probe(n, prec, mode, tile, zseg, passes=8, re=1600, u0=0.05, pad=-1, xcds=0):
    dom = {"nx": n, "ny": n, "nz": n, "tau0": tau0(u0, n, re), "rho0": 1.0, "u0": u0}
    kw = {"two_step": 0} if tile < 0 else {"two_step": 1, "two_step_tile": tile, "two_step_zseg": zseg,
                                           "two_step_xcds": xcds}
    kw["dir_pad"] = pad
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
    if info["dir_pad"]:
        key += f":p{info['dir_pad']}"
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
        prec, mode, tile, zseg, *extra = s.split(":")
        pad = int(extra[0]) if extra else -1
        xcds = int(extra[1]) if len(extra) > 1 else 0
        key, ms = probe(a.n, prec, mode, int(tile), int(zseg), a.passes, pad=pad, xcds=xcds)
        print(f"{key} {ms:.4f} ms per launch", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
