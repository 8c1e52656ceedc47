"""bench.py's roofline bookkeeping, on the CPU.

The roofline fraction of the step kernel counts the bytes one launch must
move (every population read and written once) over the launch's duration,
against the 8 TB/s peak, whether the launch performs one lattice update or
two; it cannot pass 1 for a launch slower than the peak allows.  The HBM
traffic beside it is looked up, by launch shape, in the committed rocprofv3
summaries, so those must hold the shapes the benchmark runs.
"""
import importlib.util
import os

import pytest

from tests.conftest import ROOT


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _info(two_step, tile=4, zseg=512, xcds=8, pad=0, block=256, cpt=2):
    return {"two_step": two_step, "two_step_tile": tile, "two_step_zseg": zseg, "two_step_xcds": xcds,
            "block": block, "cells_per_thread": cpt, "dir_pad": pad}


def test_shape_keys():
    b = _bench()
    assert b.shape_key("fp16", 512, _info(1)) == "fp16:512:wrap:ts4:z512:x8"
    assert b.shape_key("fp64", 512, _info(1, tile=3, pad=32)) == "fp64:512:wrap:ts3:z512:x8:p32"
    assert b.shape_key("fp64", 256, _info(1, tile=3, zseg=128), mode="tiled") == "fp64:256:tiled:ts3:z128:x8"
    assert b.shape_key("fp16", 512, _info(0)) == "fp16:512:wrap:step:b256:c2"


# launch shapes the default benchmark has chosen on MI355X, and the bytes one
# launch must move (2 * 27 * sizeof * cells)
PROFILED = [
    ("fp16:512:wrap:ts4:z512:x8", 512, 2 * 27 * 2 * 512 ** 3),
    ("bf16:512:wrap:ts4:z512:x8", 512, 2 * 27 * 2 * 512 ** 3),
    ("fp32:512:wrap:ts0:z128:x8:p32", 512, 2 * 27 * 4 * 512 ** 3),
    ("fp32:512:wrap:ts0:z512:x8:p32", 512, 2 * 27 * 4 * 512 ** 3),
    ("fp64:512:wrap:ts3:z512:x8:p32", 512, 2 * 27 * 8 * 512 ** 3),
    ("fp64:256:wrap:ts3:z256:x8:p32", 256, 2 * 27 * 8 * 256 ** 3),
    ("fp64:256:tiled:ts3:z128:x8", 256, 2 * 27 * 8 * 256 ** 3),
]


@pytest.mark.parametrize("key,n,launch_bytes", PROFILED)
def test_committed_traffic_covers_the_benchmarked_shapes(key, n, launch_bytes):
    b = _bench()
    t = b.traffic_from_profiles(key, n)
    assert t is not None, f"no traffic summary for {key}"
    # a launch moves at least what it must (rounding of the counters aside),
    # and a two-step pass re-reads its ring mostly from L2
    assert 0.99 * launch_bytes <= t <= 1.5 * launch_bytes


@pytest.mark.parametrize("ups", [1, 2])
def test_roofline_fraction_counts_the_bytes_of_one_launch(ups):
    b = _bench()
    n, bpu = 512, 2 * 27 * 2
    cells = n ** 3
    launch_bytes = bpu * cells
    launch_ms = 3.15                        # FP16 two-step pass at 512^3
    gbs = launch_bytes / (launch_ms / 1e3) / 1e9
    mlups = cells * ups / (launch_ms / 1e3) / 1e6
    rec = {"kernel_gbs": round(gbs, 1), "kernel_frac": round(gbs / b.HBM_PEAK_GBS, 4), "launch_bytes": launch_bytes,
           "launch_ms": launch_ms, "updates_per_launch": ups,
           "mlups_over_roofline": round(mlups / (b.HBM_PEAK_GBS * 1e9 / bpu / 1e6), 4),
           "shape_key": "fp16:512:wrap:ts4:z512:x8"}
    r = b.roofline(rec, n, copy_gbs=6300.0)
    # the same fraction whether the launch performs one update or two
    assert r["frac"] == pytest.approx(launch_bytes / (launch_ms / 1e3) / 8e12, abs=1e-4)
    assert r["frac"] < 1.0
    assert r["achieved"] / r["peak"] == pytest.approx(r["frac"], abs=1e-4)
    # the per-update form may pass 1 with two updates per launch
    assert r["mlups_over_roofline"] == pytest.approx(ups * r["frac"], abs=1e-3)
    assert r["traffic"] is not None
    assert r["traffic_over_launch_bytes"] == pytest.approx(r["traffic"] / launch_bytes, abs=1e-4)
