"""The C++ driver (imexlbm_amd/bin/imexlbm) run the way the reference's
tests/regression.sh runs its binary: the 32^3, Re=1600, 20-step deck
reporting every step, the same three assertions (mass drift, momentum,
decomposition independence), here with decompositions of the domain into
1, 2, 4 and 8 device tiles in place of 1, 2 and 4 MPI ranks."""
import os
import re
import subprocess

import pytest

from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "imexlbm_amd", "bin", "imexlbm")
DECK = "1.0  0.1  1600\n0.05  20  1\n32  32  32\n"     # tests/regression.sh:60
TOL_MASS_DRIFT, TOL_MOMENTUM, TOL_RANK_AGREE = 1e-12, 1e-10, 1e-12   # :41-43

CONS = re.compile(r"^cons\s+step\s+(\d+) \| mass (\S+) \| px (\S+) \| py (\S+) \| pz (\S+) \| ke (\S+)$")
STEP = re.compile(r"^step\s+(\d+) \| interval\s+\S+ s \| total\s+\S+ s \|\s+(\S+) MLUPS$")


def _run(tmp_path, deck=DECK, **env):
    (tmp_path / "input.in").write_text(deck)
    e = dict(os.environ)
    e.update({k: str(v) for k, v in env.items()})
    print(f"case: driver {env}", flush=True)
    r = subprocess.run([EXE], cwd=tmp_path, env=e, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    cons = [tuple(float(x) for x in m.groups()) for m in map(CONS.match, r.stdout.splitlines()) if m]
    steps = [m for m in map(STEP.match, r.stdout.splitlines()) if m]
    assert cons, r.stdout[-2000:]
    assert len(steps) == len(cons)
    for row in cons:
        assert all(map(lambda v: v == v and abs(v) != float("inf"), row))
    return cons, r.stdout


def _invariants(cons):
    m0 = cons[0][1]
    assert max(abs(c[1] - m0) / m0 for c in cons) <= TOL_MASS_DRIFT
    assert max(abs(v) for c in cons for v in c[2:5]) <= TOL_MOMENTUM


def test_driver_regression_deck_and_decomposition_independence(tmp_path):
    ref, out = _run(tmp_path)
    assert "3D Taylor-Green Vortex" in out and "Re    = 1600" in out
    _invariants(ref)
    assert [float(f"{c[5]:.4g}") for c in ref[:3]] == [9.852, 10.22, 9.889]
    for n in (1, 2, 4, 8):
        got, out = _run(tmp_path, IMEXLBM_MODE="tiled", IMEXLBM_TILES=n)
        assert "rx=" in out
        _invariants(got)
        for a, b in zip(ref, got):
            assert abs(b[1] - a[1]) / a[1] <= TOL_RANK_AGREE
            assert abs(b[5] - a[5]) / a[5] <= TOL_RANK_AGREE


@pytest.mark.parametrize("prec", ["fp32", "fp16", "bf16"])
def test_driver_reduced_precision_runs(tmp_path, prec):
    ref, _ = _run(tmp_path)
    got, _ = _run(tmp_path, IMEXLBM_PRECISION=prec)
    assert abs(got[-1][5] - ref[-1][5]) / ref[-1][5] < 2e-2
    assert abs(got[-1][1] - ref[-1][1]) / ref[-1][1] < 1e-4


def test_driver_autotune_matches_plain(tmp_path):
    ref, _ = _run(tmp_path)
    got, _ = _run(tmp_path, IMEXLBM_AUTOTUNE=1)
    # populations are bit-identical; the compensated sums may round differently
    # when the tuned tile size changes their order
    for a, b in zip(ref, got):
        assert abs(b[1] - a[1]) / a[1] <= 1e-14 and abs(b[5] - a[5]) / a[5] <= 1e-14
