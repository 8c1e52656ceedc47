"""Accuracy of the reduced storage formats against the FP64 path.

Runs the periodic Taylor-Green vortex in FP64 (the parity path, bit-identical
to the CPU oracle at small sizes) and in one or more reduced formats side by
side on one GPU, and reports for each reduced format:

  vel_l2      ||u - u64||_2 / ||u64||_2 over the whole field at the end
  ke_err      max over the reports of |KE - KE64| / KE64
  eps_peak    peak dissipation rate -dKE/dt (per lattice step) and its error
  mass_drift  max over the reports of |m(t) - m(0)| / m(0)
  stable      every reported KE finite and the final field finite

Usage: python -m imexlbm_amd.precision_study --n 512 --steps 2000 --every 100 \
           --prec fp32 --prec fp16 --u0 0.05
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import capi

PI_LIT = 3.1415926


def tau0(u0, n, re):
    # System::System, src/System.cpp:28 (Re may be non-integer here)
    return u0 * n / re / (1.0 / 3.0) / 2.0 / PI_LIT


def _vel_l2(ref, got):
    num = 0.0
    den = 0.0
    for a, b in zip(ref[1:], got[1:]):
        num += float(np.sum((b - a) ** 2))
        den += float(np.sum(a ** 2))
    return math.sqrt(num / den) if den > 0 else float("nan")


def _peak_eps(ts, ke):
    eps = [-(ke[k + 1] - ke[k]) / (ts[k + 1] - ts[k]) for k in range(len(ts) - 1)]
    k = int(np.nanargmax(eps))
    return eps[k], 0.5 * (ts[k] + ts[k + 1])


def study(n, steps, every, re, u0, precs, log=sys.stderr):
    dom = {"nx": n, "ny": n, "nz": n, "tau0": tau0(u0, n, re), "rho0": 1.0, "u0": u0}
    ctxs = {"fp64": capi.Context(dom, precision="fp64")}
    for p in precs:
        ctxs[p] = capi.Context(dom, precision=p)
    try:
        for c in ctxs.values():
            c.initialize()
        ts = [0]
        cons = {p: [c.conserved()] for p, c in ctxs.items()}
        t0 = time.time()
        done = 0
        while done < steps:
            k = min(every, steps - done)
            for c in ctxs.values():
                c.step(k)
            done += k
            ts.append(done)
            for p, c in ctxs.items():
                cons[p].append(c.conserved())
            print(f"study n={n} u0={u0} step {done}/{steps} "
                  + " ".join(f"{p}:ke={cons[p][-1][4]:.6e}" for p in ctxs) + f" ({time.time() - t0:.0f}s)",
                  file=log, flush=True)
        ref_macro = ctxs["fp64"].get_macro()
        ke64 = np.array([c[4] for c in cons["fp64"]])
        eps64, t_eps64 = _peak_eps(ts, ke64)
        out = []
        for p in precs:
            ke = np.array([c[4] for c in cons[p]])
            mass = np.array([c[0] for c in cons[p]])
            stable = bool(np.all(np.isfinite(ke)))
            rec = {"prec": p, "n": n, "re": re, "u0": u0, "steps": steps, "every": every,
                   "stable": stable}
            if stable:
                got = ctxs[p].get_macro()
                stable = all(bool(np.all(np.isfinite(a))) for a in got)
                rec["stable"] = stable
                if stable:
                    eps, t_eps = _peak_eps(ts, ke)
                    rec.update({
                        "vel_l2": _vel_l2(ref_macro, got),
                        "ke_err": float(np.max(np.abs(ke - ke64) / ke64)),
                        "eps_peak64": eps64, "eps_peak": eps, "t_peak64": t_eps64, "t_peak": t_eps,
                        "eps_peak_err": abs(eps - eps64) / abs(eps64),
                        "mass_drift": float(np.max(np.abs(mass - mass[0]) / mass[0])),
                        "ke_rel_curve": [float(v) for v in (ke - ke64) / ke64],
                    })
                del got
            out.append(rec)
        m64 = np.array([c[0] for c in cons["fp64"]])
        out.append({"prec": "fp64", "n": n, "re": re, "u0": u0, "steps": steps,
                    "mass_drift": float(np.max(np.abs(m64 - m64[0]) / m64[0])),
                    "eps_peak": eps64, "t_peak": t_eps64, "ke_final": float(ke64[-1]),
                    "ke_curve": [float(v) for v in ke64]})
        return out
    finally:
        for c in ctxs.values():
            c.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--re", type=float, default=1600)
    ap.add_argument("--u0", type=float, action="append")
    ap.add_argument("--prec", action="append")
    a = ap.parse_args(argv)
    for u0 in a.u0 or [0.05]:
        for rec in study(a.n, a.steps, a.every, a.re, u0, a.prec or ["fp32"]):
            print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
