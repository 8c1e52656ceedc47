"""The benchmark workloads at full size, through the C ABI.

Each test runs one of the measured configurations at its real size and step
count and checks it against its FP64 run (which is bit-identical to the
oracle at the sizes the oracle can run) and against the invariants the
reference's regression test asserts (tests/regression.sh:41-43):

  * 256^3 FP64, Re=1600, 2000 steps: periodic wrap against 2x2x2 subdomain
    tiles with a 3-layer ghost fill through pack + unpack buffers, every
    report and the final populations;
  * 512^3, Re=1600, 2000 steps: FP32 and FP16-stored populations against FP64
    (velocity L2, KE curve, mass drift);
  * 640^3, Re=1600, 2000 steps, u0 in {0.02, 0.05, 0.08}: FP16 and BF16;
  * 64 seeded batched FP32 instances, 1000 steps, against the FP64 batch, and
    the batch's FP64 instances of the two smallest edges against the oracle.

Accuracy bounds.  Velocity L2 = ||u - u64|| / ||u64|| over the final field;
KE error = max over the reports (every 100 steps) of |KE - KE64| / KE64.
FP16-stored populations must meet velocity L2 <= 1e-3 and KE error <= 0.5%.
BF16 (8-bit mantissa) does not meet them: its bounds below are the recorded
behaviour (velocity L2 about 1e-2, KE about 1%), asserted so that a
regression shows.  FP32 is held to 1e-4.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VEL_L2_FP16 = 1e-3
KE_ERR_FP16 = 5e-3
VEL_L2_BF16 = 2.5e-2     # recorded: 7e-3 .. 1.4e-2
KE_ERR_BF16 = 2.5e-2     # recorded: 3e-3 .. 1.4e-2
VEL_L2_FP32 = 1e-4       # recorded: 9e-6
KE_ERR_FP32 = 1e-4       # recorded: 2e-5
MASS_DRIFT_FP64 = 1e-12  # tests/regression.sh:41
MOMENTUM_ABS = 1e-10     # tests/regression.sh:42


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _tau0(u0, n, re):
    from imexlbm_amd.precision_study import tau0
    return tau0(u0, n, re)


def test_config_256_fp64_wrap_equals_2x2x2_pack_unpack_over_2000_steps():
    from imexlbm_amd import capi
    n, re, u0, steps, every = 256, 1600, 0.05, 2000, 100
    dom = {"nx": n, "ny": n, "nz": n, "tau0": _tau0(u0, n, re), "rho0": 1.0, "u0": u0}
    _log(f"case: 256^3 fp64 wrap vs 2x2x2 pack/unpack, {steps} steps")
    wrap = capi.Context(dom, precision="fp64")
    tiled = capi.Context(dom, precision="fp64", mode=capi.TILED, ntiles=8, halo_depth=3, halo_direct=0)
    try:
        info = tiled.info()
        assert (info["rx"], info["ry"], info["rz"]) == (2, 2, 2)
        assert info["halo_direct"] == 0 and info["halo_depth"] >= 2
        wrap.initialize()
        tiled.initialize()
        m0 = wrap.conserved()[0]
        done = 0
        while done < steps:
            wrap.step(every)
            tiled.step(every)
            done += every
            cw, ct = wrap.conserved(), tiled.conserved()
            _log(f"  step {done}: mass {cw[0]:.15e} ke {cw[4]:.15e} | tiled ke {ct[4]:.15e}")
            assert np.all(np.isfinite(cw)) and np.all(np.isfinite(ct))
            # decomposition independence (tests/regression.sh:108-126)
            assert abs(cw[0] - ct[0]) / abs(cw[0]) <= 1e-13
            assert abs(cw[4] - ct[4]) / abs(cw[4]) <= 1e-13
            assert np.abs(cw[1:4] - ct[1:4]).max() <= 1e-13
            # conservation (tests/regression.sh:91-104)
            assert abs(cw[0] - m0) / abs(m0) <= MASS_DRIFT_FP64
            assert np.abs(cw[1:4]).max() <= MOMENTUM_ABS
        fw = wrap.get_f()
        ft = tiled.get_f()
        assert np.array_equal(fw, ft), "tiled populations differ from the periodic box"
    finally:
        wrap.close()
        tiled.close()


def test_config_512_fp32_and_fp16_against_fp64_over_2000_steps():
    from imexlbm_amd import precision_study
    _log("case: 512^3 Re=1600 u0=0.05 fp32+fp16 vs fp64, 2000 steps")
    recs = precision_study.study(512, 2000, 100, 1600, 0.05, ["fp32", "fp16"])
    by = {r["prec"]: r for r in recs}
    for p in ("fp32", "fp16"):
        _log(f"  {p}: vel_l2 {by[p].get('vel_l2')} ke_err {by[p].get('ke_err')} "
             f"mass_drift {by[p].get('mass_drift')}")
    assert by["fp64"]["mass_drift"] <= MASS_DRIFT_FP64
    assert all(math.isfinite(k) for k in by["fp64"]["ke_curve"])
    assert by["fp32"]["stable"]
    assert by["fp32"]["vel_l2"] <= VEL_L2_FP32 and by["fp32"]["ke_err"] <= KE_ERR_FP32
    assert by["fp32"]["mass_drift"] <= 1e-6
    # the headline precision: the targets over 2000 steps at 512^3
    assert by["fp16"]["stable"]
    assert by["fp16"]["vel_l2"] <= VEL_L2_FP16
    assert by["fp16"]["ke_err"] <= KE_ERR_FP16
    assert by["fp16"]["mass_drift"] <= 1e-5


@pytest.mark.parametrize("u0", [0.02, 0.05, 0.08])
def test_config_640_fp16_bf16_against_fp64_over_2000_steps(u0):
    from imexlbm_amd import precision_study
    _log(f"case: 640^3 Re=1600 u0={u0} fp16+bf16 vs fp64, 2000 steps")
    recs = precision_study.study(640, 2000, 100, 1600, u0, ["fp16", "bf16"])
    by = {r["prec"]: r for r in recs}
    for p in ("fp16", "bf16"):
        _log(f"  {p}: vel_l2 {by[p].get('vel_l2')} ke_err {by[p].get('ke_err')} "
             f"mass_drift {by[p].get('mass_drift')}")
    assert by["fp64"]["mass_drift"] <= MASS_DRIFT_FP64
    assert by["fp16"]["stable"] and by["bf16"]["stable"]
    assert by["fp16"]["vel_l2"] <= VEL_L2_FP16 and by["fp16"]["ke_err"] <= KE_ERR_FP16
    assert by["fp16"]["mass_drift"] <= 1e-5
    # BF16 misses the FP16 targets; these are its recorded bounds
    assert by["bf16"]["vel_l2"] <= VEL_L2_BF16 and by["bf16"]["ke_err"] <= KE_ERR_BF16
    assert by["bf16"]["mass_drift"] <= 1e-4


def test_config_64_batched_fp32_instances_over_1000_steps(oracle_mod):
    from imexlbm_amd import batch_study, capi
    count, steps = 64, 1000
    _log(f"case: {count} batched fp32 instances, {steps} steps")
    r = batch_study.run(count, steps, "fp32")
    _log(f"  batched {r['batched_mlups']:.0f} MLUPS, one after another "
         f"{r['one_after_another_mlups']:.0f}, vel_l2 max {r['vel_l2_max']:.2e}, "
         f"ke_err max {r['ke_err_max']:.2e}, unstable fp64 decks {r['unstable_fp64']}")
    assert r["instances"] == count and len(r["per_instance"]) == count
    # no instance is unstable in FP32 where its FP64 run is stable
    assert r["unstable_only_reduced"] == 0
    for inst in r["per_instance"]:
        if inst["fp64_stable"]:
            assert inst["vel_l2"] <= VEL_L2_FP32, inst
            assert inst["ke_err"] <= KE_ERR_FP32, inst
    assert r["batched_mlups"] > r["one_after_another_mlups"]

    # the FP64 batch, instance by instance, against the oracle: the first
    # instance of each of the two smallest edges, after the same step count
    doms = batch_study.instances(count)
    picks = []
    for edge in (32, 48):
        k = next((i for i, d in enumerate(doms)
                  if d["nx"] == edge and r["per_instance"][i]["fp64_stable"]), None)
        if k is not None:
            picks.append(k)
    assert picks
    with capi.Context(doms, precision="fp64", use_graph=1) as c:
        c.initialize()
        c.step(steps + 3)
        got = {k: c.get_f(k) for k in picks}
    for k in picks:
        d = doms[k]
        _log(f"  oracle check: instance {k} n={d['nx']} re={d['re']:.1f} u0={d['u0']:.4f}")
        with oracle_mod.Oracle(d["nx"], d["ny"], d["nz"], d["tau0"], d["rho0"], d["u0"]) as o:
            o.initialize()
            o.step(steps + 3)
            want = o.get_f()
        assert np.array_equal(got[k], want), f"instance {k} differs from the oracle"
