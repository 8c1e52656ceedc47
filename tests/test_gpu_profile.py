"""The per-phase profiler's byte and flop model against an independent count.

The reference's profiler (src/profile.cpp:38-90) charges collision and stream
2 x cells x 27 x 8 bytes each, pack and unpack twice the bytes of the 26 send
or receive buffers, computeMacroscopic the populations plus four doubles per
cell, and 594 flop per collided cell, 189 per cell of moments.  Here the same
model holds with sizeof(storage) for the populations; the fused launch is
charged once per launch (each population read and written once).  The
exchange-buffer size is counted here from the tile geometry: one ghost layer,
the directions crossing each channel (9 per face, 3 per edge, 1 per corner).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dom(n, re=400, u0=0.05):
    from oracle import oracle
    return {"nx": n, "ny": n, "nz": n, "tau0": oracle.tau(u0, n, re), "rho0": 1.0, "u0": u0}


@pytest.mark.parametrize("prec,size", [("fp64", 8), ("fp32", 4), ("fp16", 2)])
def test_profile_groups_follow_the_reference_byte_and_flop_model(prec, size):
    from imexlbm_amd import capi
    n, steps = 32, 5
    cells = n ** 3
    print(f"case: profile model {prec} {n}^3 2x2x2 tiles", flush=True)
    with capi.Context(_dom(n), precision=prec, mode=capi.TILED, ntiles=8, halo_depth=1, halo_direct=0,
                      two_step=0, profile=1) as c:
        c.initialize()
        c.step(steps)
        c.get_macro()
        c.conserved()
        prof = c.profile()
        info = c.info()
    # 2x2x2 tiles of 16^3: per tile 6 faces of 16^2 cells x 9 directions,
    # 12 edges of 16 cells x 3, 8 corners x 1
    l = n // 2
    send = 8 * (6 * l * l * 9 + 12 * l * 3 + 8 * 1) * size
    assert info["halo_send_bytes"] == send
    fb = cells * 27 * size

    def per_call(name):
        calls, ms, mn, mx, by, fl = prof[name]
        assert calls > 0 and ms > 0 and 0 < mn <= mx
        return by / calls, fl / calls

    assert per_call("collision") == (2.0 * fb, 594.0 * cells)
    assert prof["collision"][0] == 1                       # the first step collides in place
    assert per_call("collision+stream (fused)") == (2.0 * fb, 594.0 * cells)
    assert prof["collision+stream (fused)"][0] == steps - 1
    assert per_call("pack (enq+fence)") == (2.0 * send, 0.0)
    assert per_call("unpack (enq+fence)") == (2.0 * send, 0.0)
    assert prof["pack (enq+fence)"][0] == prof["unpack (enq+fence)"][0] >= steps - 1
    assert per_call("computeMacroscopic") == (fb + 4.0 * 8 * cells, 189.0 * cells)
    assert prof["mpi (post+wait)"][0] == 0                  # no exchange step on one GPU
    assert prof["stream (enq+fence)"][0] == 0
    assert prof["ghost fill (direct)"][0] == 0


def test_profile_two_step_pass_is_charged_the_bytes_it_moves():
    from imexlbm_amd import capi
    print("case: profile two-step fp16 128x128x32", flush=True)
    d = {"nx": 128, "ny": 128, "nz": 32, "tau0": 0.01, "rho0": 1.0, "u0": 0.05}
    with capi.Context(d, precision="fp16", two_step=1, profile=1) as c:
        c.initialize()
        c.step(1)
        c.profile_reset()
        c.step(8)
        prof = c.profile()
        assert c.info()["two_step"] == 1
    calls, ms, mn, mx, by, fl = prof["collision+stream (fused)"]
    cells = 128 * 128 * 32
    assert calls == 4
    assert by == 4 * 2.0 * 27 * 2 * cells          # one read and one write per population and pass
    assert fl == 4 * 2 * 594.0 * cells             # two collisions per cell and pass
    assert np.isfinite(ms) and ms > 0
