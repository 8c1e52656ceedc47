"""GPU parity of the HIP path against the FP64 oracle, through the C ABI.

FP64 populations must equal the oracle's to <= 1e-13 relative per population
(they are expected to be bit-identical); conserved sums, which the library
accumulates in compensated FP64 in a different order than the reference,
must agree to 1e-12 relative (mass, KE) and 1e-12 absolute (momentum).
"""
import numpy as np
import pytest

from tests.conftest import ROOT, ke_matches_reference  # noqa: F401

pytestmark = pytest.mark.gpu

POP_TOL = 1e-13


def _dom(n, re=100, u0=0.05, rho0=1.0, ny=None, nz=None, oracle_mod=None):
    from oracle import oracle
    ny = ny or n
    nz = nz or n
    return {"nx": n, "ny": ny, "nz": nz, "tau0": oracle.tau(u0, n, re), "rho0": rho0, "u0": u0}


def _oracle(d, steps, nranks=1):
    from oracle import oracle
    o = oracle.Oracle(d["nx"], d["ny"], d["nz"], d["tau0"], d["rho0"], d["u0"], nranks)
    o.initialize()
    o.step(steps)
    return o


def _relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _ctx(d, **kw):
    from imexlbm_amd import capi
    print(f"case: {d} {kw}", flush=True)
    return capi.Context(d, **kw)


@pytest.mark.parametrize("shape,steps", [((16, 16, 16), 0), ((16, 16, 16), 1), ((32, 32, 32), 50),
                                         ((24, 20, 12), 17), ((7, 9, 5), 9)])
def test_fp64_wrap_matches_oracle(shape, steps):
    d = _dom(shape[0], ny=shape[1], nz=shape[2])
    with _ctx(d, precision="fp64") as c:
        c.initialize()
        c.step(steps)
        got = c.get_f()
        cons = c.conserved()
    o = _oracle(d, steps)
    want = o.get_f()
    assert _relmax(got, want) <= POP_TOL
    wc = o.conserved_compensated()
    assert abs(cons[0] - wc[0]) / wc[0] < 1e-12
    assert abs(cons[4] - wc[4]) / wc[4] < 1e-12
    assert np.abs(cons[1:4] - wc[1:4]).max() < 1e-12
    o.close()


def test_fp64_populations_bitwise_after_200_steps():
    # the 32^3 Re=100 deck, 200 steps
    d = _dom(32, re=100)
    with _ctx(d, precision="fp64") as c:
        c.initialize()
        c.step(200)
        got = c.get_f()
    o = _oracle(d, 200)
    want = o.get_f()
    o.close()
    assert np.array_equal(got, want)


def test_regression_deck_through_the_abi():
    # tests/regression.sh:41-43,60 and the KE transient of :28-30
    d = _dom(32, re=1600)
    cons = []
    with _ctx(d, precision="fp64") as c:
        c.initialize()
        for _ in range(20):
            c.collision()
            c.streaming()
            cons.append(c.conserved())
    cons = np.array(cons)
    assert ke_matches_reference(cons[:, 4])
    m0 = cons[0, 0]
    assert np.abs((cons[:, 0] - m0) / m0).max() < 1e-12
    assert np.abs(cons[:, 1:4]).max() < 1e-10


def test_streaming_alone_is_the_pull_permutation():
    from oracle import oracle
    _, e, _ = oracle.tables()
    d = _dom(8, ny=6, nz=5)
    tag = np.arange(27 * 8 * 6 * 5, dtype=np.float64).reshape(27, 8, 6, 5) + 0.25
    with _ctx(d, precision="fp64") as c:
        c.initialize()
        c.set_f(tag)
        c.streaming()           # no collision: a pure periodic pull
        once = c.get_f()
        c.streaming()
        twice = c.get_f()
    for ii in range(27):
        assert np.array_equal(once[ii], np.roll(tag[ii], tuple(e[ii]), axis=(0, 1, 2))), ii
        assert np.array_equal(twice[ii], np.roll(tag[ii], tuple(2 * e[ii]), axis=(0, 1, 2))), ii


def test_collision_alone_matches_oracle_on_random_state():
    from oracle import oracle
    w, _, _ = oracle.tables()
    d = _dom(6, ny=5, nz=4)
    rng = np.random.default_rng(3)
    f0 = w[:, None, None, None] * (1.0 + 0.3 * rng.standard_normal((27, 6, 5, 4)))
    with _ctx(d, precision="fp64") as c:
        c.initialize()
        c.set_f(f0)
        c.collision()
        got = c.get_f()         # forces the deferred collision, no stream
    o = oracle.Oracle(6, 5, 4, d["tau0"], d["rho0"], d["u0"])
    o.set_f(f0)
    o.collision()
    assert np.array_equal(got, o.get_f())
    o.close()


def test_interleaved_api_calls_keep_reference_semantics():
    # Collision/Streaming pairs, extra reads in between, and a bare Streaming
    # must all give what the oracle gives for the same call sequence.
    from oracle import oracle
    d = _dom(10, ny=8, nz=6)
    o = oracle.Oracle(10, 8, 6, d["tau0"], d["rho0"], d["u0"])
    o.initialize()
    with _ctx(d, precision="fp64") as c:
        c.initialize()
        seq = ["c", "s", "c", "s", "read", "c", "s", "s", "c", "read", "s", "c", "c", "s", "read"]
        for op in seq:
            if op == "c":
                c.collision(); o.collision()
            elif op == "s":
                c.streaming(); o.streaming()
            else:
                assert np.array_equal(c.get_f(), o.get_f())
    o.close()


@pytest.mark.parametrize("direct", [1, 0])
@pytest.mark.parametrize("ntiles,depth,n", [(8, 1, 16), (8, 3, 16), (2, 1, 17), (4, 2, 17),
                                            (12, 1, 18)])
def test_tiled_matches_wrap_bitwise(ntiles, depth, n, direct):
    # direct: boundary slabs copied straight into the neighbours' ghosts;
    # otherwise packed into the exchange buffer and unpacked from it
    from imexlbm_amd import capi
    d = _dom(n, re=400)
    with _ctx(d, precision="fp64") as w:
        w.initialize(); w.step(12)
        fw, cw = w.get_f(), w.conserved()
    with _ctx(d, precision="fp64", mode=capi.TILED, ntiles=ntiles, halo_depth=depth, halo_direct=direct) as t:
        t.initialize(); t.step(12)
        ft, ct = t.get_f(), t.conserved()
        info = t.info()
    assert info["ntiles"] == ntiles and info["halo_direct"] == direct and info["halo_depth"] == depth
    assert np.array_equal(fw, ft)
    assert abs(cw[0] - ct[0]) / cw[0] <= 1e-13 and abs(cw[4] - ct[4]) / cw[4] <= 1e-13


@pytest.mark.parametrize("direct", [1, 0])
@pytest.mark.parametrize("ntiles,depth,n,two", [(8, 1, 16, 0), (8, 3, 16, 0), (4, 2, 17, 0), (12, 1, 18, 0),
                                                (8, 3, 64, 1)])
def test_tiled_whole_x_matches_wrap_bitwise(ntiles, depth, n, two, direct):
    # tiles cut y and z only; x wraps inside the kernels, so no x ghosts are
    # filled.  Populations equal the periodic single-box run bit for bit.
    from imexlbm_amd import capi
    d = _dom(n, re=400)
    with _ctx(d, precision="fp64", two_step=0) as w:
        w.initialize(); w.step(12)
        fw, cw = w.get_f(), w.conserved()
    with _ctx(d, precision="fp64", mode=capi.TILED, ntiles=ntiles, halo_depth=depth, halo_direct=direct,
              whole_x=1, two_step=two) as t:
        t.initialize(); t.step(12)
        ft, ct = t.get_f(), t.conserved()
        info = t.info()
    assert info["ntiles"] == ntiles and info["whole_x"] == 1 and info["two_step"] == two
    assert info["rx"] == 1 and info["ry"] * info["rz"] == ntiles
    assert np.array_equal(fw, ft)
    assert abs(cw[0] - ct[0]) / cw[0] <= 1e-13 and abs(cw[4] - ct[4]) / cw[4] <= 1e-13


def test_tiled_whole_x_geometry():
    # setup_Cartesian's rule over y and z with rx = 1: for a cube the first
    # strict minimum of the spread puts the larger factor on z
    from imexlbm_amd import capi
    d = _dom(32)
    for n, want in ((2, (1, 1, 2)), (4, (1, 2, 2)), (8, (1, 2, 4))):
        with _ctx(d, mode=capi.TILED, ntiles=n, whole_x=1) as t:
            i = t.info()
            assert (i["rx"], i["ry"], i["rz"]) == want


def test_tiled_geometry_follows_setup_cartesian():
    from imexlbm_amd import capi
    from oracle import oracle
    d = _dom(32)
    for n in (2, 4, 8):
        with _ctx(d, mode=capi.TILED, ntiles=n) as t:
            i = t.info()
            assert (i["rx"], i["ry"], i["rz"]) == oracle.decompose(n, 32, 32, 32)


def test_graph_replay_matches_direct_launches():
    d = _dom(16, re=200)
    with _ctx(d, precision="fp64") as a, _ctx(d, precision="fp64", use_graph=1) as b:
        a.initialize(); b.initialize()
        for k in range(4):            # several replays with eager reads in between
            a.step(6); b.step(6)
            assert np.array_equal(a.conserved(), b.conserved())
            assert np.array_equal(a.get_f(), b.get_f())
        a.step(5); b.step(5)          # odd count: graph pairs plus one direct step
        assert np.array_equal(a.get_f(), b.get_f())


def test_autotune_leaves_state_and_results_unchanged():
    d = _dom(24, re=300)
    with _ctx(d, precision="fp64") as a:
        a.initialize(); a.step(3)
        before = a.get_f()
        best = a.autotune(reps=2)
        assert best > 0
        assert np.array_equal(a.get_f(), before)
        a.step(7)
        got = a.get_f()
    o = _oracle(d, 10)
    assert np.array_equal(got, o.get_f())
    o.close()


@pytest.mark.parametrize("cfg", [dict(block=64, cells_per_thread=3), dict(block=512, persistent=7),
                                 dict(block=128, nontemporal=1, dir_pad=37)])
def test_launch_shapes_do_not_change_results(cfg):
    d = _dom(20, ny=12, nz=9)
    with _ctx(d, precision="fp64", **cfg) as c:
        c.initialize(); c.step(9)
        got = c.get_f()
    o = _oracle(d, 9)
    assert np.array_equal(got, o.get_f())
    o.close()


def test_batched_instances_each_match_their_oracle():
    from oracle import oracle
    doms = [
        {"nx": 12, "ny": 12, "nz": 12, "tau0": oracle.tau(0.05, 12, 100), "rho0": 1.0, "u0": 0.05},
        {"nx": 16, "ny": 10, "nz": 8, "tau0": 0.0123, "rho0": 1.1, "u0": 0.03},
        {"nx": 9, "ny": 9, "nz": 9, "tau0": oracle.tau(0.08, 9, 250), "rho0": 0.9, "u0": 0.08},
    ]
    for seq in (0, 1):
        with _ctx(doms, precision="fp64", sequential=seq) as c:
            c.initialize(); c.step(11)
            for k, d in enumerate(doms):
                o = _oracle(d, 11)
                assert np.array_equal(c.get_f(k), o.get_f()), (seq, k)
                wc = o.conserved_compensated()
                assert abs(c.conserved(k)[0] - wc[0]) / wc[0] < 1e-12
                o.close()


@pytest.mark.parametrize("prec,pop_tol", [("fp32", 2e-6), ("fp16", 2e-4), ("bf16", 2e-3)])
def test_reduced_precision_tracks_fp64(prec, pop_tol):
    d = _dom(32, re=400)
    with _ctx(d, precision="fp64") as a:
        a.initialize(); a.step(40)
        f64, c64 = a.get_f(), a.conserved()
    with _ctx(d, precision=prec) as b:
        b.initialize()
        c0 = b.conserved()
        b.step(40)
        fr, cr = b.get_f(), b.conserved()
    # absolute error per population relative to the largest population
    assert np.abs(fr - f64).max() / np.abs(f64).max() < pop_tol
    # mass is conserved up to the storage rounding of the shifted populations
    assert abs(cr[0] - c0[0]) / c0[0] < 50 * pop_tol
    assert abs(cr[4] - c64[4]) / c64[4] < 200 * pop_tol


def test_large_domain_invariants_fp64():
    # 192^3 FP64, 30 steps: size-independent checks at a large size.
    d = _dom(192, re=1600)
    with _ctx(d, precision="fp64") as c:
        c.initialize()
        m0 = c.conserved()
        c.step(30)
        m1 = c.conserved()
    assert abs(m1[0] - m0[0]) / m0[0] < 1e-12
    assert np.abs(m1[1:4]).max() < 1e-10
    assert 0.9 * m0[4] < m1[4] < 1.01 * m0[4]


def test_get_macro_matches_moments_of_oracle_state():
    from oracle import oracle
    _, e, _ = oracle.tables()
    d = _dom(12, ny=10, nz=8)
    with _ctx(d, precision="fp64") as c:
        c.initialize(); c.step(5)
        p, u, v, w = c.get_macro()
    o = _oracle(d, 5)
    f = o.get_f()
    o.close()
    mom = np.einsum("ia,ixyz->axyz", e.astype(float), f)
    assert np.abs(p - f.sum(0) / 3.0).max() < 1e-15
    for got, want in zip((u, v, w), mom):
        assert np.abs(got - want).max() < 1e-15


@pytest.mark.parametrize("direct,two,whole_x", [(1, -1, 0), (0, 0, 0), (1, -1, 1)])
def test_tiled_256_three_layer_halo_matches_wrap(direct, two, whole_x):
    # the 256^3 deck's decomposition check: 2x2x2 tiles (1x2x4 keeping x
    # whole), 3-layer ghost fill (direct with two steps per pass, or
    # pack/unpack one step per pass)
    from imexlbm_amd import capi
    d = _dom(256, re=1600)
    with _ctx(d, precision="fp64") as w:
        w.initialize(); w.step(6)
        cw = w.conserved()
        fw = w.get_f()
    with _ctx(d, precision="fp64", mode=capi.TILED, ntiles=8, halo_depth=3, halo_direct=direct,
              two_step=two, whole_x=whole_x) as t:
        t.initialize(); t.step(6)
        ct = t.conserved()
        info = t.info()
        assert info["ntiles"] == 8 and info["two_step"] == (1 if two else 0)
        assert (info["rx"], info["ry"], info["rz"]) == ((1, 2, 4) if whole_x else (2, 2, 2))
        assert np.array_equal(fw, t.get_f())
    del fw
    assert abs(cw[0] - ct[0]) / cw[0] <= 1e-13 and abs(cw[4] - ct[4]) / cw[4] <= 1e-13


@pytest.mark.parametrize("whole_x", [1, 0])
@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_tiled_256_16bit_whole_x_matches_wrap(prec, whole_x):
    # the tiled benchmark's shapes: 1x2x4 tiles keeping x whole, or the
    # reference's 2x2x2 with the pass writing its x neighbours' ghosts; two
    # steps per pass in 64x16 columns (FP16 with the partial prefetch and its
    # mixed-precision collision), against the periodic box one step per
    # pass, bit for bit
    from imexlbm_amd import capi
    d = _dom(256, re=1600)
    pf = 2 if prec == "fp16" else 0
    with _ctx(d, precision=prec, two_step=0) as w:
        w.initialize(); w.step(6)
        cw = w.conserved()
        fw = w.get_f()
    with _ctx(d, precision=prec, mode=capi.TILED, ntiles=8, halo_depth=3, whole_x=whole_x, two_step=1,
              two_step_tile=4, two_step_prefetch=pf) as t:
        t.initialize(); t.step(6)
        ct = t.conserved()
        info = t.info()
        assert (info["rx"], info["ry"], info["rz"]) == ((1, 2, 4) if whole_x else (2, 2, 2))
        assert info["two_step_tile"] == 4 and info["two_step_prefetch"] == pf
        assert np.array_equal(fw, t.get_f())
    del fw
    assert abs(cw[0] - ct[0]) / cw[0] <= 1e-13 and abs(cw[4] - ct[4]) / cw[4] <= 1e-13


def test_precision_study_small():
    from imexlbm_amd import precision_study
    recs = precision_study.study(24, 40, 10, 400, 0.05, ["fp32", "fp16", "bf16"])
    by = {r["prec"]: r for r in recs}
    assert by["fp32"]["stable"] and by["fp32"]["vel_l2"] < 1e-5
    assert by["fp16"]["stable"] and by["fp16"]["vel_l2"] < 5e-3
    assert by["bf16"]["stable"] and by["bf16"]["vel_l2"] < 5e-2
    assert by["fp64"]["mass_drift"] < 1e-13


def test_batch_study_small():
    from imexlbm_amd import batch_study
    r = batch_study.run(count=5, steps=20, prec="fp32")
    assert r["vel_l2_max"] < 1e-5 and r["ke_err_max"] < 1e-5
    assert r["batched_mlups"] > 0 and r["one_after_another_mlups"] > 0


def test_fp32_uniform_flow_keeps_its_momentum():
    # A uniform flow is a fixed point of stream + collide.  With float-rounded
    # weights the equilibrium's momentum is u(1+d), d ~ 3e-8, and the flow
    # would gain ~6e-5 over 1000 steps; with the exact momentum weights only
    # unbiased rounding remains.
    from oracle import oracle
    w, e, _ = oracle.tables()
    n = 16
    u = np.array([0.05, -0.03, 0.02])
    eu = e.astype(float) @ u
    feq = w * (1.0 + 3.0 * eu + 4.5 * eu ** 2 - 1.5 * (u @ u))
    f0 = np.broadcast_to(feq[:, None, None, None], (27, n, n, n)).copy()
    d = {"nx": n, "ny": n, "nz": n, "tau0": 0.002, "rho0": 1.0, "u0": 0.05}
    with _ctx(d, precision="fp32") as c:
        c.initialize()
        c.set_f(f0)
        p0 = c.conserved()[1:4]
        c.step(1000)
        p1 = c.conserved()[1:4]
    assert np.abs(p1 - p0).max() / np.abs(p0).max() < 2e-6


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16", "bf16"])
def test_paired_cells_match_single_cells_bitwise(prec):
    # cells_per_thread >= 2 selects the kernel that loads two cells per thread
    d = _dom(24, ny=18, nz=7, re=300)
    out = []
    for cpt, blk in ((1, 256), (2, 256), (3, 64)):
        with _ctx(d, precision=prec, cells_per_thread=cpt, block=blk) as c:
            c.initialize(); c.step(11)
            out.append(c.get_f())
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


@pytest.mark.parametrize("prec", ["fp32", "fp16", "bf16"])
def test_tile_order_and_access_path_do_not_change_results(prec):
    d = _dom(20, ny=16, nz=6, re=300)
    out = []
    for cfg in ({}, {"io16": 0}, {"io16": 1}, {"xcd_groups": 8, "xcd_run": 3},
                {"xcd_groups": 8, "xcd_run": 0}, {"xcd_groups": 1}):
        with _ctx(d, precision=prec, **cfg) as c:
            c.initialize(); c.step(9)
            out.append(c.get_f())
    for o in out[1:]:
        assert np.array_equal(out[0], o)
