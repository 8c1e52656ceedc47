"""Two steps per pass (temporal blocking) against one pass per step.

The two-step kernel keeps the intermediate step in LDS rounded to the storage
format, so its populations must be bit-identical to two single-step launches
in every format, for every column tile, z-segmentation and tile order, and
through every entry point that advances the state (imex_step, the deferred
Collision/Streaming pairs, graph replay, the autotuner).  FP64 is also checked
directly against the oracle.
"""
import numpy as np
import pytest

from tests.conftest import ROOT  # noqa: F401

pytestmark = pytest.mark.gpu


def _dom(nx, ny=None, nz=None, re=400, u0=0.05, rho0=1.0):
    from oracle import oracle
    return {"nx": nx, "ny": ny or nx, "nz": nz or nx, "tau0": oracle.tau(u0, nx, re), "rho0": rho0, "u0": u0}


def _ctx(d, **kw):
    from imexlbm_amd import capi
    print(f"case: {d} {kw}", flush=True)
    return capi.Context(d, **kw)


def _pair(d, prec, steps, **two):
    """get_f and conserved after `steps` steps, two-step on (with `two`) and off."""
    with _ctx(d, precision=prec, two_step=1, **two) as a:
        a.initialize()
        a.step(steps)
        info = a.info()
        fa, ca = a.get_f(), a.conserved()
    with _ctx(d, precision=prec, two_step=0) as b:
        b.initialize()
        b.step(steps)
        assert b.info()["two_step"] == 0
        fb, cb = b.get_f(), b.conserved()
    assert info["two_step"] == 1
    return fa, fb, ca, cb


# pf 2: the partial prefetch of 2-byte formats (tile 3, whose threads
# outnumber its ring cells, falls back to the full one)
CASES = ([("fp64", t, pf) for t in (2, 3) for pf in (0, 1)] +
         [("fp32", t, pf) for t in (0, 1, 2, 3) for pf in (0, 1)] +
         [(p, t, pf) for p in ("fp16", "bf16") for t in (0, 1, 2, 3, 4) for pf in (0, 1, 2)])


TILE_DIMS = {0: (64, 8), 1: (32, 16), 2: (32, 8), 3: (32, 8), 4: (64, 16)}
SHAPES = [((64, 64, 64), 0, 7), ((128, 64, 24), 7, 6), ((64, 32, 3), 2, 5), ((64, 16, 1), 0, 4)]
# every case whose column tile divides the domain
DIVIDING = [(prec, tile, pf, shape, zseg, steps) for (shape, zseg, steps) in SHAPES for (prec, tile, pf) in CASES
            if shape[0] % TILE_DIMS[tile][0] == 0 and shape[1] % TILE_DIMS[tile][1] == 0]


@pytest.mark.parametrize("prec,tile,pf,shape,zseg,steps", DIVIDING)
def test_two_step_bitwise_equals_single_steps(prec, tile, pf, shape, zseg, steps):
    d = _dom(*shape)
    io = (tile + pf + shape[2]) % 2          # both access paths across the cases
    fa, fb, ca, cb = _pair(d, prec, steps, two_step_tile=tile, two_step_zseg=zseg, two_step_prefetch=pf,
                           two_step_io=io)
    assert np.array_equal(fa, fb)
    assert np.array_equal(np.array(ca), np.array(cb))


@pytest.mark.parametrize("prec", ["fp64", "fp16"])
def test_two_step_tile_order(prec):
    d = _dom(64, nz=20)
    tile = 2 if prec == "fp64" else 0
    outs = []
    for x in (1, 8):
        fa, fb, _, _ = _pair(d, prec, 9, two_step_tile=tile, two_step_zseg=6, two_step_xcds=x)
        assert np.array_equal(fa, fb)
        outs.append(fa)
    assert np.array_equal(outs[0], outs[1])


def test_two_step_fp64_matches_oracle_bitwise():
    from oracle import oracle
    d = _dom(32, re=1600)
    with _ctx(d, precision="fp64", two_step=1) as c:
        c.initialize()
        assert c.info()["two_step"] == 1
        c.step(40)
        got = c.get_f()
    o = oracle.Oracle(d["nx"], d["ny"], d["nz"], d["tau0"], d["rho0"], d["u0"])
    o.initialize()
    o.step(40)
    want = o.get_f()
    o.close()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
def test_deferred_collision_streaming_pairs(prec):
    # the driver's call pattern: Collision + Streaming per step, reads at odd
    # and even points; Streaming defers one fused step and launches pairs
    d = _dom(64, nz=16)
    with _ctx(d, precision=prec, two_step=1) as a, _ctx(d, precision=prec, two_step=0) as b:
        a.initialize()
        b.initialize()
        for it in range(1, 12):
            a.collision(); a.streaming()
            b.collision(); b.streaming()
            if it in (3, 4, 7, 11):
                assert np.array_equal(np.array(a.conserved()), np.array(b.conserved()))
            if it == 8:
                a.synchronize()
        # a pending collision, a bare streaming and the get on top
        a.collision(); b.collision()
        assert np.array_equal(a.get_f(), b.get_f())
        a.streaming(); b.streaming()
        a.streaming(); b.streaming()
        assert np.array_equal(a.get_f(), b.get_f())


@pytest.mark.parametrize("prec", ["fp32", "fp16"])   # fp16: the partial prefetch
def test_two_step_graph_replay(prec):
    d = _dom(64, nz=12)
    with _ctx(d, precision=prec, two_step=1, use_graph=1) as a, _ctx(d, precision=prec, two_step=0) as b:
        a.initialize(); b.initialize()
        a.step(1); b.step(1)
        for k in range(4):
            a.step(6); b.step(6)
            a.synchronize()
            assert np.array_equal(a.get_f(), b.get_f())


@pytest.mark.parametrize("prec,whole_x", [("fp16", 1), ("fp64", 0)])
def test_tiled_two_step_graph_replay(prec, whole_x):
    # the ghost fill and the two-step pass captured together: every replay,
    # with eager reads and a synchronisation between replays, must equal the
    # periodic box run one step per pass
    from imexlbm_amd import capi
    d = _dom(64)
    with _ctx(d, precision=prec, mode=capi.TILED, ntiles=8, halo_depth=2, whole_x=whole_x, two_step=1,
              use_graph=1) as a, _ctx(d, precision=prec, two_step=0) as b:
        a.initialize(); b.initialize()
        a.step(1); b.step(1)
        assert a.info()["use_graph"] == 1 and a.info()["two_step"] == 1
        for k in range(4):
            a.step(6); b.step(6)
            a.synchronize()
            assert np.array_equal(a.get_f(), b.get_f()), k
            ca, cb = a.conserved(), b.conserved()
            for i in (0, 4):   # the tiles sum in another order
                assert abs(ca[i] - cb[i]) <= 1e-12 * abs(cb[i]), k


@pytest.mark.parametrize("prec", ["fp64", "fp16"])
def test_autotune_with_two_step_candidates(prec):
    d = _dom(64, nz=32)
    with _ctx(d, precision=prec) as a, _ctx(d, precision=prec, two_step=0) as b:
        a.initialize(); b.initialize()
        a.step(3); b.step(3)
        before = a.get_f()
        assert a.autotune(reps=2) > 0
        assert np.array_equal(a.get_f(), before)
        a.step(9); b.step(9)
        assert np.array_equal(a.get_f(), b.get_f())


@pytest.mark.parametrize("fail_at", [3, 40])       # inside the one-step, then the two-step candidates
def test_failed_autotune_restores_the_launch_shape(fail_at, monkeypatch):
    from imexlbm_amd import capi
    d = _dom(128, ny=64, nz=32)
    with _ctx(d, precision="fp16", use_graph=1) as a, _ctx(d, precision="fp16", two_step=0) as b:
        a.initialize(); b.initialize()
        a.step(3); b.step(3)                      # a graph captured for the current shape
        shape = a.info()
        monkeypatch.setenv("IMEXLBM_AUTOTUNE_FAIL_AT", str(fail_at))
        with pytest.raises(capi.ImexError, match="launch shape restored"):
            a.autotune(reps=1)
        monkeypatch.delenv("IMEXLBM_AUTOTUNE_FAIL_AT")
        assert a.info() == shape
        a.step(9); b.step(9)                      # the graph is captured again
        assert np.array_equal(a.get_f(), b.get_f())


@pytest.mark.parametrize("prec,whole_x", [("fp16", 1), ("bf16", 0), ("fp64", 1)])
def test_tiled_autotune_keeps_state(prec, whole_x):
    # the autotuner times candidates on scratch pools: the tiled state and
    # its ghosts must come back untouched, and the chosen shape must still
    # give the periodic box's populations
    from imexlbm_amd import capi
    d = _dom(64)
    with _ctx(d, precision=prec, mode=capi.TILED, ntiles=8, halo_depth=3, whole_x=whole_x) as a, \
            _ctx(d, precision=prec, two_step=0) as b:
        a.initialize(); b.initialize()
        a.step(3); b.step(3)
        before = a.get_f()
        assert a.autotune(reps=2) > 0
        assert np.array_equal(a.get_f(), before)
        a.step(9); b.step(9)
        assert np.array_equal(a.get_f(), b.get_f())


def test_tiled_autotune_keeps_one_step_ghost_tables():
    # two tiles of 64 x 64 x 32: too few workgroups for the model to pick two
    # steps per pass, so the ghost tables are built one layer deep; the
    # autotuner must not switch to two steps on them
    from imexlbm_amd import capi
    d = _dom(64)
    with _ctx(d, precision="fp32", mode=capi.TILED, ntiles=2, halo_depth=1) as a, \
            _ctx(d, precision="fp32", two_step=0) as b:
        assert a.info()["two_step"] == 0
        a.initialize(); b.initialize()
        a.step(3); b.step(3)
        assert a.autotune(reps=2) > 0
        assert a.info()["two_step"] == 0
        a.step(8); b.step(8)
        assert np.array_equal(a.get_f(), b.get_f())


def test_two_step_eligibility():
    from imexlbm_amd import capi
    with _ctx(_dom(32), precision="fp64") as c:   # auto: 4 column tiles are too few workgroups
        assert c.info()["two_step"] == 0
    with _ctx(_dom(256, nz=64), precision="fp16") as c:   # 128 tiles x 4 sweeps of 16 planes
        info = c.info()
        assert info["two_step"] == 1 and info["two_step_zseg"] == 16
    d = _dom(48, ny=40, nz=8)
    with _ctx(d, precision="fp32") as c:          # 6 column tiles: too few workgroups, auto stays off
        assert c.info()["two_step"] == 0
    with pytest.raises(capi.ImexError):           # narrower than half of every column tile
        _ctx(_dom(12, ny=40, nz=8), precision="fp32", two_step=1)
    with pytest.raises(capi.ImexError):           # the 64x8 tile over a box 31 cells across
        _ctx(_dom(31, ny=40, nz=8), precision="fp32", two_step=1, two_step_tile=0)
    with pytest.raises(capi.ImexError):
        _ctx(_dom(64), precision="fp64", two_step=1, two_step_tile=0)   # FP64 builds only 32x8
    with pytest.raises(capi.ImexError):      # tiles of 24 x 40 cells: no column tile divides them
        _ctx(_dom(48, ny=40, nz=8), precision="fp32", two_step=1, mode=capi.TILED, ntiles=2)
    with pytest.raises(capi.ImexError):      # a batch member of 12 cells across
        _ctx([_dom(64), _dom(12, ny=64)], precision="fp32", two_step=1)
    with pytest.raises(capi.ImexError):      # one launch per instance: no pass table
        _ctx([_dom(64), _dom(64)], precision="fp32", two_step=1, sequential=1)
    with _ctx(_dom(64, nz=1), precision="fp32", mode=capi.TILED, ntiles=2) as c:   # tiles one plane deep
        assert c.info()["two_step"] == 0


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16", "bf16"])
@pytest.mark.parametrize("ntiles,halo,direct,whole_x", [(2, 1, 1, 0), (4, 2, 0, 0), (8, 3, 1, 0), (8, 1, 0, 0),
                                                        (8, 3, 1, 1), (4, 2, 0, 1)])
def test_tiled_two_step_equals_wrap_single_steps(prec, ntiles, halo, direct, whole_x):
    # tiles exchange two ghost layers once per pass (whole_x: in y and z
    # only, x wraps in the kernel); populations must equal the periodic
    # single-box run bit for bit
    from imexlbm_amd import capi
    d = _dom(64, ny=64, nz=24)
    with _ctx(d, precision=prec, mode=capi.TILED, ntiles=ntiles, halo_depth=halo, two_step=1,
              halo_direct=direct, whole_x=whole_x) as a:
        a.initialize()
        a.step(7)
        info = a.info()
        assert info["two_step"] == 1 and info["halo_depth"] == 2 and info["halo_direct"] == direct
        assert info["whole_x"] == whole_x and (info["rx"] == 1 or not whole_x)
        fa, ca = a.get_f(), a.conserved()
        a.collision(); a.streaming(); a.collision(); a.streaming()
        a.collision(); a.streaming()
        fa2 = a.get_f()
    with _ctx(d, precision=prec, two_step=0) as b:
        b.initialize()
        b.step(7)
        fb, cb = b.get_f(), b.conserved()
        b.step(3)
        fb2 = b.get_f()
    assert np.array_equal(fa, fb)
    assert np.array_equal(fa2, fb2)
    for k in (0, 4):
        assert abs(ca[k] - cb[k]) <= 1e-12 * abs(cb[k])


@pytest.mark.parametrize("prec,shape,tile", [("fp64", (96, 96, 48), 3), ("fp16", (192, 96, 96), 4)])
def test_tiled_two_step_three_boxes_across_x(prec, shape, tile):
    # 27 tiles in a 3 x 3 x 3 split: each box's lower and upper x neighbours
    # are different boxes, so the pass fills two different boxes' side
    # buffers (in 2 x 2 x 2 both neighbours are the same box)
    from imexlbm_amd import capi
    d = _dom(*shape)
    with _ctx(d, precision=prec, mode=capi.TILED, ntiles=27, halo_depth=3, two_step=1, two_step_tile=tile,
              halo_direct=1) as a:
        a.initialize()
        a.step(9)
        info = a.info()
        fa, ca = a.get_f(), a.conserved()
    assert (info["rx"], info["ry"], info["rz"]) == (3, 3, 3)
    assert info["two_step"] == 1 and info["two_step_tile"] == tile and info["halo_direct"] == 1
    with _ctx(d, precision=prec, two_step=0) as b:
        b.initialize()
        b.step(9)
        fb, cb = b.get_f(), b.conserved()
    assert np.array_equal(fa, fb)
    for k in (0, 4):
        assert abs(ca[k] - cb[k]) <= 1e-12 * abs(cb[k])


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("two,tile,pf", [(0, None, None), (1, 0, 2), (1, 1, 2), (1, 2, 2), (1, 4, 2),
                                         (1, 4, 0), (1, 2, 0)])
def test_tiled_16bit_collision_equals_wrap(prec, two, tile, pf):
    # over ghost-bounded boxes the FP16 collision reads the stored halves
    # through mixed-precision FMAs, the periodic kernels widen them first:
    # the populations must still agree bit for bit, one step per pass and
    # in every column tile and prefetch mode of the two-step pass (1x2x4
    # tiles of 64x32x16, which all four tiles divide)
    from imexlbm_amd import capi
    d = _dom(64)
    kw = dict(two_step=two)
    if tile is not None:
        kw.update(two_step_tile=tile, two_step_prefetch=pf)
    with _ctx(d, precision=prec, mode=capi.TILED, ntiles=8, halo_depth=3, whole_x=1, **kw) as a:
        a.initialize()
        a.step(7)
        info = a.info()
        fa = a.get_f()
    assert info["two_step"] == two and (info["rx"], info["ry"], info["rz"]) == (1, 2, 4)
    if tile is not None:
        assert info["two_step_tile"] == tile and info["two_step_prefetch"] == pf
    with _ctx(d, precision=prec, two_step=0) as b:
        b.initialize()
        b.step(7)
        fb = b.get_f()
    assert np.array_equal(fa, fb)


def test_two_step_full_size_fp16_512():
    # the bench configuration: conserved sums of the two paths bit-identical
    # after 1 + 4 steps (each context holds 29 GB)
    from imexlbm_amd import capi
    n = 512
    d = {"nx": n, "ny": n, "nz": n, "tau0": 0.05 * n / 1600 / (1.0 / 3.0) / 2.0 / 3.1415926,
         "rho0": 1.0, "u0": 0.05}
    out = []
    for two in (1, 0):
        with _ctx(d, precision="fp16", two_step=two) as c:
            c.initialize()
            c.step(5)
            out.append(np.array(c.conserved()))
            m = c.get_macro()
            out.append(np.array([float(np.sum(a, dtype=np.float64)) for a in m]))
            assert np.all(np.isfinite(out[-1]))
            assert c.info()["two_step"] == two
            del m
    assert np.array_equal(out[0], out[2])
    assert np.array_equal(out[1], out[3])


@pytest.mark.parametrize("prec", ["fp64", "fp16"])
def test_batched_two_step_equals_instances_alone(prec):
    # three periodic instances of different sizes, Re and rho0 in one pass
    # per two steps; each must equal its own single-instance one-step run
    doms = [_dom(64, ny=32, nz=16, re=300), _dom(32, ny=64, nz=8, re=900, rho0=1.1),
            _dom(96, ny=16, nz=5, re=150, u0=0.03)]
    with _ctx(doms, precision=prec, two_step=1) as a:
        a.initialize()
        a.step(9)
        assert a.info()["two_step"] == 1
        got = [a.get_f(k) for k in range(3)]
    for k, d in enumerate(doms):
        with _ctx(d, precision=prec, two_step=0) as b:
            b.initialize()
            b.step(9)
            assert np.array_equal(got[k], b.get_f())


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_batched_64x16_two_step_equals_instances_alone(prec):
    # a batch whose instances are all at least 64 x 16 across, in the 64x16
    # column tile with the partial prefetch: ragged last columns (100, 80)
    # and sweeps cut at each instance's own depth
    doms = [_dom(100, ny=40, nz=12, re=300), _dom(64, ny=64, nz=9, re=900, rho0=1.1),
            _dom(80, ny=16, nz=5, re=150, u0=0.03)]
    with _ctx(doms, precision=prec, two_step=1, two_step_tile=4, two_step_zseg=4, two_step_prefetch=2) as a:
        a.initialize()
        a.step(9)
        info = a.info()
        assert info["two_step"] == 1 and info["two_step_tile"] == 4 and info["two_step_prefetch"] == 2
        got = [a.get_f(k) for k in range(3)]
    for k, d in enumerate(doms):
        with _ctx(d, precision=prec, two_step=0) as b:
            b.initialize()
            b.step(9)
            assert np.array_equal(got[k], b.get_f()), k


RAGGED_TILES = [("fp64", 3), ("fp64", 2), ("fp32", 0), ("fp32", 1), ("fp16", 0), ("fp16", 3), ("bf16", 1),
                ("fp16", 4), ("bf16", 4)]


@pytest.mark.parametrize("prec,tile", [("fp16", 4), ("bf16", 4), ("fp32", 0), ("fp64", 3)])
def test_batched_narrow_instances_in_wide_column_tiles(prec, tile):
    # a batch with instances narrower than the column tile (32 and 48 cells
    # across for the 64-wide tiles, 16 and 24 for the 32-wide FP64 one) in
    # one pass per two steps; each must equal its own one-step run
    tx = TILE_DIMS[tile][0]
    doms = [_dom(tx // 2, ny=32, nz=10, re=300), _dom(3 * tx // 4, ny=40, nz=7, re=900, rho0=1.1),
            _dom(tx + 36, ny=24, nz=5, re=150, u0=0.03), _dom(2 * tx, ny=16, nz=12, re=600)]
    with _ctx(doms, precision=prec, two_step=1, two_step_tile=tile, two_step_zseg=4) as a:
        a.initialize()
        a.step(9)
        info = a.info()
        assert info["two_step"] == 1 and info["two_step_tile"] == tile
        got = [a.get_f(k) for k in range(len(doms))]
    for k, d in enumerate(doms):
        with _ctx(d, precision=prec, two_step=0) as b:
            b.initialize()
            b.step(9)
            assert np.array_equal(got[k], b.get_f()), k
# every box at least half a column tile across and one tile deep in y; 32
# and 16 cells across are exactly half the 64- and 32-wide tiles, whose ring
# then wraps twice in x
RAGGED = [(prec, tile, shape)
          for shape in [(100, 40, 12), (48, 72, 7), (96, 50, 5), (33, 17, 4), (32, 24, 6), (16, 40, 5)]
          for (prec, tile) in RAGGED_TILES
          if 2 * shape[0] >= TILE_DIMS[tile][0] and shape[1] >= TILE_DIMS[tile][1]]


@pytest.mark.parametrize("prec,tile,shape", RAGGED)
def test_two_step_ragged_column_tiles(prec, tile, shape):
    # the last column tile runs past nx / ny: its ring wraps onto real cells
    # (twice in x for a box narrower than the tile) and its cells beyond the
    # box are not stored
    d = _dom(*shape)
    for io in (0, 1):
        fa, fb, ca, cb = _pair(d, prec, 7, two_step_tile=tile, two_step_zseg=3, two_step_io=io)
        assert np.array_equal(fa, fb)
    assert np.array_equal(fa, fb)
    assert np.array_equal(np.array(ca), np.array(cb))


def test_two_step_fp64_128_matches_oracle_bitwise():
    # the default FP64 pass (32x8 columns, 16-plane sweeps) at 128^3 against
    # the CPU restatement of the reference, populations bit for bit
    from oracle import oracle
    d = _dom(128, re=1600)
    with _ctx(d, precision="fp64") as c:
        c.initialize()
        assert c.info()["two_step"] == 1
        c.step(10)
        got = c.get_f()
        cons = c.conserved()
    o = oracle.Oracle(d["nx"], d["ny"], d["nz"], d["tau0"], d["rho0"], d["u0"])
    o.initialize()
    o.step(10)
    want = o.get_f()
    wc = o.conserved_compensated()
    o.close()
    assert np.array_equal(got, want)
    assert abs(cons[4] - wc[4]) / wc[4] < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16"])
def test_direction_padding_defaults_and_bit_identity(prec):
    # periodic FP64/FP32 contexts pad their direction planes by 32 elements
    # (the planes of a 64^3 box are a power of two apart); the padding is a
    # layout choice only: populations are the same with and without it
    from imexlbm_amd import capi
    d = _dom(64, nz=20, re=800)
    with _ctx(d, precision=prec) as a:
        a.initialize()
        a.step(9)
        pad = a.info()["dir_pad"]
        fa = a.get_f()
    assert pad == (32 if prec in ("fp64", "fp32") else 0)
    with _ctx(d, precision=prec, dir_pad=0 if pad else 64) as b:
        b.initialize()
        b.step(9)
        assert b.info()["dir_pad"] == (0 if pad else 64)
        fb = b.get_f()
    assert np.array_equal(fa, fb)
    with _ctx(d, precision=prec, mode=capi.TILED, ntiles=2) as t:   # tiled boxes: no padding
        assert t.info()["dir_pad"] == 0
