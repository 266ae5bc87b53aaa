"""Batched boxes (nmhd_batch_*): many boxes of uneven sizes (periodic or
shearing) in per-tile-shape or persistent launches over a work table,
per-box dt and ghost fills, HIP-graph replay and the autotuner.

The batched kernels run the single-box tile functions, and a cell's
arithmetic does not depend on the tiling, so the invariant checked here is
bit-identity: a box advanced in a batch (eager or graph, any tile shape, z
chunk, launch mode or workgroup count, tuned or not) equals the same box advanced alone
by a Context on the tiled path, state and time, bit for bit.  The tiled path
itself is checked against the oracle in test_gpu_parity.py; one batched box
is compared with the oracle here as well.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from mhd_checks import bitwise_equal, div_b, group_rel_err, perturb_ulps
from nmhd_amd import capi, problems as P

pytestmark = pytest.mark.gpu

SMALL = [(16, 24, 8), (32, 16, 16), (24, 24, 24), (48, 16, 8), (16, 16, 32), (64, 8, 12)]


def announce(msg):
    print("\n[case] " + msg, flush=True)


def small_boxes(seed=11):
    out = []
    for b, shape in enumerate(SMALL):
        cfg = P.make_config(shape, cfl=0.4)
        if b % 2 == 0:
            W = P.orszag_tang_noisy(cfg, seed=seed + b, b0=0.3 + 0.05 * b)
        else:
            u = P.uniform01(seed, b, 0)
            W = P.blast(cfg, (float(np.cos(6.0 * u)), float(np.sin(6.0 * u)), 0.3))
        out.append((cfg, W))
    return out


def make_batch(boxes):
    B = capi.Batch([c for c, _ in boxes])
    for i, (_, W) in enumerate(boxes):
        B.set_state(i, W)
    return B


def single_runs(boxes, goal, nsteps):
    out = []
    for cfg, W in boxes:
        g = capi.Context(cfg, path=capi.PATH_TILED)
        g.set_state(W)
        k, _ = g.run(goal, nsteps)
        out.append((g.get_state(), g.get_time(), k))
        g.close()
    return out


def assert_batch_equals(B, ref):
    for i, (W, (t, gen), k) in enumerate(ref):
        got = B.get_state(i)
        assert bitwise_equal(got, W), "box %d state differs" % i
        tb, gb = B.get_time(i)
        assert tb == t and gb == gen, (i, tb, t, gb, gen)


def test_batch_eager_equals_single_boxes(gpu_ready):
    announce("batch eager 6 boxes x 5 steps vs single contexts")
    boxes = small_boxes()
    ref = single_runs(boxes, 1e30, 5)
    B = make_batch(boxes)
    k, reached = B.run(1e30, 5, use_graph=False)
    assert k == 5 and not reached
    assert_batch_equals(B, ref)
    # the boxes took different dt
    dts = {B.status(i)[2] for i in range(B.nbox)}
    assert len(dts) > 1


def shearing_mix():
    """Shearing boxes of uneven sizes (one with vertical ghost planes too,
    one with all three quirk corrections, one with first-order
    reconstruction) mixed with periodic boxes (one first-order)."""
    out = []
    for k, shape in enumerate([(16, 16, 8), (24, 16, 16), (32, 24, 8)]):
        cfg = P.config_c2(shape)
        if k == 2:
            cfg = dict(cfg, fixes=capi.FIX_ALL)
        out.append((cfg, P.shearing_box_mri(cfg, seed=40 + k)))
    cfg = P.make_config((16, 24, 8), margin=(0, 4, 2), cfl=0.4, source=1, shear_y=1)
    out.append((cfg, P.shearing_box_mri(cfg, seed=44)))
    # first-order reconstruction in one shearing and one periodic box
    cfg = dict(P.config_c2((16, 16, 16)), recon=1)
    out.append((cfg, P.shearing_box_mri(cfg, seed=45)))
    per = small_boxes(seed=31)[:3]
    per[1] = (dict(per[1][0], recon=1), per[1][1])
    return out + per


def test_batch_shearing_boxes_equal_single_boxes(gpu_ready):
    """Shearing boxes in a batch (per-box shear remap of the radial ghost
    rows at each box's own time, periodic vertical ghost planes, the ghost
    cells' CFL added to each box's minimum), mixed with periodic boxes:
    every box equals its single-box tiled context bit for bit, eagerly and
    replayed from the graph, and across a goal that some boxes reach."""
    announce("batch of shearing and periodic boxes vs single contexts")
    boxes = shearing_mix()
    ref = single_runs(boxes, 1e30, 6)
    B = make_batch(boxes)
    k, _ = B.run(1e30, 3, use_graph=False)
    assert k == 3
    k, _ = B.run(1e30, 3, use_graph=True)
    assert k == 3
    assert_batch_equals(B, ref)
    # a goal ahead of every box's time: each box clips its own last step
    # and stops there
    goal = max(t for _, (t, _), _ in ref) + 0.02
    ref2 = []
    for (cfg, W), (Wr, (t, gen), _) in zip(boxes, ref):
        g = capi.Context(cfg, path=capi.PATH_TILED)
        g.set_state(Wr)
        g.set_time(t, gen)
        g.run(goal, 5)
        ref2.append((g.get_state(), g.get_time(), 0))
        g.close()
    B.run(goal, 5)
    assert_batch_equals(B, ref2)
    reached = [B.status(i)[1] for i in range(B.nbox)]
    assert any(reached), reached
    B.close()


def test_batch_with_cfl_speed_fix_equals_single_boxes(gpu_ready):
    """Boxes with NMHD_FIX_CFL_SPEED (|v| + c) in a batch, mixed with boxes
    without it: each equals its single-box tiled context bit for bit, and the
    fixed boxes take their own, shorter steps (each box record carries its
    own corrections; a bulk flow of -0.3 along x makes |v| + c exceed
    |v + c| in the fastest cells)."""
    announce("batch with per-box CFL-speed fix")
    boxes = []
    for i, (cfg, W) in enumerate(small_boxes(seed=23)):
        if i % 2:
            W = W.copy()
            W[1] -= np.float32(0.3)
            cfg = dict(cfg, fixes=capi.FIX_CFL_SPEED)
        boxes.append((cfg, W))
    ref = single_runs(boxes, 1e30, 4)
    B = make_batch(boxes)
    k, _ = B.run(1e30, 4)
    assert k == 4
    assert_batch_equals(B, ref)
    plain = single_runs([(dict(c, fixes=0), W) for c, W in boxes], 1e30, 4)
    for i in range(1, len(boxes), 2):
        assert ref[i][1][0] < plain[i][1][0], i


def test_batch_graph_equals_eager(gpu_ready):
    announce("batch graph vs eager, 7 + 4 steps (every rotation phase, replays)")
    boxes = small_boxes(seed=3)
    A = make_batch(boxes)
    G = make_batch(boxes)
    A.run(1e30, 7, use_graph=False)
    G.run(1e30, 7, use_graph=True)
    for i in range(A.nbox):
        assert bitwise_equal(A.get_state(i), G.get_state(i)), i
    A.run(1e30, 4, use_graph=False)
    G.run(1e30, 4, use_graph=True)
    ref = single_runs(boxes, 1e30, 11)
    assert_batch_equals(G, ref)
    assert_batch_equals(A, ref)


@pytest.mark.parametrize("ft,ut,fz,uz,fw,uw", [
    (capi.TILE_64x4, capi.TILE_64x4, 8, 32, 1, 4),
    (capi.TILE_32x8, capi.TILE_16x16, 5, 7, 3, 1),
    (capi.TILE_16x16, capi.TILE_32x8, 32, 3, 2, 2),
    (capi.TILE_32x8, capi.TILE_64x4, 8, 16, 0, 0),
    (capi.TILE_64x4, capi.TILE_16x16, 3, 8, 0, 3),
    (capi.TILE_16x16, capi.TILE_32x8, 16, 5, 2, 0),
    (capi.TILE_32x8, capi.TILE_64x4, 1, 2, 0, 0),
    (capi.TILE_16x16, capi.TILE_32x8, 2, 1, 2, 2),
])
def test_batch_tilings_bit_identical(gpu_ready, ft, ut, fz, uz, fw, uw):
    announce("batch tiling faces %d update %d zc %d/%d wg %d/%d" % (ft, ut, fz, uz, fw, uw))
    boxes = small_boxes(seed=5)
    ref = single_runs(boxes, 1e30, 3)
    B = make_batch(boxes)
    B.set_tiles(ft, ut)
    B.set_launch(fz, uz, fw, uw)
    B.run(1e30, 3, use_graph=True)
    assert_batch_equals(B, ref)


@pytest.mark.parametrize("fw,uw", [(0, 0), (2, 2)])
def test_batch_mixed_tile_shapes_bit_identical(gpu_ready, fw, uw):
    announce("batch with a different tile shape per box, wg %d/%d" % (fw, uw))
    boxes = small_boxes(seed=9)
    ref = single_runs(boxes, 1e30, 3)
    B = make_batch(boxes)
    for i in range(B.nbox):
        B.set_tiles(i % 3, (i + 1) % 3, box=i)
    B.set_launch(8, 16, fw, uw)
    B.run(1e30, 3, use_graph=True)
    assert_batch_equals(B, ref)


def test_batch_per_box_goal_clipping(gpu_ready):
    announce("batch goal clipping: boxes reach the goal at different steps")
    boxes = small_boxes(seed=8)
    # a goal that some boxes reach within 6 steps and others do not
    probe = make_batch(boxes)
    probe.run(1e30, 1, use_graph=False)
    dts = sorted(probe.status(i)[2] for i in range(probe.nbox))
    goal = float(np.float32(3.5 * dts[len(dts) // 2]))
    ref = single_runs(boxes, goal, 6)
    B = make_batch(boxes)
    kmin, all_reached = B.run(goal, 6, use_graph=True)
    assert_batch_equals(B, ref)
    reached = [B.status(i)[1] for i in range(B.nbox)]
    assert any(reached) and not all(reached)
    for i, r in enumerate(reached):
        t, _ = B.get_time(i)
        if r:
            assert t == pytest.approx(goal, rel=0, abs=1e-6)
    assert kmin == min(k for _, _, k in ref)


def test_batch_autotune_cached_and_identical(gpu_ready, tmp_path):
    announce("batch autotune (cache in tmp) and results after tuning")
    boxes = small_boxes(seed=9)
    ref = single_runs(boxes, 1e30, 3)
    B = make_batch(boxes)
    cache = str(tmp_path / "tune.txt")
    n1 = B.autotune(cache)
    assert n1 == len(set(SMALL))
    assert os.path.exists(cache)
    text = open(cache).read()
    assert text.count("\nshape ") == len(set(SMALL)) and "\nlaunch " in text
    B2 = make_batch(boxes)
    assert B2.autotune(cache) == 0  # every shape from the cache
    assert B2.get_launch() == B.get_launch()
    B.run(1e30, 3, use_graph=True)
    assert_batch_equals(B, ref)


def test_batch_box_against_oracle(gpu_ready):
    announce("batched box vs oracle, 3 steps")
    cfg = P.make_config((16, 24, 8), cfl=0.4)
    W = P.orszag_tang_noisy(cfg, seed=21)
    other_cfg = P.make_config((32, 16, 16), cfl=0.4)
    B = capi.Batch([other_cfg, cfg])
    B.set_state(0, P.orszag_tang(other_cfg))
    B.set_state(1, W)
    B.run(1e30, 3, use_graph=True)
    o = po.Oracle(cfg, W)
    o.run(1e30, 3)
    got, want = B.get_state(1), o.get_state()
    errs = group_rel_err(got, want)
    assert errs.max() < 1e-3, errs
    assert B.get_time(1)[0] == pytest.approx(o.time, rel=1e-6)
    bmax = np.abs(got[5:8]).max()
    assert np.abs(div_b(got, cfg["d"])).max() * max(cfg["d"]) < 1e-6 * bmax



def test_batch_thin_boxes_equal_single_boxes(gpu_ready):
    """Boxes one cell thick (the reference's shock-tube demo along x, y and
    z, 128 cells long, one with first-order reconstruction) next to a 3D box:
    tiles far wider than the box, periodic wrap onto the same cell."""
    announce("batch of thin shock-tube boxes + a 3D box vs single contexts")
    boxes = [P.shock_tube(0, 128), P.shock_tube(1, 128), P.shock_tube(2, 128, recon=1)]
    cfg, W = P.shock_tube(1, 64)
    boxes.append((cfg, W))
    boxes.append(small_boxes()[2])
    ref = single_runs(boxes, 0.05, 20)
    B = make_batch(boxes)
    B.run(0.05, 20, use_graph=True)
    assert_batch_equals(B, ref)


def test_c4_workload_autotuned_graph_steps(gpu_ready, tmp_path):
    """The multi-box workload itself: 64 boxes of edges from {16, ..., 96}
    (seed 1004; Orszag-Tang and blast boxes), autotuned, then 5 steps
    replayed from the captured graph.  Every box equals a single-box tiled
    context bit for bit (state and time), and two Orszag-Tang and two blast
    boxes stay within three times the oracle's own one-ulp sensitivity of
    the oracle (both states are mirror symmetric, so CT upwind sums are
    exactly zero on whole rows of edges and their choice is a round-off
    tie)."""
    announce("C4: 64 boxes, autotune, 5 graph steps vs single contexts and oracle")
    boxes = P.batch_boxes(seed=1004, nbox=64)
    B = make_batch(boxes)
    B.autotune(str(tmp_path / "tune.txt"))
    print("launch (faces zchunk, update zchunk, faces wg/CU, update wg/CU):", B.get_launch(),
          flush=True)
    k, reached = B.run(1e30, 5, use_graph=True)
    assert k == 5 and not reached
    ref = single_runs(boxes, 1e30, 5)
    assert_batch_equals(B, ref)
    checked = {0: [], 1: []}
    for i, (cfg, W) in enumerate(boxes):
        kind = i % 2
        if len(checked[kind]) == 2:
            continue
        o = po.Oracle(cfg, W)
        o2 = po.Oracle(cfg, perturb_ulps(W))
        o.run(1e30, 5)
        o2.run(1e30, 5)
        want = o.get_state()
        env = group_rel_err(o2.get_state(), want)
        err = group_rel_err(B.get_state(i), want)
        print("box %d %s %s: err %s, oracle 1-ulp envelope %s"
              % (i, "OT" if kind == 0 else "blast", cfg["global_size"],
                 np.array2string(err, precision=2), np.array2string(env, precision=2)))
        assert np.all(err <= 3.0 * env + 1e-6), (i, err, env)
        # the dt sequence follows the state's round-off ties: 1e-5 per step
        assert B.get_time(i)[0] == pytest.approx(o.time, rel=1e-5)
        checked[kind].append(i)
    assert len(checked[0]) == len(checked[1]) == 2


def test_batch_second_run_after_goal(gpu_ready):
    """Boxes that reached the goal keep the CFL minimum of their last active
    step through the trailing inactive steps, so a second run to a later goal
    takes the same steps as single contexts whose second run recomputes the
    CFL over the whole box."""
    announce("batch run(g1) then run(g2) vs single contexts")
    boxes = small_boxes(seed=8)
    probe = make_batch(boxes)
    probe.run(1e30, 1, use_graph=False)
    dts = sorted(probe.status(i)[2] for i in range(probe.nbox))
    g1 = float(np.float32(3.5 * dts[len(dts) // 2]))
    g2 = float(np.float32(2.0 * g1))
    ref = []
    for cfg, W in boxes:
        g = capi.Context(cfg, path=capi.PATH_TILED)
        g.set_state(W)
        g.run(g1, 40)
        g.set_path(capi.PATH_TILED)  # the second run starts from a full CFL pass
        k, _ = g.run(g2, 40)
        ref.append((g.get_state(), g.get_time(), k))
        g.close()
    B = make_batch(boxes)
    B.run(g1, 40, use_graph=True)
    assert all(B.status(i)[1] for i in range(B.nbox))
    B.run(g2, 40, use_graph=True)
    assert_batch_equals(B, ref)
    for i in range(B.nbox):
        assert B.status(i)[0] == ref[i][2], i
