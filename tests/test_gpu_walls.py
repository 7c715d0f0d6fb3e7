"""GPU parity with mirror walls: the engine's direct and Barnes-Hut paths
(point and group traversals) against the CPU oracle on the same tangle,
with the reference's wall image tables.  The oracle is pinned bit for bit
to the compiled reference and to the reflected-copy identity in
tests/test_walls.py.

Tolerances: FP64 velocities within 1e-10 relative L2 (summation order),
tangents and normals 1e-14, positions after one RK4 step 1e-14 cm; mixed
precision within 1e-5 of the FP64 path."""
import sys

import numpy as np
import pytest

import openvort_amd as ov
from conftest import rel_l2
from test_walls import wall_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


CASES = [
    # (table, method, traverse, precision)
    (ov.IMAGES_WALL_CANAL, ov.METHOD_DIRECT, ov.TRAVERSE_POINT, ov.PREC_FP64),
    (ov.IMAGES_WALL_CANAL, ov.METHOD_BH, ov.TRAVERSE_POINT, ov.PREC_FP64),
    (ov.IMAGES_WALL_CANAL, ov.METHOD_BH, ov.TRAVERSE_GROUP, ov.PREC_FP64),
    (ov.IMAGES_WALL_1_18, ov.METHOD_BH, ov.TRAVERSE_POINT, ov.PREC_FP64),
    (ov.IMAGES_WALL_1_18, ov.METHOD_DIRECT, ov.TRAVERSE_POINT, ov.PREC_FP64),
    (ov.IMAGES_WALL_PARALLEL_PLANES, ov.METHOD_BH, ov.TRAVERSE_GROUP, ov.PREC_FP64),
    (ov.IMAGES_WALL_2_2, ov.METHOD_BH, ov.TRAVERSE_POINT, ov.PREC_FP64),
    (ov.IMAGES_WALL_1_OPEN, ov.METHOD_BH, ov.TRAVERSE_GROUP, ov.PREC_FP64),
]


@pytest.mark.parametrize("config,method,traverse,precision", CASES)
def test_wall_update_and_rk4_vs_oracle(oracle, config, method, traverse, precision):
    t, p = wall_case(config, method)
    p.traverse = traverse
    p.precision = precision
    live = t.live
    say(f"walls table {config} method {method} traverse {traverse}: {len(live)} points")
    want = oracle.update_tangle(t, p)
    with ov.Engine(p) as e:
        e.upload(t)
        e.update_tangle(0.0)
        o = e.download()
        assert rel_l2(o["tangents"][live], want["tangents"][live]) < 1e-14
        assert rel_l2(o["normals"][live], want["normals"][live]) < 1e-14
        assert rel_l2(o["vs"][live], want["vs"][live]) < 1e-10
        assert rel_l2(o["vels"][live], want["vels"][live]) < 1e-10
        dt = 1e-4
        e.rk4_step(0.0, dt)
        x = e.download()["vnodes"][live]
    w = ov.Tangle(t.vnodes.copy(), t.conn.copy(), t.status.copy())
    oracle.rk4_step(w, p, 0.0, dt, want["vels"])
    assert np.abs(x - w.vnodes[live]).max() < 1e-14


def test_wall_group_mixed_precision(oracle):
    t, p = wall_case(ov.IMAGES_WALL_CANAL, ov.METHOD_BH)
    p.traverse = ov.TRAVERSE_GROUP
    live = t.live
    say(f"walls canal, group traversal FP64 and mixed: {len(live)} points")
    out = {}
    for prec in (ov.PREC_FP64, ov.PREC_MIXED):
        p.precision = prec
        with ov.Engine(p) as e:
            e.upload(t)
            e.update_tangle(0.0)
            out[prec] = e.download()
    assert rel_l2(out[ov.PREC_MIXED]["vs"][live], out[ov.PREC_FP64]["vs"][live]) < 1e-5


@pytest.mark.parametrize("config", [ov.IMAGES_WALL_CANAL, ov.IMAGES_WALL_1_18,
                                    ov.IMAGES_WALL_PARALLEL_PLANES])
def test_wall_list_small_theta_matches_point(config):
    """theta -> 0: the list traversal (mirrored copies as reflected, reversed
    sources) and the point traversal (reflected targets, reflected
    velocities) reduce to the same segment sums; summation order and the
    frame of the rounding alone separate them."""
    t, p = wall_case(config, ov.METHOD_BH, theta=1e-6)
    live = t.live
    say(f"walls table {config}: list vs point traversal at theta 1e-6, FP64")
    out = {}
    for trav in (ov.TRAVERSE_LIST, ov.TRAVERSE_POINT):
        p.traverse = trav
        with ov.Engine(p) as e:
            e.upload(t)
            e.update_tangle(0.0)
            out[trav] = e.download()
    for key in ("vs", "vels"):
        assert rel_l2(out[ov.TRAVERSE_LIST][key][live], out[ov.TRAVERSE_POINT][key][live]) < 1e-11


def test_pinned_list_small_theta_matches_point():
    from test_walls import pinned_case
    t, p = pinned_case(ov.METHOD_BH, True)
    p.theta = 1e-6
    live = t.live
    say("pinned lines: list vs point traversal at theta 1e-6, FP64")
    out = {}
    for trav in (ov.TRAVERSE_LIST, ov.TRAVERSE_POINT):
        p.traverse = trav
        with ov.Engine(p) as e:
            e.upload(t)
            e.update_tangle(0.0)
            out[trav] = e.download()
    assert rel_l2(out[ov.TRAVERSE_LIST]["vels"][live], out[ov.TRAVERSE_POINT]["vels"][live]) < 1e-11


@pytest.mark.parametrize("ppw", [16, 32])
def test_wall_list_mixed_vs_fp64(ppw):
    t, p = wall_case(ov.IMAGES_WALL_CANAL, ov.METHOD_BH)
    p.traverse = ov.TRAVERSE_LIST
    p.points_per_wave = ppw
    live = t.live
    say(f"walls canal, list traversal {ppw} points per wave: FP64 and mixed")
    out = {}
    for prec in (ov.PREC_FP64, ov.PREC_MIXED):
        p.precision = prec
        with ov.Engine(p) as e:
            e.upload(t)
            e.update_tangle(0.0)
            out[prec] = e.download()
    assert rel_l2(out[ov.PREC_MIXED]["vs"][live], out[ov.PREC_FP64]["vs"][live]) < 1e-6


def test_wall_reuse_lists_is_rejected():
    t, p = wall_case(ov.IMAGES_WALL_CANAL, ov.METHOD_BH)
    p.traverse = ov.TRAVERSE_LIST
    p.reuse_lists = 1
    say("walls with reuse_lists: EUNSUPPORTED")
    with pytest.raises(ov.OvgError) as ei:
        with ov.Engine(p):
            pass
    assert ei.value.code == ov.EUNSUPPORTED


PINNED_CASES = [
    # (method, traverse, slip)
    (ov.METHOD_DIRECT, ov.TRAVERSE_POINT, False),
    (ov.METHOD_DIRECT, ov.TRAVERSE_POINT, True),
    (ov.METHOD_BH, ov.TRAVERSE_POINT, False),
    (ov.METHOD_BH, ov.TRAVERSE_GROUP, True),
]


@pytest.mark.parametrize("method,traverse,slip", PINNED_CASES)
def test_pinned_lines_vs_oracle(oracle, method, traverse, slip):
    from test_walls import pinned_case
    t, p = pinned_case(method, slip)
    p.traverse = traverse
    live = t.live
    say(f"pinned lines (slip {slip}) method {method} traverse {traverse}: {len(live)} points")
    want = oracle.update_tangle(t, p)
    with ov.Engine(p) as e:
        e.upload(t)
        e.update_tangle(0.0)
        o = e.download()
        assert rel_l2(o["tangents"][live], want["tangents"][live]) < 1e-14
        assert rel_l2(o["normals"][live], want["normals"][live]) < 1e-14
        assert rel_l2(o["vs"][live], want["vs"][live]) < 1e-10
        assert rel_l2(o["vels"][live], want["vels"][live]) < 1e-10
        ends = np.where(t.status[:, 0] != ov.FREE)[0]
        if slip:
            assert np.all(o["vels"][ends, 2] == 0)
        else:
            assert np.all(o["vels"][ends] == 0)
        dt = 1e-4
        e.rk4_step(0.0, dt)
        x = e.download()["vnodes"][live]
    w = ov.Tangle(t.vnodes.copy(), t.conn.copy(), t.status.copy())
    oracle.rk4_step(w, p, 0.0, dt, want["vels"])
    assert np.abs(x - w.vnodes[live]).max() < 1e-14


@pytest.mark.parametrize("orders", [None, "mixed"])
def test_smooth_open_lines_vs_oracle(orders):
    """ovg_smooth on rings and lines pinned to the walls (2 to 300 points:
    the kept-band DFT and the batched transform) against the oracle's
    high_freq_cutoff: an open line is found from its smallest slot by its
    reverse links, detrended and odd-extended, filtered and retrended, and
    keeps its end points (src/vortex_smooth.c:316-444)."""
    from oracle import smooth as S
    from test_walls import lines_and_rings
    lengths = [2, 3, 5, 6, 7, 40, 4, 300]
    od = None if orders is None else ["mid", "up", "down", "mid", "up", "mid", "up", "mid"]
    t, p = lines_and_rings(lengths, orders=od)
    rng = np.random.default_rng(8)
    st = t.status[:, 0]
    free = np.where(st == ov.FREE)[0]
    t.vnodes[free] += rng.normal(scale=3e-4, size=(len(free), 3))
    say(f"smoothing with open lines: {len(t.live)} points")
    with ov.Engine(p) as e:
        e.upload(t)
        e.smooth()
        got = e.download()["vnodes"]
    want = S.smooth_tangle(t.vnodes, t.conn, t.status, p.box_lo[:], p.box_hi[:],
                           [1, 1, 0], p.freq_to_cutoff)
    live = t.live
    assert np.abs(got[live] - want[live]).max() < 1e-14
    pinned = (st == ov.PINNED) | (st == ov.PINNED_SLIP)
    assert np.array_equal(got[pinned], t.vnodes[pinned])
    assert np.abs(got[live] - t.vnodes[live]).max() > 1e-5


def test_pinned_velocity_at_matches_oracle(oracle):
    """ovg_velocity_at (the reference's update_velocity with tree == NULL,
    which wall clipping calls) on a tangle with pinned lines: slots on the
    lines, at and next to the pinned ends, and on the rings."""
    from test_walls import pinned_case
    t, p = pinned_case(ov.METHOD_DIRECT, True)
    ends = np.where(t.status[:, 0] != ov.FREE)[0]
    slots = np.unique(np.concatenate([ends, t.conn[ends[::2], 0], t.live[::29]]))
    say(f"velocity_at on pinned lines: {len(slots)} slots")
    tn, nm = oracle.tangents_normals(t, p)
    with ov.Engine(p) as e:
        e.upload(t)
        e.update_tangents_normals()
        vs, vels = e.velocity_at(slots, 0.0)
    for i, k in enumerate(slots):
        ws, wv = oracle.velocity_direct_one(t, p, int(k), 0.0, tn, nm)
        assert np.abs(vs[i] - ws).max() <= 1e-12 * max(1.0, np.abs(ws).max()), k
        assert np.abs(vels[i] - wv).max() <= 1e-12 * max(1.0, np.abs(wv).max()), k


@pytest.mark.parametrize("lengths,orders", [
    ([2, 3, 5, 6, 7, 40, 4], None),
    ([3, 5, 6, 4, 6, 9, 5], ["mid", "up", "mid", "up", "down", "mid", "up"])])
def test_pinned_small_lines_vs_oracle(oracle, lengths, orders):
    """Lines whose lowest slot is their DOWN end, their UP end or mid-line
    (the pass then counts back from its start point before erasing)."""
    from test_walls import lines_and_rings
    t, p = lines_and_rings(lengths, seed=5 if orders is None else 11, orders=orders)
    say(f"small-loop removal with pinned lines: {len(t.live)} points")
    w = ov.Tangle(t.vnodes.copy(), t.conn.copy(), t.status.copy())
    want = oracle.eliminate_small_loops(w, 6)
    with ov.Engine(p) as e:
        e.upload(t)
        got = e.eliminate_small_loops(6)
        conn, st = e.download_connections()
        # the evaluation still runs on what is left
        e.update_tangle(0.0)
        o = e.download()
    assert got == want
    assert np.array_equal(conn, w.conn)
    assert np.array_equal(st[:, 0], w.status[:, 0])
    live = w.live
    ref = oracle.update_tangle(w, p)
    assert rel_l2(o["vels"][live], ref["vels"][live]) < 1e-10


@pytest.mark.parametrize("config,leaf", [(ov.IMAGES_WALL_CANAL, 1), (ov.IMAGES_WALL_CANAL, 16),
                                         (ov.IMAGES_WALL_1_18, 1),
                                         (ov.IMAGES_WALL_PARALLEL_PLANES, 1),
                                         (ov.IMAGES_WALL_PARALLEL_PLANES, 8), ("pinned", 1)])
def test_wall_list_truncation_vs_exact_sum(config, leaf):
    """theta 0.4 with mirror images: the list traversal's acceptance of a
    mirrored copy (its reflected centre against the group box) and its
    reflected multi-segment monopoles, against the exact reflected sum.  The
    exact sum is the reference-acceptance point traversal at theta 1e-6
    (every source a single segment; equal to the oracle's reflected copies,
    tests/test_walls.py).  The direct path is no reference here: it adds
    mirror images unreflected, as the reference's direct sum does.  The list
    error must not exceed the point traversal's at theta 0.4: its group test
    never accepts a cell that the per-point test opens, so a wrong
    reflection or a misplaced mirrored centre shows as a larger error."""
    from test_walls import pinned_case
    if config == "pinned":
        t, p = pinned_case(ov.METHOD_BH, True)
    else:
        t, p = wall_case(config, ov.METHOD_BH, n=24)
    live = t.live
    say(f"walls {config}, leaf {leaf}: list and point at theta 0.4 vs the exact sum, "
        f"{len(live)} points")

    def run(trav, theta, leaf_size=1, prec=ov.PREC_FP64):
        q = ov.Params.from_buffer_copy(p)
        q.traverse, q.theta, q.leaf_size, q.precision = trav, theta, leaf_size, prec
        with ov.Engine(q) as e:
            e.upload(t)
            e.update_tangle(0.0)
            s = e.stats()
            return e.download()["vs"][live], s

    exact, _ = run(ov.TRAVERSE_POINT, 1e-6)
    point, sp = run(ov.TRAVERSE_POINT, 0.4)
    lst, sl = run(ov.TRAVERSE_LIST, 0.4, leaf)
    mixed, _ = run(ov.TRAVERSE_LIST, 0.4, leaf, ov.PREC_MIXED)
    e_point, e_list, e_mixed = rel_l2(point, exact), rel_l2(lst, exact), rel_l2(mixed, exact)
    say(f"vs exact: point {e_point:.3e} ({sp.interactions:.0f} pairs), list {e_list:.3e} "
        f"({sl.interactions:.0f} pairs), list mixed {e_mixed:.3e}")
    # the tangle is large enough that theta 0.4 accepts cells
    assert e_point > 1e-6
    assert e_list <= e_point
    assert e_mixed <= e_point


@pytest.mark.parametrize("pin", [ov.PINNED, ov.PINNED_SLIP])
def test_wall_clipping_vs_reference(ref, pin):
    """ovg_enforce_walls against the compiled reference's
    enforce_wall_boundaries (src/tangle.c:966-1016) on rings crossing and
    approaching the two walls of the parallel-planes geometry: the same
    clips (new pinned nodes in the same slots, through a slot doubling), the
    same removals, links, statuses and positions exactly, the velocities the
    pass leaves at the points it tested within 1e-10, and the evaluation of
    the clipped tangle within 1e-10 of the reference's."""
    from test_walls import rings_at_walls
    t, p = rings_at_walls()
    p.pin_mode = pin
    say(f"wall clipping, pin mode {pin}: {len(t.live)} points in {t.n} slots")
    r = ref.tangle(t, p, wall_images=ov.IMAGES_WALL_PARALLEL_PLANES)
    r.enforce_walls(0.0, pin)
    want, want_st = r.get(), r.status()
    with ov.Engine(p) as e:
        e.upload(t)
        clipped, removed = e.enforce_walls(0.0)
        assert e.n == r.n
        got = e.download()
        conn, st = e.download_connections()
        e.update_tangle(0.0)
        ev = e.download()
    say(f"clipped {clipped}, removed {removed}, slots {e.n}")
    assert clipped > 0 and removed > 0
    assert np.array_equal(conn, want["conn"])
    assert np.array_equal(st, want_st)
    live = st[:, 0] != ov.EMPTY
    assert np.array_equal(got["vnodes"][live], want["vnodes"][live])
    # the state the pass wrote, at the points that existed before it (the
    # reference's new slots come from realloc and hold no values)
    old = np.zeros(len(live), dtype=bool)
    old[t.live] = True
    old &= live
    for key in ("vels", "vs", "tangents", "normals"):
        scale = np.abs(want[key][old]).max()
        assert scale > 0, key
        assert np.abs(got[key][old] - want[key][old]).max() <= 1e-10 * scale, key
    r.update_tangle(0.0)
    w2 = r.get()
    # two-node lines clipped onto one wall point have no stencil (0/0) in
    # the reference as here: the same NaN rows
    for key in ("vels", "tangents"):
        fin = np.isfinite(w2[key][live]).all(1)
        assert np.array_equal(fin, np.isfinite(ev[key][live]).all(1)), key
        assert rel_l2(ev[key][live][fin], w2[key][live][fin]) < 1e-10, key


@pytest.mark.parametrize("case", ["lines", "clipped"])
def test_remesh_pinned_vs_reference(ref, case):
    """ovg_remesh on tangles with lines pinned to the walls against the
    compiled reference's remesh (src/tangle.c:1018-1061): lines from wall to
    wall whose segments are split (and the new points split again later in
    the same pass), and a tangle just clipped to the walls, whose short wall
    segments are removed and whose insertions are followed by the wall
    clipping remesh runs after them.  Links, statuses and positions exact."""
    from test_walls import lines_and_rings, rings_at_walls
    if case == "lines":
        t, p = lines_and_rings([2, 3, 5, 6, 7, 40, 4])
        bounds = (1.5e-3, 2.4e-3)
    else:
        t, p = rings_at_walls()
        bounds = (1.0e-3, 2.03e-3)
    r = ref.tangle(t, p, wall_images=ov.IMAGES_WALL_PARALLEL_PLANES)
    with ov.Engine(p) as e:
        e.upload(t)
        if case == "clipped":
            r.enforce_walls(0.0, ov.PINNED)
            e.enforce_walls(0.0)
        added, removed = e.remesh(*bounds)
        conn, st = e.download_connections()
        got = e.download()
    r.remesh(*bounds)
    want, want_st = r.get(), r.status()
    say(f"remesh {case}: {added} added, {removed} removed, {e.n} slots")
    assert added > 0
    assert e.n == r.n
    assert np.array_equal(conn, want["conn"])
    assert np.array_equal(st, want_st)
    live = st[:, 0] != ov.EMPTY
    assert np.array_equal(got["vnodes"][live], want["vnodes"][live])


def test_reconnect_pinned_vs_reference(ref):
    """ovg_reconnect on a tangle just clipped to the walls (open lines ending
    on pinned nodes, rings crossing them) against the compiled reference's
    reconnect (src/vortex_dynamics.c:186-288) with its direct-sum velocities:
    pinned nodes never reconnect, stencils next to a pinned end use
    step_node's mirrored ghosts; the same rewirings, links exact."""
    from test_walls import rings_at_walls
    t, p = rings_at_walls(seed=9, n_cross=30, n_inside=20)
    p.rec_dist = 0.003
    p.rec_velocity = ov.RECVEL_DIRECT
    r = ref.tangle(t, p, wall_images=ov.IMAGES_WALL_PARALLEL_PLANES)
    r.enforce_walls(0.0, ov.PINNED)
    with ov.Engine(p) as e:
        e.upload(t)
        e.enforce_walls(0.0)
        cand = e.reconnect_candidates()
        n = e.reconnect(0.0)
        conn, st = e.download_connections()
    want_n = r.reconnect(0.0)
    want = r.get()
    say(f"pinned reconnection: {len(cand)} candidate pairs, {n} reconnections "
        f"(reference {want_n})")
    assert n == want_n and n > 0
    assert np.array_equal(conn, want["conn"])
    assert np.array_equal(st, r.status())
    free = st[:, 0] == ov.FREE
    assert all(free[k] and free[l] for k, l in cand)


def test_inject_line_pairs_vs_reference(ref):
    """The reference's line-pair injection (inject_line_pairs,
    src/vortex_injection.c:760-800) through ovg_inject_line_pairs_draw and
    ovg_add_line_z on a resident tangle with holes (the slots double on the
    way): slots, links, statuses and positions bit for bit; the tangle with
    its new wall-to-wall lines then evaluates as the reference's does."""
    from test_walls import rings_at_walls
    t, p = rings_at_walls(spare=40)
    dl_min, dl_max, freq, seed, n_pairs = 5e-4, 1.5e-3, 1000.0, 4242, 2
    times = np.arange(1, 6) * 4e-4 + 5.0  # after any earlier call's static
    r = ref.tangle(t, p, wall_images=ov.IMAGES_WALL_PARALLEL_PLANES)
    r.inject_line_pairs(seed, times, freq, n_pairs, 0, dl_min, dl_max)
    rng, last = ov.rand48_state(seed), 0.0
    added = 0
    with ov.Engine(p) as e:
        e.upload(t)
        for tm in times:
            due, xy, last = ov.inject_line_pairs_draw(tm, freq, n_pairs, 0, p.box_lo[:],
                                                      p.box_hi[:], rng, last)
            if due:
                for x1, y1, x2, y2 in xy:
                    added += e.add_line_z(x1, y1, +1, dl_min, dl_max)
                    added += e.add_line_z(x2, y2, -1, dl_min, dl_max)
        conn, st = e.download_connections()
        got = e.download()["vnodes"]
        e.update_tangle(0.0)
        ev = e.download()
    say(f"line injection: {added} points in {e.n} slots")
    assert e.n == r.n and added > 0
    want = r.get()
    assert np.array_equal(conn, want["conn"])
    assert np.array_equal(st, r.status())
    live = st[:, 0] != ov.EMPTY
    assert np.array_equal(got[live], want["vnodes"][live])
    r.update_tangle(0.0)
    w2 = r.get()
    assert rel_l2(ev["vels"][live], w2["vels"][live]) < 1e-10


def test_wall_driver_steps_vs_reference(ref):
    """Three steps of the reference driver's order with walls
    (src/main.c:125-138 without injection and smoothing, whose FFTW is
    absent): enforce_wall_boundaries, small loops, reconnect, small loops,
    wrap, remesh, small loops, update_tangle, rk4_step, wrap; direct FP64
    sums as the reference's update_velocity with tree == NULL.  The same
    topology after every step and positions within 1e-10 cm."""
    from test_walls import rings_at_walls
    t, p = rings_at_walls(seed=9, n_cross=12, n_inside=6)
    p.rec_dist = 0.003
    p.rec_velocity = ov.RECVEL_DIRECT
    dt, bounds = 1e-4, (1.0e-3, 2.6e-3)
    r = ref.tangle(t, p, wall_images=ov.IMAGES_WALL_PARALLEL_PLANES)
    stats = []
    with ov.Engine(p) as e:
        e.upload(t)
        tm = 0.0
        for s in range(3):
            clip = e.enforce_walls(0.0)
            e.eliminate_small_loops(5)
            nrec = e.reconnect(0.0)
            e.eliminate_small_loops(5)
            e.enforce_periodic()
            rem = e.remesh(*bounds)
            e.eliminate_small_loops(5)
            e.update_tangle(tm)
            e.rk4_step(tm, dt)
            e.enforce_periodic()
            r.enforce_walls(0.0)
            r.eliminate_small_loops(5)
            want_rec = r.reconnect(0.0)
            r.eliminate_small_loops(5)
            r.enforce_periodic()
            r.remesh(*bounds)
            r.eliminate_small_loops(5)
            r.update_tangle(tm)
            r.rk4_step(tm, dt)
            r.enforce_periodic()
            tm += dt
            conn, st = e.download_connections()
            got = e.download()["vnodes"]
            want = r.get()
            stats.append((clip, nrec, rem))
            assert nrec == want_rec, s
            assert e.n == r.n, s
            assert np.array_equal(conn, want["conn"]), s
            assert np.array_equal(st, r.status()), s
            live = st[:, 0] != ov.EMPTY
            assert np.abs(got[live] - want["vnodes"][live]).max() < 1e-10, s
    say(f"wall driver steps (clipped/removed, reconnections, remesh added/removed): {stats}")
