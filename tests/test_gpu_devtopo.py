"""Device-resident topology (ovg_params.topology = TOPO_DEVICE): reconnection,
small-loop removal and remesh as device kernels against the CPU oracle, the
reference's golden run and the host passes.

Every comparison is exact: the device passes keep the reference's visiting
order and the host passes' FP64 operation order, so links, status,
positions and counts must be identical.  Tangles with more free slots on
the device (it grows its slots before it could run out) are compared on the
slots both have, the rest EMPTY.
"""
import sys

import numpy as np
import pytest

import openvort_amd as ov
from openvort_amd import configs
from conftest import load_golden

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def params_from(blob):
    p = ov.Params()
    raw = blob.tobytes()
    ov.C.memmove(ov.C.byref(p), raw, min(len(raw), ov.C.sizeof(p)))
    return p


def engine(t, p, topology):
    q = ov.Params()
    ov.C.memmove(ov.C.byref(q), ov.C.byref(p), ov.C.sizeof(p))
    q.topology = topology
    e = ov.Engine(q)
    e.upload(t)
    return e


def same_state(e, want_x, want_conn, want_st, tag):
    """Engine state equals (x, conn, status) on the common slots; slots the
    engine has beyond them are EMPTY."""
    o = e.download()
    conn, st = e.download_connections()
    n = min(len(st), len(want_st))
    assert np.array_equal(conn[:n], want_conn[:n]), tag
    assert np.array_equal(st[:n], want_st[:n]), tag
    assert (st[n:, 0] == -1).all() and (want_st[n:, 0] == -1).all(), tag
    lv = np.nonzero(st[:n, 0] != -1)[0]
    assert np.array_equal(o["vnodes"][lv], want_x[lv]), tag


def jittered(rings, seed, scale=2e-4):
    t, p, _ = configs.make("c2", rings=rings)
    rng = np.random.default_rng(seed)
    live = t.live
    t.vnodes[live] = (t.vnodes[live] + rng.normal(scale=scale, size=(len(live), 3))) % 0.1
    return t, p


def test_devtopo_sequence_vs_oracle(oracle):
    """remesh, small loops and reconnection (direct-sum velocities) in
    sequence on the device against the oracle's sequential passes."""
    t, p = jittered(60, 31)
    p.method = ov.METHOD_DIRECT
    p.rec_dist = 4e-4
    say(f"device topology sequence on {len(t.live)} jittered points")
    want, ne = t.copy(), 0
    with engine(t, p, ov.TOPO_DEVICE) as e:
        for step, (lo, hi) in enumerate(((4e-4, 1.2e-3), (5e-4, 1.0e-3))):
            want, added, removed, ne = oracle.remesh(want, p, lo, hi, next_empty=ne)
            a, r = e.remesh(lo, hi)
            assert (a, r) == (added, removed) and added > 0 and removed > 0
            same_state(e, want.vnodes, want.conn, want.status, f"remesh {step}")
            n_small = oracle.eliminate_small_loops(want, 5)
            assert e.eliminate_small_loops(5) == n_small
            same_state(e, want.vnodes, want.conn, want.status, f"small loops {step}")
            n_rec = oracle.reconnect(want, p)
            assert e.reconnect(0.0) == n_rec
            same_state(e, want.vnodes, want.conn, want.status, f"reconnect {step}")
            n_small = oracle.eliminate_small_loops(want, 5)
            assert e.eliminate_small_loops(5) == n_small
            same_state(e, want.vnodes, want.conn, want.status, f"small loops 2 {step}")


def test_devtopo_c1_100_steps_golden():
    """The reference's 100-step run of the 16-ring tangle with the driver's
    passes on the device, queued without count pointers: live points,
    reconnections, final links and positions as the compiled reference."""
    g = load_golden("c1_100steps.npz")
    t, _, _ = configs.make(str(g["config"]), rings=int(g["rings"]))
    p = params_from(g["params"])
    dt, lo, hi = float(g["dt"]), float(g["dl_min"]), float(g["dl_max"])
    say("c1 direct FP64, 100 steps, device topology")
    tm, recs = 0.0, 0
    with engine(t, p, ov.TOPO_DEVICE) as e:
        for step in range(100):
            recs += e.reconnect(tm)
            e.eliminate_small_loops(5, counts=False)
            e.enforce_periodic()
            e.remesh(lo, hi, counts=False)
            e.eliminate_small_loops(5, counts=False)
            e.update_tangle(tm)
            e.rk4_step(tm, dt)
            e.enforce_periodic()
            tm += dt
            if (step + 1) % 10 == 0:
                i = (step + 1) // 10 - 1
                assert e.num_points() == int(g["points"][i]), step
                assert recs == int(g["reconnections"][i]), step
        x = e.download()["vnodes"]
        conn, st = e.download_connections()
    n = len(g["status_final"])
    assert np.array_equal(conn[:n], g["conn_final"]) and np.array_equal(st[:n], g["status_final"])
    assert (st[n:, 0] == -1).all()
    lv = np.nonzero(st[:n, 0] != -1)[0]
    assert np.abs(x[lv] - g["x_final"][lv]).max() < 1e-9


def test_devtopo_remesh_capacity(oracle):
    """A tangle with no free slot: the device grows its slots before the
    pass; every insertion lands in the slot the reference's doubling gives."""
    seed, _, D, rmin, rmax, dxi = configs.CONFIGS["c1"][:6]
    t0, p, _ = configs.make("c1", rings=40)
    t = ov.generate_rings(seed, 40, D, rmin, rmax, dxi, n_cap=len(t0.live))
    say("device remesh on a full tangle")
    want, added, removed, _ = oracle.remesh(t, p, 0.4e-3, 0.9e-3)
    with engine(t, p, ov.TOPO_DEVICE) as e:
        a, r = e.remesh(0.4e-3, 0.9e-3)
        assert (a, r) == (added, removed) and a > 0 and e.n >= want.n
        same_state(e, want.vnodes, want.conn, want.status, "remesh on full tangle")


@pytest.fixture(scope="module")
def c2_list():
    t, p, dt = configs.make("c2", precision=ov.PREC_MIXED, traverse=ov.TRAVERSE_LIST)
    p.points_per_wave = 16
    return t, p, dt


@pytest.mark.parametrize("rec_velocity", [ov.RECVEL_TREE, ov.RECVEL_DIRECT])
def test_devtopo_reconnect_equals_host_c2(c2_list, rec_velocity):
    """One reconnection pass on the full 200-ring tangle: device pass and
    host pass give the same rewirings, with the tree's (list traversal)
    endpoint velocities and with the FP64 direct sum."""
    t, p, _ = c2_list
    p.rec_velocity = rec_velocity
    say(f"c2 reconnect, {len(t.live)} points, rec_velocity {rec_velocity}")
    with engine(t, p, ov.TOPO_HOST) as h, engine(t, p, ov.TOPO_DEVICE) as d:
        nh, nd = h.reconnect(0.0), d.reconnect(0.0)
        say(f"  reconnections host {nh} device {nd}")
        assert nh == nd and nh > 0
        x = h.download()["vnodes"]
        conn, st = h.download_connections()
        same_state(d, x, conn, st, "c2 reconnect")


def test_devtopo_driver_steps_equal_host(c2_list):
    """Five steps of the reference driver's order on a jittered 100-ring
    tangle, list traversal in mixed precision, tree reconnection velocities,
    smoothing every other step: device topology and host topology reach the
    same links, status and positions bit for bit."""
    _, p, dt = c2_list
    t, _ = jittered(100, 7, scale=1.5e-4)
    p.rec_velocity = ov.RECVEL_TREE
    lo, hi = 4.5e-4, 1.0e-3
    say(f"driver steps host vs device topology, {len(t.live)} points")
    counts = {}
    engines = {}
    for mode in (ov.TOPO_HOST, ov.TOPO_DEVICE):
        e = engine(t, p, mode)
        tm, tot = 0.0, np.zeros(4, dtype=np.int64)
        for step in range(5):
            tot[0] += e.reconnect(tm)
            tot[1] += e.eliminate_small_loops(5)
            if step % 2 == 0:
                e.smooth()
            e.enforce_periodic()
            a, r = e.remesh(lo, hi)
            tot[2] += a
            tot[3] += r
            e.eliminate_small_loops(5)
            e.update_tangle(tm)
            e.rk4_step(tm, dt)
            e.enforce_periodic()
            tm += dt
        counts[mode] = tot
        engines[mode] = e
    try:
        say(f"  counts (rec, loops, added, removed) host {counts[0]} device {counts[1]}")
        assert np.array_equal(counts[0], counts[1]) and counts[0][0] > 0 and counts[0][2] > 0
        h = engines[ov.TOPO_HOST]
        x = h.download()["vnodes"]
        conn, st = h.download_connections()
        same_state(engines[ov.TOPO_DEVICE], x, conn, st, "driver steps")
        assert h.num_points() == engines[ov.TOPO_DEVICE].num_points()
    finally:
        for e in engines.values():
            e.close()


def test_devtopo_queued_counts_and_mirrors():
    """Passes queued without count pointers: the point count, the host-side
    calls (download, save-format arrays, add_loop) and a later host-mode
    pass all see the device's edits."""
    t, p = jittered(40, 3)
    p.method = ov.METHOD_DIRECT
    lo, hi = 4e-4, 1.0e-3
    with engine(t, p, ov.TOPO_HOST) as h, engine(t, p, ov.TOPO_DEVICE) as d:
        h.remesh(lo, hi)
        h.eliminate_small_loops(5)
        d.remesh(lo, hi, counts=False)
        d.eliminate_small_loops(5, counts=False)
        assert d.num_points() == h.num_points()
        for e in (h, d):
            e.add_loop((0.05, 0.05, 0.05), (0.0, 0.0, 1.0), 0.01, lo, hi)
        x = h.download()["vnodes"]
        conn, st = h.download_connections()
        same_state(d, x, conn, st, "after add_loop")
        # back to host passes on the device engine's state
        q = d.get_params()
        q.topology = ov.TOPO_HOST
        d.set_params(q)
        assert d.reconnect(0.0) == h.reconnect(0.0)
        conn, st = h.download_connections()
        same_state(d, h.download()["vnodes"], conn, st, "host pass after device passes")


@pytest.mark.parametrize("rec_velocity", [ov.RECVEL_TREE, ov.RECVEL_DIRECT])
def test_devtopo_releases_its_buffers(rec_velocity):
    """A context that ran the device passes (both endpoint-velocity modes)
    gives back every device byte when it is destroyed: a serving loop
    creates and destroys one per job."""
    t, p = jittered(30, 5)
    p.method = ov.METHOD_BH
    p.precision = ov.PREC_MIXED
    p.traverse = ov.TRAVERSE_LIST
    p.rec_velocity = rec_velocity
    before = ov.device_memory_in_use()
    with engine(t, p, ov.TOPO_DEVICE) as e:
        for step in range(2):
            e.reconnect(0.0)
            e.eliminate_small_loops(5)
            e.smooth()
            e.remesh(4e-4, 1.0e-3)
            e.update_tangle(0.0)
        assert ov.device_memory_in_use() > before
    say(f"device bytes in use before {before}, after {ov.device_memory_in_use()}")
    assert ov.device_memory_in_use() == before


def test_devtopo_remesh_chains_of_splits(oracle):
    """Coarse rings (spacing 3e-3 cm) remeshed to max_dist 1e-3 cm with no
    free slot: a segment takes a chain of insertions in one pass (a new point
    whose slot lies ahead of the cursor is visited and split again), so the
    clusters hand the pass to the sequential kernel.  The device reads its
    insertion bound, grows its slots once, and places every point in the
    slot the reference's doublings give."""
    seed, _, D, rmin, rmax = configs.CONFIGS["c1"][:5]
    t0 = ov.generate_rings(seed, 10, D, rmin, rmax, 3e-3)
    t = ov.generate_rings(seed, 10, D, rmin, rmax, 3e-3, n_cap=len(t0.live))
    _, p, _ = configs.make("c1", rings=10)
    say(f"device remesh of {len(t.live)} coarse points with no free slot")
    want, added, removed, _ = oracle.remesh(t, p, 0.3e-3, 1.0e-3)
    with engine(t, p, ov.TOPO_DEVICE) as e:
        a, r = e.remesh(0.3e-3, 1.0e-3)
        say(f"  added {a} (oracle {added}), removed {r} (oracle {removed}), slots {e.n}")
        assert (a, r) == (added, removed) and a > len(t.live)
        assert e.n >= want.n
        same_state(e, want.vnodes, want.conn, want.status, "chains of splits")
