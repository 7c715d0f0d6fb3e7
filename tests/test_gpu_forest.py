"""Forests (ovg_upload_forest): many tangles evaluated and stepped in one
context, one tree build and one list-traversal launch per evaluation.  Each
tangle's velocities and positions must be bit-identical to the same tangle
run alone (same tree cells, groups, walk and sums: its keys only gain the
tangle index above the Morton bits, and its fixed-point units are its
own).  HIP graphs of a step replay bit-identically to eager steps; each
replay is checked, with eager work and synchronisation between replays."""
import sys

import numpy as np
import pytest

import openvort_amd as ov
from openvort_amd import configs

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


# (target points, D, dxi, alpha pair, v_ns): uneven sizes, two boxes and
# spacings (so two group levels and FP64 zones), both temperatures
MEMBERS = [(2000, 0.1, 8e-4, configs.T16, 0.5), (9000, 0.1, 8e-4, configs.T19, 1.2),
           (30000, 0.1, 8e-4, configs.T16, 2.0), (4000, 0.15, 6e-4, configs.T19, 0.8),
           (15000, 0.15, 6e-4, configs.T16, 1.5), (600, 0.1, 8e-4, configs.T19, 1.0)]


def members(precision=ov.PREC_MIXED, seed=0xF0E57):
    out = []
    for j, (pts, D, dxi, ap, vns) in enumerate(MEMBERS):
        s = seed * 100 + j
        n_rings = ov.rings_for_points(s, pts, D, 0.005, 0.03, dxi)
        t = ov.generate_rings(s, n_rings, D, 0.005, 0.03, dxi)
        p = configs.params(D, dxi, ap, vns, precision=precision, leaf_size=1,
                           traverse=ov.TRAVERSE_LIST)
        p.points_per_wave = 16
        out.append((t, p))
    return out


def step(e, tm, dt):
    e.update_tangle(tm)
    e.rk4_step(tm, dt)
    e.enforce_periodic()


def alone(t, p, n_steps, dt, first=False):
    with ov.Engine(p) as e:
        e.upload(t)
        outs = []
        tm = 0.0
        if first:
            e.update_tangle(0.0)
            outs.append(e.download())
        for _ in range(n_steps):
            step(e, tm, dt)
            tm += dt
            outs.append(e.download())
        return outs


@pytest.mark.parametrize("precision", [ov.PREC_MIXED, ov.PREC_FP64])
def test_forest_equals_alone(precision):
    ms = members(precision)
    dt = 5e-5
    say(f"forest of {len(ms)} tangles ({sum(len(t.live) for t, _ in ms)} points), "
        f"precision {precision}: update_tangle, two steps, against each tangle alone")
    with ov.Engine(ms[0][1]) as e:
        off = e.upload_forest([t for t, _ in ms], [p for _, p in ms])
        assert e.forest_size() == len(ms)
        e.update_tangle(0.0)
        got = [e.download()]
        tm = 0.0
        for _ in range(2):
            step(e, tm, dt)
            tm += dt
            got.append(e.download())
        st = e.stats()
    pairs = 0.0
    for j, (t, p) in enumerate(ms):
        want = alone(t, p, 2, dt, first=True)
        sl = slice(off[j], off[j + 1])
        live = t.live
        for g, w in zip(got, want):
            for key in ("vnodes", "vels", "vs", "tangents", "normals"):
                assert np.array_equal(g[key][sl][live], w[key][live]), (j, key)
    assert st.interactions > 0
    assert (st.interactions_fp32 > 0.5 * st.interactions) == (precision == ov.PREC_MIXED)


def test_forest_graph_replay_equals_eager():
    ms = members()
    dt = 5e-5
    tangles, params = [t for t, _ in ms], [p for _, p in ms]
    say("forest: a captured step replayed 4 times against eager steps")
    with ov.Engine(params[0]) as a, ov.Engine(params[0]) as b:
        a.upload_forest(tangles, params)
        b.upload_forest(tangles, params)
        tm = 0.0
        step(a, tm, dt)
        step(b, tm, dt)  # every buffer at its size before the capture
        tm += dt
        b.graph_begin()
        step(b, tm, dt)
        nodes = b.graph_end()
        assert nodes > 10
        # the capture queued nothing: b is still one step in
        assert np.array_equal(a.download()["vnodes"], b.download()["vnodes"])
        for r in range(4):
            step(a, tm, dt)
            tm += dt
            b.graph_launch(1)
            b.synchronize()
            ga, gb = a.download(), b.download()
            for key in ("vnodes", "vels", "vs"):
                assert np.array_equal(ga[key], gb[key]), (r, key)
            b.line_length()  # eager work on b between replays
        b.graph_release()


def test_single_tangle_graph_replay_equals_eager():
    t, p, dt = configs.make("c2", rings=80, precision=ov.PREC_MIXED,
                            traverse=ov.TRAVERSE_LIST, leaf_size=1)
    say(f"one tangle ({len(t.live)} points): a captured step replayed 3 times")
    with ov.Engine(p) as a, ov.Engine(p) as b:
        a.upload(t)
        b.upload(t)
        step(a, 0.0, dt)
        step(b, 0.0, dt)
        b.graph_begin()
        step(b, dt, dt)
        b.graph_end()
        for r in range(3):
            step(a, dt, dt)
            b.graph_launch(1)
            b.synchronize()
            assert np.array_equal(a.download()["vnodes"], b.download()["vnodes"]), r
            assert np.array_equal(a.download()["vels"], b.download()["vels"]), r


def test_graph_capture_refuses_buffer_growth():
    t, p, dt = configs.make("c2", rings=20, precision=ov.PREC_MIXED,
                            traverse=ov.TRAVERSE_LIST, leaf_size=1)
    say("graph capture: a buffer that would grow inside it is refused")
    with ov.Engine(p) as e:
        e.upload(t)
        e.graph_begin()
        with pytest.raises(ov.OvgError) as ei:
            e.update_tangle(0.0)  # the tree buffers do not exist yet
        assert ei.value.code == ov.EINVAL
        with pytest.raises(ov.OvgError) as ei:
            e.graph_end()
        assert ei.value.code == ov.EINVAL
        with pytest.raises(ov.OvgError):
            e.graph_launch(1)  # nothing was instantiated
        e.update_tangle(0.0)  # eager again: works
        o = e.download()
    assert np.isfinite(o["vels"][t.live]).all()


def test_forest_refuses_single_tangle_calls():
    ms = members()[:2]
    say("forest: topology passes and direct sums refused")
    with ov.Engine(ms[0][1]) as e:
        e.upload_forest([t for t, _ in ms], [p for _, p in ms])
        for call in (lambda: e.reconnect(0.0), lambda: e.remesh(4e-4, 1.2e-3), e.smooth,
                     lambda: e.eliminate_small_loops(5), lambda: e.velocity_at([0, 1], 0.0),
                     lambda: e.set_params(ms[1][1])):
            with pytest.raises(ov.OvgError) as ei:
                call()
            assert ei.value.code == ov.EUNSUPPORTED
        # a plain upload makes it one tangle again
        e.upload(ms[1][0])
        assert e.forest_size() == 1
        e.update_tangle(0.0)


def test_forest_autotuned_equals_alone():
    ms = members()
    say("forest: autotune, then the tuned setting against each tangle alone")
    with ov.Engine(ms[0][1]) as e:
        off = e.upload_forest([t for t, _ in ms], [p for _, p in ms])
        rows, best = e.autotune(repeats=1)
        q = e.get_params()
        e.update_tangle(0.0)
        got = e.download()
    say(f"forest autotune pick: {rows[best]}")
    for j, (t, p) in enumerate(ms):
        pj = ov.Params.from_buffer_copy(p)
        for f in ("points_per_wave", "leaf_size", "group_level", "list_stack", "list_tile",
                  "list_blocks"):
            setattr(pj, f, getattr(q, f))
        with ov.Engine(pj) as a:
            a.upload(t)
            a.update_tangle(0.0)
            w = a.download()
        sl = slice(off[j], off[j + 1])
        assert np.array_equal(got["vels"][sl][t.live], w["vels"][t.live]), j
