"""Mirror walls (images with reflections): the engine's tables against the
compiled reference's, the oracle against the compiled reference bit for
bit, and the image construction against what the mathematics fixes (a
mirror image is the reflected tangle with reversed circulation, so with
theta -> 0 it equals an explicit reflected and reversed copy)."""
import numpy as np
import pytest

import openvort_amd as ov
from openvort_amd import configs
from conftest import rel_l2

P_, M_, O_ = ov.WALL_PERIODIC, ov.WALL_MIRROR, ov.WALL_OPEN
# walls of each table's geometry (src/initial_conditions.c:135-165 and the
# table comments of src/boundary_images.c:75-230), faces LEFT..UP
WALLS = {
    ov.IMAGES_WALL_1_OPEN: (O_, O_, O_, O_, M_, O_),
    ov.IMAGES_WALL_1_6: (P_, P_, P_, P_, M_, P_),
    ov.IMAGES_WALL_1_18: (P_, P_, P_, P_, M_, P_),
    ov.IMAGES_WALL_1_26: (P_, P_, P_, P_, M_, P_),
    ov.IMAGES_WALL_2_4: (P_, P_, P_, P_, M_, M_),
    ov.IMAGES_WALL_2_2: (P_, P_, M_, M_, M_, M_),
    ov.IMAGES_WALL_PARALLEL_PLANES: (P_, P_, P_, P_, M_, M_),
    ov.IMAGES_WALL_CANAL: (P_, P_, M_, M_, M_, M_),
}


def rings_inside(seed, n, D, R, dxi):
    """Rings wholly inside the box [0, D]^3."""
    rng = np.random.default_rng(seed)
    xs, conns = [], []
    off = 0
    for _ in range(n):
        c = rng.uniform(R * 1.1, D - R * 1.1, 3)
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        u = np.cross(nrm, [1.0, 0, 0] if abs(nrm[0]) < 0.9 else [0, 1.0, 0])
        u /= np.linalg.norm(u)
        v = np.cross(nrm, u)
        m = int(2 * np.pi * R / dxi)
        ph = 2 * np.pi * np.arange(m) / m
        xs.append(c + R * (np.cos(ph)[:, None] * u + np.sin(ph)[:, None] * v))
        conns.append(np.stack([off + (np.arange(m) + 1) % m, off + (np.arange(m) - 1) % m], 1))
        off += m
    x = np.concatenate(xs)
    st = np.zeros((off, 2), dtype=np.int32)
    st[:, 1] = -1
    return ov.Tangle(x, np.concatenate(conns).astype(np.int32), st)


def rings_at_walls(seed=3, n_cross=6, n_inside=3, D=0.1, R=0.012, dxi=2e-3, spare=16):
    """Rings inside the box plus rings whose centres lie within 0.9 R of the
    DOWN or UP plane (so they cross it or come within rec_dist of it), and
    `spare` EMPTY slots after them (the allocator's first choices)."""
    rng = np.random.default_rng(seed)
    t = rings_inside(seed, n_inside, D, R, dxi)
    xs, conns = [t.vnodes], [t.conn]
    off = t.n
    for i in range(n_cross):
        z = rng.uniform(-0.3 * R, 0.9 * R)
        c = np.array([rng.uniform(0, D), rng.uniform(0, D), z if i % 2 == 0 else D - z])
        nrm = rng.normal(size=3)
        nrm[2] *= 0.3  # mostly upright rings, which cut the plane
        nrm /= np.linalg.norm(nrm)
        u = np.cross(nrm, [1.0, 0, 0] if abs(nrm[0]) < 0.9 else [0, 1.0, 0])
        u /= np.linalg.norm(u)
        v = np.cross(nrm, u)
        m = int(2 * np.pi * R / dxi)
        ph = 2 * np.pi * np.arange(m) / m
        xs.append(c + R * (np.cos(ph)[:, None] * u + np.sin(ph)[:, None] * v))
        conns.append(np.stack([off + (np.arange(m) + 1) % m, off + (np.arange(m) - 1) % m], 1))
        off += m
    xs.append(np.zeros((spare, 3)))
    conns.append(np.full((spare, 2), -1))
    st = np.zeros((off + spare, 2), dtype=np.int32)
    st[:, 1] = -1
    st[off:, 0] = ov.EMPTY
    t = ov.Tangle(np.concatenate(xs), np.concatenate(conns).astype(np.int32), st)
    p = configs.params(D, dxi, configs.T16, 0.5, method=ov.METHOD_DIRECT)
    for f, w in enumerate(WALLS[ov.IMAGES_WALL_PARALLEL_PLANES]):
        p.wall[f] = w
    ov.set_wall_images(p, ov.IMAGES_WALL_PARALLEL_PLANES)
    return t, p


def test_reference_wall_clipping_scenario(ref):
    """The scenario of the engine's wall-clipping parity test clips and
    removes in the compiled reference (enforce_wall_boundaries,
    src/tangle.c:966-1016), for both pin modes: new nodes pinned to the DOWN
    and UP walls, open lines ending on them, nothing left behind a wall."""
    t, p = rings_at_walls()
    for pin in (ov.PINNED, ov.PINNED_SLIP):
        r = ref.tangle(t, p, wall_images=ov.IMAGES_WALL_PARALLEL_PLANES)
        r.enforce_walls(0.0, pin)
        st, g = r.status(), r.get()
        live = st[:, 0] != ov.EMPTY
        pinned = st[:, 0] == pin
        assert pinned.sum() >= 4 and pinned.sum() % 2 == 0
        assert set(st[pinned, 1]) <= {ov.FACE_DOWN, ov.FACE_UP}
        z = g["vnodes"][live, 2]
        assert (z >= 0).all() and (z <= p.box_hi[2]).all()
        # every pinned node ends a line: exactly one of its links is -1
        c = g["conn"][pinned]
        assert ((c[:, 0] < 0) ^ (c[:, 1] < 0)).all()


def wall_case(config, method, seed=7, n=6, theta=0.4):
    D, dxi = 0.1, 2e-3
    t = rings_inside(seed, n, D, 0.012, dxi)
    p = configs.params(D, dxi, configs.T16, 0.5, method=method, theta=theta)
    for f, w in enumerate(WALLS[config]):
        p.wall[f] = w
    ov.set_wall_images(p, config)
    return t, p


@pytest.mark.parametrize("config", sorted(WALLS))
def test_wall_tables_are_the_reference_tables(ref, config):
    p = ov.Params()
    ov.set_wall_images(p, config)
    shifts, masks = ref.wall_table(config)
    n = p.n_images
    assert n == len(shifts)
    got = np.array([[p.images[j][k] for k in range(3)] for j in range(n)])
    assert np.array_equal(got, shifts)
    assert np.array_equal(np.array(p.image_reflect[:n]), masks)
    assert all(p.image_reflect[j] == 0 for j in range(n, 26))


def test_periodic_table_clears_reflections():
    p = ov.Params()
    ov.set_wall_images(p, ov.IMAGES_WALL_CANAL)
    ov.set_images(p, 26)
    assert all(v == 0 for v in p.image_reflect)


@pytest.mark.parametrize("config", [ov.IMAGES_WALL_CANAL, ov.IMAGES_WALL_1_18,
                                    ov.IMAGES_WALL_PARALLEL_PLANES, ov.IMAGES_WALL_1_OPEN,
                                    ov.IMAGES_WALL_2_2])
@pytest.mark.parametrize("method", [ov.METHOD_DIRECT, ov.METHOD_BH])
def test_wall_velocities_bitexact_vs_reference(oracle, ref, config, method):
    """update_tangle with mirror images: the BH path reflects each image's
    velocity (src/tangle.c:644-646), the direct path adds it unreflected
    (src/tangle.c:678-683, kept as the reference has it)."""
    t, p = wall_case(config, method)
    want = oracle.update_tangle(t, p)
    r = ref.tangle(t, p, wall_images=config)
    r.update_tangle(0.0)
    got = r.get()
    live = t.live
    for key in ("tangents", "normals", "vs", "vels"):
        assert np.array_equal(got[key][live], want[key][live]), key


def test_mirror_image_is_the_reflected_reversed_copy(oracle):
    """One mirror wall at z = 0 (wall_1_open), theta -> 0: the field of the
    image equals that of the tangle reflected through the wall with its
    orientation reversed, placed explicitly in an open box.  A dropped
    reflection of the velocity, a wrong wall coordinate or an unreversed
    copy each fail this by O(1)."""
    D = 0.1
    t, p = wall_case(ov.IMAGES_WALL_1_OPEN, ov.METHOD_BH, theta=1e-6)
    base = oracle.update_tangle(t, p)
    live = t.live
    nl = len(live)
    x = t.vnodes[live]
    c = t.conn[live]
    mx = x.copy()
    mx[:, 2] = -mx[:, 2]                 # 2 * 0 - z
    mc = c[:, ::-1] + nl                 # forward and reverse swapped
    v = np.concatenate([x, mx])
    cc = np.concatenate([c, mc]).astype(np.int32)
    st = np.zeros((2 * nl, 2), dtype=np.int32)
    st[:, 1] = -1
    big = ov.Tangle(v, cc, st)
    q = ov.Params.from_buffer_copy(p)
    q.n_images = 0
    for f in range(6):
        q.wall[f] = ov.WALL_OPEN
    q.box_lo[2] = -D
    o = oracle.update_tangle(big, q)
    assert rel_l2(o["vs"][:nl], base["vs"][live]) < 1e-12
    assert rel_l2(o["vels"][:nl], base["vels"][live]) < 1e-12
    # the image matters: without it the velocities differ at O(1e-2)
    q2 = ov.Params.from_buffer_copy(p)
    q2.n_images = 0
    alone = oracle.update_tangle(t, q2)
    assert rel_l2(alone["vs"][live], base["vs"][live]) > 1e-3


def pinned_case(method, slip=False, seed=11, lines=3, rings=3):
    """Wavy lines from the DOWN wall (z = 0) to the UP wall (z = D), their
    end points pinned there, plus rings; parallel-planes geometry."""
    D, dxi = 0.1, 2e-3
    t = rings_inside(seed, rings, D, 0.012, dxi)
    rng = np.random.default_rng(seed + 1)
    xs, conns, sts = [t.vnodes], [t.conn], [t.status]
    off = t.n
    # 52 points: spacing D/51, so no point but the pinned ends lies on a
    # dyadic plane of the box, where the tuned trees break ties upward
    # and the reference tree downward
    m = 52
    pin = ov.PINNED_SLIP if slip else ov.PINNED
    for _ in range(lines):
        x0, y0 = rng.uniform(0.02, 0.08, 2)
        amp, ph = rng.uniform(0.002, 0.006), rng.uniform(0, 2 * np.pi)
        z = np.linspace(0.0, D, m)
        pts = np.stack([x0 + amp * np.sin(2 * np.pi * z / D + ph),
                        y0 + amp * np.cos(3 * np.pi * z / D + ph), z], 1)
        j = np.arange(m)
        c = np.stack([np.where(j < m - 1, off + j + 1, -1), np.where(j > 0, off + j - 1, -1)], 1)
        st = np.zeros((m, 2), dtype=np.int32)
        st[:, 1] = -1
        st[0] = (pin, ov.FACE_DOWN)
        st[-1] = (pin, ov.FACE_UP)
        xs.append(pts)
        conns.append(c)
        sts.append(st)
        off += m
    tt = ov.Tangle(np.concatenate(xs), np.concatenate(conns).astype(np.int32),
                   np.concatenate(sts).astype(np.int32))
    p = configs.params(D, dxi, configs.T16, 0.5, method=method, theta=0.4)
    for f, w in enumerate(WALLS[ov.IMAGES_WALL_PARALLEL_PLANES]):
        p.wall[f] = w
    ov.set_wall_images(p, ov.IMAGES_WALL_PARALLEL_PLANES)
    return tt, p


@pytest.mark.parametrize("slip", [False, True])
@pytest.mark.parametrize("method", [ov.METHOD_DIRECT, ov.METHOD_BH])
def test_pinned_lines_bitexact_vs_reference(oracle, ref, method, slip):
    """Lines pinned to mirror walls: the stencil walks past an open end into
    the mirrored ghost (step_node, src/tangle.c:170-185), PINNED nodes take
    the boundary velocity (:608-615), PINNED_SLIP ones lose their normal
    velocity (:718-725)."""
    t, p = pinned_case(method, slip)
    want = oracle.update_tangle(t, p)
    r = ref.tangle(t, p, wall_images=ov.IMAGES_WALL_PARALLEL_PLANES)
    r.update_tangle(0.0)
    got = r.get()
    live = t.live
    for key in ("tangents", "normals", "vs", "vels"):
        assert np.array_equal(got[key][live], want[key][live]), key
    ends = np.where(t.status[:, 0] != ov.FREE)[0]
    if slip:  # no velocity through the wall
        assert np.all(want["vels"][ends, 2] == 0)
    else:
        assert np.all(want["vels"][ends] == 0) and np.all(want["vs"][ends] == 0)


def test_pinned_ghost_is_the_mirror_image(oracle):
    """A straight line normal to the wall: its ghost continues it, so the
    pinned end's tangent is the line direction and its normal vanishes."""
    D, dxi = 0.1, 2e-3
    m = 21
    z = np.linspace(0.0, D, m)
    x = np.stack([np.full(m, 0.05), np.full(m, 0.04), z], 1)
    j = np.arange(m)
    c = np.stack([np.where(j < m - 1, j + 1, -1), np.where(j > 0, j - 1, -1)], 1)
    st = np.zeros((m, 2), dtype=np.int32)
    st[:, 1] = -1
    st[0] = (ov.PINNED, ov.FACE_DOWN)
    st[-1] = (ov.PINNED, ov.FACE_UP)
    t = ov.Tangle(x, c.astype(np.int32), st)
    p = configs.params(D, dxi, configs.T16, 0.5, method=ov.METHOD_DIRECT)
    for f, w in enumerate(WALLS[ov.IMAGES_WALL_PARALLEL_PLANES]):
        p.wall[f] = w
    ov.set_wall_images(p, ov.IMAGES_WALL_PARALLEL_PLANES)
    tn, nm = oracle.tangents_normals(t, p)
    assert np.allclose(tn[[0, -1]], [[0, 0, 1], [0, 0, 1]], atol=1e-12)
    assert np.abs(nm[[0, -1]]).max() < 1e-9


@pytest.mark.parametrize("faces,reverse", [((ov.FACE_DOWN,), True),
                                           ((ov.FACE_BACK, ov.FACE_DOWN), False)])
def test_mirror_image_orientation_follows_reflection_parity(oracle, faces, reverse):
    """An image mirrored through k walls is the tangle reflected through
    them, reversed when k is odd: with M the reflection, M v(M r + c) is
    det(M) times the field of the reflected copy at r.  Checked with one
    custom image at theta -> 0 against an explicit copy."""
    D = 0.1
    t, p = wall_case(ov.IMAGES_WALL_1_OPEN, ov.METHOD_BH, theta=1e-6)
    p.n_images = 1
    p.images[0][0], p.images[0][1], p.images[0][2] = 0, 0, -1
    p.image_reflect[0] = sum(1 << f for f in faces)
    for f in range(6):
        p.wall[f] = ov.WALL_OPEN
    base = oracle.update_tangle(t, p)
    live = t.live
    nl = len(live)
    x, c = t.vnodes[live], t.conn[live]
    mx = x.copy()
    for f in faces:  # walls at box_lo = 0
        mx[:, f // 2] = -mx[:, f // 2]
    mc = (c[:, ::-1] if reverse else c) + nl
    big = ov.Tangle(np.concatenate([x, mx]), np.concatenate([c, mc]).astype(np.int32),
                    np.tile(np.array([[0, -1]], dtype=np.int32), (2 * nl, 1)))
    q = ov.Params.from_buffer_copy(p)
    q.n_images = 0
    for a in range(3):
        q.box_lo[a] = -D
    o = oracle.update_tangle(big, q)
    assert rel_l2(o["vs"][:nl], base["vs"][live]) < 1e-12


def lines_and_rings(lengths, seed=5, rings=3, orders=None):
    """Rings plus lines of the given point counts from the DOWN wall to the
    UP wall, pinned at both ends (parallel-planes geometry).  orders[i] puts
    line i's lowest slot at its DOWN end ("down", the default), at its UP
    end ("up": slots in reverse) or mid-line ("mid": slots rotated)."""
    D, dxi = 0.1, 2e-3
    t = rings_inside(seed, rings, D, 0.012, dxi)
    rng = np.random.default_rng(seed)
    xs, conns, sts = [t.vnodes], [t.conn], [t.status]
    off = t.n
    for i, m in enumerate(lengths):
        x0, y0 = rng.uniform(0.02, 0.08, 2)
        z = np.linspace(0.0, D, m)
        j = np.arange(m)
        order = orders[i] if orders else "down"
        # pos[s]: position along the line (0 = DOWN end) held by slot off + s
        pos = {"down": j, "up": m - 1 - j, "mid": (j + m // 2) % m}[order]
        slot = np.empty(m, dtype=np.int64)
        slot[pos] = off + j
        xs.append(np.stack([np.full(m, x0), np.full(m, y0), z[pos]], 1))
        conns.append(np.stack([np.where(pos < m - 1, slot[np.minimum(pos + 1, m - 1)], -1),
                               np.where(pos > 0, slot[np.maximum(pos - 1, 0)], -1)], 1))
        st = np.zeros((m, 2), dtype=np.int32)
        st[:, 1] = -1
        st[pos == 0] = (ov.PINNED, ov.FACE_DOWN)
        st[pos == m - 1] = (ov.PINNED_SLIP, ov.FACE_UP)
        sts.append(st)
        off += m
    t = ov.Tangle(np.concatenate(xs), np.concatenate(conns).astype(np.int32),
                  np.concatenate(sts).astype(np.int32))
    p = configs.params(D, dxi, configs.T16, 0.5, method=ov.METHOD_DIRECT)
    for f, w in enumerate(WALLS[ov.IMAGES_WALL_PARALLEL_PLANES]):
        p.wall[f] = w
    ov.set_wall_images(p, ov.IMAGES_WALL_PARALLEL_PLANES)
    return t, p


def test_small_open_lines_removed_as_reference(oracle, ref):
    """eliminate_small_loops with lines pinned to walls: an open line counts
    its forward steps to the pinned end plus the steps back from the start
    point (src/tangle.c:1081-1093) and is erased forward from its far end."""
    t, p = lines_and_rings([2, 3, 5, 6, 7, 40, 4])
    r = ref.tangle(t, p, wall_images=ov.IMAGES_WALL_PARALLEL_PLANES)
    r.eliminate_small_loops(6)
    want_conn, want_st = r.get()["conn"], r.status()
    w = ov.Tangle(t.vnodes.copy(), t.conn.copy(), t.status.copy())
    removed = oracle.eliminate_small_loops(w, 6)
    assert np.array_equal(w.conn, want_conn)
    assert np.array_equal(w.status[:, 0], want_st[:, 0])
    assert removed == 2 + 3 + 5 + 6 + 4  # lines of at most 6 points


def test_small_open_lines_started_mid_line_or_far_end(oracle, ref):
    """The same pass on short lines whose lowest slot is mid-line or at the
    UP end, so the visit counts back from its start point before erasing
    forward from the far end (src/tangle.c:1081-1093)."""
    lengths = [3, 5, 6, 4, 6, 9, 5]
    orders = ["mid", "up", "mid", "up", "down", "mid", "up"]
    t, p = lines_and_rings(lengths, seed=11, orders=orders)
    r = ref.tangle(t, p, wall_images=ov.IMAGES_WALL_PARALLEL_PLANES)
    r.eliminate_small_loops(6)
    want_conn, want_st = r.get()["conn"], r.status()
    w = ov.Tangle(t.vnodes.copy(), t.conn.copy(), t.status.copy())
    removed = oracle.eliminate_small_loops(w, 6)
    assert np.array_equal(w.conn, want_conn)
    assert np.array_equal(w.status[:, 0], want_st[:, 0])
    assert removed == 3 + 5 + 6 + 4 + 6 + 5  # every line but the 9-point one
