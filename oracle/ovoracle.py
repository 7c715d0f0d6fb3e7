"""ctypes wrapper of the CPU oracle.

TEST INFRASTRUCTURE ONLY.  Used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; never by the product.

* :class:`Oracle` wraps ``oracle/libovoracle.so``, the plain-C FP64
  restatement of the reference algorithm (oracle/ovoracle.c).
* :class:`RefLib` wraps ``oracle/_ref/libovref.so``, the reference's own
  sources compiled in place by ``make -C oracle ref`` (present only where
  it was built; it travels to the GPU box as a built file).
"""
import ctypes as C
import os

import numpy as np

from openvort_amd import Params, Tangle  # struct layouts only

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "libovoracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libovref.so")
REF_BENCH = os.path.join(HERE, "_ref", "ovref_bench")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


class Oracle:
    def __init__(self, path=ORACLE_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path}: run `make -C oracle`")
        L = self.lib = C.CDLL(path)
        P = C.POINTER(Params)
        L.ovo_tangents_normals.argtypes = [C.c_int, _dp, _ip, _ip, P, _dp, _dp]
        L.ovo_velocities.argtypes = [C.c_int, _dp, _ip, _ip, P, C.c_double, _dp, _dp, _dp, _dp,
                                     C.c_int, _dp]
        L.ovo_velocity_direct_one.argtypes = [C.c_int, _dp, _ip, _ip, P, C.c_double, _dp, _dp,
                                              C.c_int, _dp, _dp]
        L.ovo_bs_direct_one.argtypes = [C.c_int, _dp, _ip, P, C.c_int, _dp]
        L.ovo_segment_field.argtypes = [_dp, _dp, _dp, C.c_double, _dp]
        L.ovo_rk4_step.argtypes = [C.c_int, _dp, _ip, _ip, P, C.c_double, C.c_double, _dp, C.c_int]
        L.ovo_enforce_periodic.argtypes = [C.c_int, _dp, _ip, P]
        L.ovo_reconnect_candidates.argtypes = [C.c_int, _dp, _ip, _ip, P, C.c_long, _ip]
        L.ovo_reconnect_candidates.restype = C.c_long
        L.ovo_reconnect.argtypes = [C.c_int, _dp, _ip, _ip, P, C.c_double, C.c_int]
        L.ovo_line_length.argtypes = [C.c_int, _dp, _ip, _ip, P]
        L.ovo_line_length.restype = C.c_double
        L.ovo_eliminate_small_loops.argtypes = [C.c_int, _ip, _ip, C.c_int]
        L.ovo_remesh.argtypes = [C.POINTER(C.c_int), C.c_int, _dp, _ip, _ip, C.POINTER(C.c_int),
                                 P, C.c_double, C.c_double, C.POINTER(C.c_int)]
        L.ovo_bh_field.argtypes = [C.c_int, _dp, _ip, _ip, P, C.c_int, _dp, _dp]
        L.ovo_bh_tree_stats.argtypes = [C.c_int, _dp, _ip, _ip, P, C.POINTER(C.c_long),
                                        C.POINTER(C.c_int), C.POINTER(C.c_long)]

    def tangents_normals(self, t: Tangle, p):
        tn = np.zeros((t.n, 3))
        nm = np.zeros((t.n, 3))
        self.lib.ovo_tangents_normals(t.n, _d(t.vnodes), _i(t.conn), _i(t.status), C.byref(p),
                                      _d(tn), _d(nm))
        return tn, nm

    def velocities(self, t: Tangle, p, time=0.0, tangents=None, normals=None, threads=8):
        if tangents is None:
            tangents, normals = self.tangents_normals(t, p)
        vs = np.zeros((t.n, 3))
        vels = np.zeros((t.n, 3))
        cnt = C.c_double(0)
        self.lib.ovo_velocities(t.n, _d(t.vnodes), _i(t.conn), _i(t.status), C.byref(p), time,
                                _d(tangents), _d(normals), _d(vs), _d(vels), threads,
                                C.byref(cnt))
        return vs, vels, cnt.value

    def update_tangle(self, t: Tangle, p, time=0.0, threads=8):
        tn, nm = self.tangents_normals(t, p)
        vs, vels, cnt = self.velocities(t, p, time, tn, nm, threads)
        return dict(tangents=tn, normals=nm, vs=vs, vels=vels, interactions=cnt)

    def velocity_direct_one(self, t: Tangle, p, k, time=0.0, tangents=None, normals=None):
        if tangents is None:
            tangents, normals = self.tangents_normals(t, p)
        vs = np.zeros(3)
        vels = np.zeros(3)
        self.lib.ovo_velocity_direct_one(t.n, _d(t.vnodes), _i(t.conn), _i(t.status), C.byref(p),
                                         time, _d(tangents), _d(normals), k, _d(vs), _d(vels))
        return vs, vels

    def segment_field(self, r1, r2, r, kappa):
        out = np.zeros(3)
        a, b, c = (np.ascontiguousarray(v, dtype=np.float64) for v in (r1, r2, r))
        self.lib.ovo_segment_field(_d(a), _d(b), _d(c), kappa, _d(out))
        return out

    def rk4_step(self, t: Tangle, p, time, dt, vels, threads=8):
        """In place on t.vnodes; vels = k1."""
        vels = np.ascontiguousarray(vels)
        self.lib.ovo_rk4_step(t.n, _d(t.vnodes), _i(t.conn), _i(t.status), C.byref(p), time, dt,
                              _d(vels), threads)

    def enforce_periodic(self, t: Tangle, p):
        self.lib.ovo_enforce_periodic(t.n, _d(t.vnodes), _i(t.status), C.byref(p))

    def reconnect_candidates(self, t: Tangle, p):
        n = self.lib.ovo_reconnect_candidates(t.n, _d(t.vnodes), _i(t.conn), _i(t.status),
                                              C.byref(p), 0, None)
        pairs = np.zeros((max(n, 1), 2), dtype=np.int32)
        self.lib.ovo_reconnect_candidates(t.n, _d(t.vnodes), _i(t.conn), _i(t.status),
                                          C.byref(p), n, _i(pairs))
        return pairs[:n]

    def reconnect(self, t: Tangle, p, time=0.0):
        """In place on t.conn; returns the number of reconnections."""
        return self.lib.ovo_reconnect(t.n, _d(t.vnodes), _i(t.conn), _i(t.status), C.byref(p),
                                      time, 8)

    def line_length(self, t: Tangle, p):
        return self.lib.ovo_line_length(t.n, _d(t.vnodes), _i(t.conn), _i(t.status), C.byref(p))

    def eliminate_small_loops(self, t: Tangle, cutoff=5):
        """In place on t.conn / t.status; returns the deleted point count."""
        return self.lib.ovo_eliminate_small_loops(t.n, _i(t.conn), _i(t.status), cutoff)

    def remesh(self, t: Tangle, p, min_dist, max_dist, next_empty=0, cap_factor=4):
        """Returns (new Tangle, added, removed, next_empty); capacity doubles as
        the reference's expand_tangle."""
        cap = t.n * cap_factor
        v = np.zeros((cap, 3)); v[:t.n] = t.vnodes
        c = -np.ones((cap, 2), dtype=np.int32); c[:t.n] = t.conn
        s = np.zeros((cap, 2), dtype=np.int32); s[:, 0] = -1; s[:, 1] = -1; s[:t.n] = t.status
        N, ne, rem = C.c_int(t.n), C.c_int(next_empty), C.c_int(0)
        added = self.lib.ovo_remesh(C.byref(N), cap, _d(v), _i(c), _i(s), C.byref(ne),
                                    C.byref(p), min_dist, max_dist, C.byref(rem))
        if added < 0:
            raise RuntimeError("remesh capacity")
        n = N.value
        return Tangle(v[:n], c[:n], s[:n]), added, rem.value, ne.value

    def bh_field(self, t: Tangle, p, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        out = np.zeros_like(pts)
        self.lib.ovo_bh_field(t.n, _d(t.vnodes), _i(t.conn), _i(t.status), C.byref(p), len(pts),
                              _d(pts), _d(out))
        return out

    def tree_stats(self, t: Tangle, p):
        nodes, depth, ne = C.c_long(0), C.c_int(0), C.c_long(0)
        self.lib.ovo_bh_tree_stats(t.n, _d(t.vnodes), _i(t.conn), _i(t.status), C.byref(p),
                                   C.byref(nodes), C.byref(depth), C.byref(ne))
        return nodes.value, depth.value, ne.value


def ref_available():
    return os.path.exists(REF_LIB)


class RefLib:
    """The compiled reference (strict IEEE flavour) through oracle/ref_harness.c."""

    def __init__(self, path=REF_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path}: run `make -C oracle ref` where the reference exists")
        L = self.lib = C.CDLL(path)
        P = C.POINTER(Params)
        vp = C.c_void_p
        L.ref_setup.argtypes = [P, C.c_int]
        L.ref_image_table.argtypes = [C.c_int, _ip]
        L.ref_wall_table.argtypes = [C.c_int, _ip, _ip]
        L.ref_set_wall_images.argtypes = [C.c_void_p, C.c_int]
        L.ref_tangle_new.argtypes = [C.c_int, _dp, _ip, _ip, P]
        L.ref_tangle_new.restype = vp
        L.ref_tangle_free.argtypes = [vp]
        L.ref_set_vels.argtypes = [vp, _dp]
        L.ref_set_tangents.argtypes = [vp, _dp, _dp]
        L.ref_get.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, _ip]
        for f in ("ref_update_tangents_normals", "ref_enforce_periodic"):
            getattr(L, f).argtypes = [vp]
        for f in ("ref_update_velocities", "ref_update_tangle"):
            getattr(L, f).argtypes = [vp, C.c_double]
        L.ref_rk4_step.argtypes = [vp, C.c_double, C.c_double]
        L.ref_reconnect.argtypes = [vp, C.c_double]
        L.ref_velocity_one.argtypes = [vp, C.c_int, C.c_double, _dp, _dp]
        L.ref_bh_field.argtypes = [vp, C.c_int, _dp, C.c_double, _dp]
        L.ref_eliminate_small_loops.argtypes = [vp, C.c_int]
        L.ref_remesh.argtypes = [vp, C.c_double, C.c_double]
        L.ref_enforce_walls.argtypes = [vp, C.c_double, C.c_int]
        L.ref_num_slots.argtypes = [vp]
        L.ref_save_tangle.argtypes = [vp, C.c_char_p]
        L.ref_load_tangle.argtypes = [C.c_char_p, C.c_void_p]
        L.ref_load_tangle.restype = vp
        L.ref_get_status.argtypes = [vp, _ip]
        L.ref_time_phases.argtypes = [vp, C.c_double, C.c_double, _dp]
        L.ref_add_loop.argtypes = [vp, _dp, _dp, C.c_double, C.c_double, C.c_double]
        L.ref_inject_loops.argtypes = [vp, C.c_long, _dp, C.c_int, C.c_double, C.c_double,
                                       C.c_double]
        L.ref_inject_line_pairs.argtypes = [vp, C.c_long, _dp, C.c_int, C.c_double, C.c_int,
                                            C.c_int, C.c_double, C.c_double, C.c_int]

    def image_table(self, n):
        out = np.zeros((26, 3), dtype=np.int32)
        m = self.lib.ref_image_table(n, _i(out))
        return out[:m]

    def wall_table(self, config):
        """(shifts, face masks) of the reference's mirror-wall table."""
        sh = np.zeros((26, 3), dtype=np.int32)
        mk = np.zeros(26, dtype=np.int32)
        m = self.lib.ref_wall_table(config, _i(sh), _i(mk))
        if m < 0:
            raise ValueError(f"no wall table {config}")
        return sh[:m], mk[:m]

    def tangle(self, t: Tangle, p, threads=8, wall_images=None):
        """wall_images: an OVG_IMAGES_WALL_* id, the reference's own table
        (p's images must then be that table, as set_wall_images loads it)."""
        self.lib.ref_setup(C.byref(p), threads)
        h = self.lib.ref_tangle_new(t.n, _d(t.vnodes), _i(t.conn), _i(t.status), C.byref(p))
        if not h:
            raise ValueError("reference has no such image table")
        if wall_images is not None and self.lib.ref_set_wall_images(h, wall_images) != 0:
            self.lib.ref_tangle_free(h)
            raise ValueError(f"no wall table {wall_images}")
        return RefTangle(self, h, t.n)


def ref_load(ref, path, p):
    """The reference's load_tangle into create_tangle(512)."""
    h = ref.lib.ref_load_tangle(str(path).encode(), C.byref(p))
    if not h:
        raise ValueError("reference has no such image table")
    return RefTangle(ref, h, ref.lib.ref_num_slots(h))


class RefTangle:
    def __init__(self, ref, h, n):
        self.ref, self.h, self.n = ref, h, n

    def __del__(self):
        if self.h:
            self.ref.lib.ref_tangle_free(self.h)
            self.h = None

    def get(self):
        out = {k: np.zeros((self.n, 3)) for k in ("vnodes", "vels", "vs", "tangents", "normals")}
        conn = np.zeros((self.n, 2), dtype=np.int32)
        self.ref.lib.ref_get(self.h, _d(out["vnodes"]), _d(out["vels"]), _d(out["vs"]),
                             _d(out["tangents"]), _d(out["normals"]), _i(conn))
        out["conn"] = conn
        return out

    def update_tangle(self, time=0.0):
        self.ref.lib.ref_update_tangle(self.h, time)

    def update_tangents_normals(self):
        self.ref.lib.ref_update_tangents_normals(self.h)

    def update_velocities(self, time=0.0):
        self.ref.lib.ref_update_velocities(self.h, time)

    def rk4_step(self, time, dt):
        self.ref.lib.ref_rk4_step(self.h, time, dt)

    def reconnect(self, time=0.0):
        return self.ref.lib.ref_reconnect(self.h, time)

    def enforce_periodic(self):
        self.ref.lib.ref_enforce_periodic(self.h)

    def eliminate_small_loops(self, cutoff=5):
        self.ref.lib.ref_eliminate_small_loops(self.h, cutoff)

    def remesh(self, min_dist, max_dist):
        self.ref.lib.ref_remesh(self.h, min_dist, max_dist)
        self.n = self.ref.lib.ref_num_slots(self.h)

    def enforce_walls(self, time=0.0, pin_mode=1):
        """enforce_wall_boundaries (src/tangle.c:966-1016); pin_mode: status
        of the new wall nodes."""
        self.ref.lib.ref_enforce_walls(self.h, time, pin_mode)
        self.n = self.ref.lib.ref_num_slots(self.h)

    def add_loop(self, center, direction, r, dl_min, dl_max):
        c = np.ascontiguousarray(center, dtype=np.float64)
        d = np.ascontiguousarray(direction, dtype=np.float64)
        self.ref.lib.ref_add_loop(self.h, _d(c), _d(d), r, dl_min, dl_max)
        self.n = self.ref.lib.ref_num_slots(self.h)

    def inject_loops(self, seed, times, frequency, dl_min, dl_max):
        """srand48(seed), then inject_loop at each time (its static last
        injection time persists for the life of the library)."""
        ts = np.ascontiguousarray(times, dtype=np.float64)
        self.ref.lib.ref_inject_loops(self.h, seed, _d(ts), len(ts), frequency, dl_min, dl_max)
        self.n = self.ref.lib.ref_num_slots(self.h)

    def inject_line_pairs(self, seed, times, frequency, n_pairs, polarized, dl_min, dl_max,
                          pin_mode=1):
        """srand48(seed), then inject_line_pairs at each time."""
        ts = np.ascontiguousarray(times, dtype=np.float64)
        self.ref.lib.ref_inject_line_pairs(self.h, seed, _d(ts), len(ts), frequency, n_pairs,
                                           int(polarized), dl_min, dl_max, pin_mode)
        self.n = self.ref.lib.ref_num_slots(self.h)

    def save(self, path):
        """The reference's save_tangle (src/utils.c:132-184)."""
        self.ref.lib.ref_save_tangle(self.h, str(path).encode())

    def status(self):
        st = np.zeros((self.n, 2), dtype=np.int32)
        self.ref.lib.ref_get_status(self.h, _i(st))
        return st

    def velocity_one(self, k, time=0.0):
        vs, vels = np.zeros(3), np.zeros(3)
        self.ref.lib.ref_velocity_one(self.h, k, time, _d(vs), _d(vels))
        return vs, vels

    def bh_field(self, pts, theta):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        out = np.zeros_like(pts)
        self.ref.lib.ref_bh_field(self.h, len(pts), _d(pts), theta, _d(out))
        return out
