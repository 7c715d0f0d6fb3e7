"""Python bindings (ctypes) of the MI355X vortex-filament engine.

The product is the C ABI of ``include/openvort_gpu.h`` implemented by the
in-tree ``openvort_amd/libopenvort_gpu.so`` (hand-written HIP for gfx950).
This module only restates its structs and wraps its entry points with numpy
arrays for tests and benchmarks.  Importing it fails loudly when the
library has not been built; creating an :class:`Engine` fails loudly when
no gfx950 GPU is present.  There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libopenvort_gpu.so")
DROPIN_PATH = os.path.join(_HERE, "libopenvort_dropin.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `make -C openvort_amd` "
        "(or __graft_entry__.build())")

_lib = C.CDLL(LIB_PATH)

# status codes (include/openvort_gpu.h)
OK, EINVAL, ENOMEM, EDEVICE, EUNSUPPORTED, ELIMIT = 0, 1, 2, 3, 4, 5
EMPTY, FREE, PINNED, PINNED_SLIP = -1, 0, 1, 2
NOALLOC_MAX_TARGETS = 131072  # OVG_NOALLOC_MAX_TARGETS
WALL_OPEN, WALL_PERIODIC, WALL_MIRROR = 0, 1, 2
FACE_LEFT, FACE_RIGHT, FACE_BACK, FACE_FRONT, FACE_DOWN, FACE_UP = range(6)
# mirror-wall image tables (include/openvort_gpu.h OVG_IMAGES_WALL_*)
(IMAGES_WALL_1_OPEN, IMAGES_WALL_1_6, IMAGES_WALL_1_18, IMAGES_WALL_1_26, IMAGES_WALL_2_4,
 IMAGES_WALL_2_2, IMAGES_WALL_PARALLEL_PLANES, IMAGES_WALL_CANAL) = range(1, 9)
FLOW_NONE, FLOW_SIMPLE, FLOW_DECREASING = 0, 1, 2
METHOD_DIRECT, METHOD_BH = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
TRAVERSE_POINT, TRAVERSE_GROUP, TRAVERSE_LIST = 0, 1, 2
FIELD_VELS, FIELD_VS, FIELD_TANGENTS, FIELD_NORMALS = 0, 1, 2, 3
RECVEL_DIRECT, RECVEL_TREE = 0, 1
TOPO_HOST, TOPO_DEVICE = 0, 1


class Flow(C.Structure):
    _fields_ = [("type", C.c_int32), ("pad_", C.c_int32), ("strength", C.c_double),
                ("time_of_linear_decrease", C.c_double),
                ("starting_strength", C.c_double), ("direction", C.c_double * 3)]


class Params(C.Structure):
    _fields_ = [
        ("kappa", C.c_double), ("vortex_width", C.c_double),
        ("use_mutual_friction", C.c_int32), ("pad0_", C.c_int32),
        ("alpha", C.c_double), ("alpha_p", C.c_double),
        ("vn", Flow), ("vs", Flow), ("vb", Flow),
        ("box_lo", C.c_double * 3), ("box_hi", C.c_double * 3),
        ("wall", C.c_int32 * 6), ("n_images", C.c_int32),
        ("images", (C.c_int32 * 3) * 26),
        ("method", C.c_int32), ("precision", C.c_int32), ("theta", C.c_double),
        ("leaf_size", C.c_int32), ("traverse", C.c_int32),
        ("rec_dist", C.c_double), ("rec_angle", C.c_double),
        ("freq_to_cutoff", C.c_double), ("rk4_classic", C.c_int32),
        ("rec_velocity", C.c_int32), ("mixed_near", C.c_double),
        ("points_per_wave", C.c_int32), ("group_level", C.c_int32),
        ("reuse_lists", C.c_int32), ("image_reflect", C.c_int32 * 26),
        ("list_stack", C.c_int32), ("list_tile", C.c_int32), ("list_blocks", C.c_int32),
        ("pin_mode", C.c_int32), ("topology", C.c_int32),
    ]


class TuneRow(C.Structure):
    _fields_ = [("points_per_wave", C.c_int32), ("leaf_size", C.c_int32),
                ("group_level", C.c_int32), ("list_stack", C.c_int32),
                ("list_tile", C.c_int32), ("list_blocks", C.c_int32), ("ms", C.c_double),
                ("pairs", C.c_double)]


class Tune(C.Structure):
    _fields_ = [("n_rows", C.c_int32), ("best", C.c_int32), ("rows", TuneRow * 96)]


class Stats(C.Structure):
    _fields_ = [("interactions", C.c_double), ("interactions_fp32", C.c_double),
                ("tree_nodes", C.c_int64), ("tree_depth", C.c_int32),
                ("n_points", C.c_int32), ("ms_tree", C.c_double), ("ms_bs", C.c_double),
                ("ms_local", C.c_double), ("bs_launches", C.c_int32), ("pad_", C.c_int32)]


_P = C.POINTER
_dp = _P(C.c_double)
_ip = _P(C.c_int32)


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("ovg_last_error", C.c_char_p)
_sig("ovg_version", C.c_char_p)
_sig("ovg_params_sizeof", C.c_size_t)
_sig("ovg_stats_sizeof", C.c_size_t)
_sig("ovg_tune_sizeof", C.c_size_t)
_sig("ovg_params_default", None, _P(Params), C.c_double)
_sig("ovg_params_set_periodic_images", C.c_int, _P(Params), C.c_int)
_sig("ovg_params_set_wall_images", C.c_int, _P(Params), C.c_int)
_sig("ovg_create", C.c_int, _P(C.c_void_p), C.c_int)
_sig("ovg_destroy", C.c_int, C.c_void_p)
_sig("ovg_set_params", C.c_int, C.c_void_p, _P(Params))
_sig("ovg_get_params", C.c_int, C.c_void_p, _P(Params))
_sig("ovg_upload", C.c_int, C.c_void_p, C.c_int, _dp, _ip, _ip)
_sig("ovg_upload_positions", C.c_int, C.c_void_p, _dp)
_sig("ovg_upload_field", C.c_int, C.c_void_p, C.c_int, _dp)
_sig("ovg_upload_connections", C.c_int, C.c_void_p, _ip, _ip)
_sig("ovg_download", C.c_int, C.c_void_p, _dp, _dp, _dp, _dp, _dp)
_sig("ovg_download_connections", C.c_int, C.c_void_p, _ip, _ip)
_sig("ovg_num_slots", C.c_int, C.c_void_p)
_sig("ovg_num_points", C.c_int, C.c_void_p)
_sig("ovg_write_tangle_file", C.c_int, C.c_char_p, C.c_int, _dp, _dp, _dp, _dp, _ip, _ip,
     C.c_int)
_sig("ovg_read_tangle_file", C.c_int, C.c_char_p, C.c_int, _ip, _dp, _dp, _dp, _dp, _ip, _ip)
_sig("ovg_save_tangle", C.c_int, C.c_void_p, C.c_char_p, C.c_int)
_sig("ovg_autotune", C.c_int, C.c_void_p, C.c_int, _P(Tune))
_sig("ovg_load_tangle", C.c_int, C.c_void_p, C.c_char_p)
_sig("ovg_update_tangents_normals", C.c_int, C.c_void_p)
_sig("ovg_update_velocities", C.c_int, C.c_void_p, C.c_double)
_sig("ovg_update_tangle", C.c_int, C.c_void_p, C.c_double)
_sig("ovg_rk4_step", C.c_int, C.c_void_p, C.c_double, C.c_double)
_sig("ovg_enforce_periodic", C.c_int, C.c_void_p)
_sig("ovg_reconnect_candidates", C.c_int, C.c_void_p, C.c_int, _ip, _ip)
_sig("ovg_reconnect", C.c_int, C.c_void_p, C.c_double, _ip)
_sig("ovg_smooth", C.c_int, C.c_void_p)
_sig("ovg_eliminate_small_loops", C.c_int, C.c_void_p, C.c_int, _ip)
_sig("ovg_remesh", C.c_int, C.c_void_p, C.c_double, C.c_double, _ip, _ip)
_sig("ovg_topology_counters", C.c_int, C.c_void_p, _ip)
_sig("ovg_enforce_walls", C.c_int, C.c_void_p, C.c_double, _ip, _ip)
_sig("ovg_loop_points", C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp, C.c_int, _ip)
_sig("ovg_inject_loop_draw", C.c_int, C.c_double, C.c_double, _dp, _dp, _P(C.c_ushort), _dp, _dp,
     _dp, _dp, _ip)
_sig("ovg_add_loop", C.c_int, C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_double, _ip)
_sig("ovg_inject_line_pairs_draw", C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, _dp, _dp,
     _P(C.c_ushort), _dp, _dp, _ip)
_sig("ovg_add_line_z", C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_double,
     C.c_double, _ip)
_sig("ovg_velocity_at", C.c_int, C.c_void_p, C.c_int, _ip, C.c_double, _dp, _dp)
_sig("ovg_velocity_at_noalloc", C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_double,
     C.c_void_p, C.c_void_p)
_sig("ovg_line_length", C.c_int, C.c_void_p, _dp, _ip)
_sig("ovg_get_stats", C.c_int, C.c_void_p, _P(Stats))
_sig("ovg_set_timing", C.c_int, C.c_void_p, C.c_int)
_sig("ovg_synchronize", C.c_int, C.c_void_p)
_sig("ovg_marker_record", C.c_int, C.c_void_p, C.c_int)
_sig("ovg_marker_elapsed", C.c_int, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double))
_sig("ovg_generate_rings", C.c_int, C.c_uint64, C.c_int, C.c_double, C.c_double, C.c_double,
     C.c_double, C.c_int, _dp, _ip, _ip)
_sig("ovg_rings_for_points", C.c_int, C.c_uint64, C.c_int, C.c_double, C.c_double,
     C.c_double, C.c_double)
_sig("ovg_upload_forest", C.c_int, C.c_void_p, C.c_int, _P(Params), _ip, _dp, _ip, _ip)
_sig("ovg_forest_size", C.c_int, C.c_void_p)
_sig("ovg_graph_begin", C.c_int, C.c_void_p)
_sig("ovg_graph_end", C.c_int, C.c_void_p, _ip)
_sig("ovg_graph_launch", C.c_int, C.c_void_p, C.c_int)
_sig("ovg_graph_release", C.c_int, C.c_void_p)
_sig("ovg_set_device_memory_limit", C.c_size_t, C.c_size_t)
_sig("ovg_device_memory_in_use", C.c_size_t)

if (_lib.ovg_params_sizeof() != C.sizeof(Params) or _lib.ovg_stats_sizeof() != C.sizeof(Stats)
        or _lib.ovg_tune_sizeof() != C.sizeof(Tune)):
    raise ImportError("ctypes struct layout does not match libopenvort_gpu.so")


class OvgError(RuntimeError):
    def __init__(self, code, where):
        msg = _lib.ovg_last_error().decode(errors="replace")
        super().__init__(f"{where} failed with code {code}: {msg}")
        self.code = code


def _check(code, where):
    if code != OK:
        raise OvgError(code, where)


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def lib():
    """The loaded ctypes CDLL of libopenvort_gpu.so."""
    return _lib


def set_device_memory_limit(nbytes):
    """Cap the device bytes all contexts of this process may hold (0 = no
    cap); returns the previous cap."""
    return _lib.ovg_set_device_memory_limit(int(nbytes))


def device_memory_in_use():
    return _lib.ovg_device_memory_in_use()


def default_params(D, **kw):
    """Reference defaults (src/initial_conditions.c:12-91) for [0,D]^3 with
    26 periodic images; keyword arguments override fields."""
    p = Params()
    _lib.ovg_params_default(C.byref(p), float(D))
    for k, v in kw.items():
        if k in ("vn", "vs", "vb"):
            f = getattr(p, k)
            typ, strength, direction = v
            f.type = typ
            f.strength = strength
            for i in range(3):
                f.direction[i] = direction[i]
        elif not hasattr(p, k):
            raise AttributeError(k)
        else:
            setattr(p, k, v)
    return p


def set_images(p, n):
    _check(_lib.ovg_params_set_periodic_images(C.byref(p), n), "set_images")


def set_wall_images(p, config):
    """Load one of the reference's mirror-wall image tables (IMAGES_WALL_*)."""
    _check(_lib.ovg_params_set_wall_images(C.byref(p), config), "set_wall_images")


class Tangle:
    """Host tangle in the reference slot layout (include/tangle.h:47-66)."""

    def __init__(self, vnodes, conn, status):
        self.vnodes = np.ascontiguousarray(vnodes, dtype=np.float64).reshape(-1, 3)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, 2)
        self.status = np.ascontiguousarray(status, dtype=np.int32).reshape(-1, 2)

    @property
    def n(self):
        return self.vnodes.shape[0]

    @property
    def live(self):
        return np.nonzero(self.status[:, 0] != EMPTY)[0]

    def copy(self):
        return Tangle(self.vnodes.copy(), self.conn.copy(), self.status.copy())


def generate_rings(seed, n_rings, D, r_min, r_max, dxi, n_cap=None):
    """Seeded ring tangle (ovg_generate_rings); n_cap defaults to the next
    power of two at or above the point count, as the reference's doubling
    capacity (src/tangle.c:248)."""
    total = _lib.ovg_generate_rings(seed, n_rings, D, r_min, r_max, dxi, 0, None, None, None)
    if total < 0:
        raise OvgError(EINVAL, "generate_rings(count)")
    if n_cap is None:
        n_cap = 1
        while n_cap < total:
            n_cap *= 2
    v = np.zeros((n_cap, 3))
    c = np.zeros((n_cap, 2), dtype=np.int32)
    s = np.zeros((n_cap, 2), dtype=np.int32)
    got = _lib.ovg_generate_rings(seed, n_rings, D, r_min, r_max, dxi, n_cap, _d(v), _i(c), _i(s))
    if got != total:
        raise OvgError(ELIMIT, "generate_rings")
    return Tangle(v, c, s)


def rings_for_points(seed, target, D, r_min, r_max, dxi):
    return _lib.ovg_rings_for_points(seed, target, D, r_min, r_max, dxi)


class Engine:
    """One device context (one HIP stream) holding one tangle."""

    def __init__(self, params, device=0):
        h = C.c_void_p()
        _check(_lib.ovg_create(C.byref(h), device), "ovg_create")
        self._h = h
        self.set_params(params)
        self.n = 0

    def close(self):
        if self._h:
            _lib.ovg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def get_params(self):
        """The context's parameters (after ovg_autotune, with its pick)."""
        p = Params()
        _check(_lib.ovg_get_params(self._h, C.byref(p)), "ovg_get_params")
        return p

    def set_params(self, p):
        self.params = p
        _check(_lib.ovg_set_params(self._h, C.byref(p)), "ovg_set_params")

    def upload(self, t):
        self.n = t.n
        _check(_lib.ovg_upload(self._h, t.n, _d(t.vnodes), _i(t.conn), _i(t.status)), "ovg_upload")

    def upload_positions(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        _check(_lib.ovg_upload_positions(self._h, _d(v)), "ovg_upload_positions")

    def upload_field(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(_lib.ovg_upload_field(self._h, field, _d(v)), "ovg_upload_field")

    def download(self):
        out = {k: np.zeros((self.n, 3)) for k in ("vnodes", "vels", "vs", "tangents", "normals")}
        _check(_lib.ovg_download(self._h, _d(out["vnodes"]), _d(out["vels"]), _d(out["vs"]),
                                 _d(out["tangents"]), _d(out["normals"])), "ovg_download")
        return out

    def download_connections(self):
        c = np.zeros((self.n, 2), dtype=np.int32)
        s = np.zeros((self.n, 2), dtype=np.int32)
        _check(_lib.ovg_download_connections(self._h, _i(c), _i(s)), "ovg_download_connections")
        return c, s

    def update_tangents_normals(self):
        _check(_lib.ovg_update_tangents_normals(self._h), "ovg_update_tangents_normals")

    def update_velocities(self, t=0.0):
        _check(_lib.ovg_update_velocities(self._h, t), "ovg_update_velocities")

    def update_tangle(self, t=0.0):
        _check(_lib.ovg_update_tangle(self._h, t), "ovg_update_tangle")

    def rk4_step(self, t, dt):
        _check(_lib.ovg_rk4_step(self._h, t, dt), "ovg_rk4_step")

    def enforce_periodic(self):
        _check(_lib.ovg_enforce_periodic(self._h), "ovg_enforce_periodic")

    def reconnect_candidates(self, max_pairs=1 << 20):
        pairs = np.zeros((max_pairs, 2), dtype=np.int32)
        n = C.c_int32(0)
        _check(_lib.ovg_reconnect_candidates(self._h, max_pairs, _i(pairs), C.byref(n)),
               "ovg_reconnect_candidates")
        return pairs[: n.value].copy()

    # The topology passes return their counts; counts=False passes no count
    # pointer, so that with topology = TOPO_DEVICE the pass queues on the
    # GPU without waiting (returns None).
    def reconnect(self, t=0.0, counts=True):
        n = C.c_int32(0)
        _check(_lib.ovg_reconnect(self._h, t, C.byref(n) if counts else None), "ovg_reconnect")
        return n.value if counts else None

    def eliminate_small_loops(self, cutoff=5, counts=True):
        n = C.c_int32(0)
        _check(_lib.ovg_eliminate_small_loops(self._h, cutoff, C.byref(n) if counts else None),
               "ovg_eliminate_small_loops")
        return n.value if counts else None

    def remesh(self, min_dist, max_dist, counts=True):
        a, r = C.c_int32(0), C.c_int32(0)
        _check(_lib.ovg_remesh(self._h, min_dist, max_dist, C.byref(a) if counts else None,
                               C.byref(r) if counts else None), "ovg_remesh")
        self.n = _lib.ovg_num_slots(self._h)
        return (a.value, r.value) if counts else None

    def topology_counters(self):
        """The device topology passes' 22 counters (ovg_topology_counters)."""
        out = np.zeros(22, dtype=np.int32)
        _check(_lib.ovg_topology_counters(self._h, _i(out)), "ovg_topology_counters")
        return out

    def enforce_walls(self, t=0.0):
        """enforce_wall_boundaries (src/tangle.c:966-1016): returns (clipped,
        removed)."""
        a, r = C.c_int32(0), C.c_int32(0)
        _check(_lib.ovg_enforce_walls(self._h, t, C.byref(a), C.byref(r)), "ovg_enforce_walls")
        self.n = _lib.ovg_num_slots(self._h)
        return a.value, r.value

    def smooth(self):
        _check(_lib.ovg_smooth(self._h), "ovg_smooth")

    def add_line_z(self, x, y, direction, dl_min, dl_max):
        """add_line_in_zdir (src/vortex_injection.c:267-318): a line from
        wall to wall with pinned ends; returns the number of points."""
        n = C.c_int32(0)
        _check(_lib.ovg_add_line_z(self._h, x, y, direction, dl_min, dl_max, C.byref(n)),
               "ovg_add_line_z")
        self.n = _lib.ovg_num_slots(self._h)
        return n.value

    def add_loop(self, center, direction, r, dl_min, dl_max):
        """add_loop (src/vortex_injection.c:44-103) on the resident tangle;
        returns the number of points added."""
        c = np.ascontiguousarray(center, dtype=np.float64)
        d = np.ascontiguousarray(direction, dtype=np.float64)
        n = C.c_int32(0)
        _check(_lib.ovg_add_loop(self._h, _d(c), _d(d), r, dl_min, dl_max, C.byref(n)),
               "ovg_add_loop")
        self.n = _lib.ovg_num_slots(self._h)
        return n.value

    def velocity_at(self, slots, t=0.0):
        slots = np.ascontiguousarray(slots, dtype=np.int32)
        vs = np.zeros((len(slots), 3))
        vels = np.zeros((len(slots), 3))
        _check(_lib.ovg_velocity_at(self._h, len(slots), _i(slots), t, _d(vs), _d(vels)),
               "ovg_velocity_at")
        return vs, vels

    def upload_forest(self, tangles, params):
        """Many tangles in this context (ovg_upload_forest): tangles[j] in
        its own slot numbering with params[j].  Returns the slot offsets;
        downloads hold tangle j at slots offsets[j] .. offsets[j + 1] - 1."""
        assert len(tangles) == len(params) and tangles
        off = np.zeros(len(tangles) + 1, dtype=np.int32)
        off[1:] = np.cumsum([t.n for t in tangles])
        v = np.ascontiguousarray(np.concatenate([t.vnodes for t in tangles]))
        c = np.ascontiguousarray(np.concatenate([t.conn for t in tangles]), dtype=np.int32)
        s = np.ascontiguousarray(np.concatenate([t.status for t in tangles]), dtype=np.int32)
        arr = (Params * len(params))(*params)
        _check(_lib.ovg_upload_forest(self._h, len(tangles), arr, _i(off), _d(v), _i(c), _i(s)),
               "ovg_upload_forest")
        self.n = int(off[-1])
        self.offsets = off
        return off

    def forest_size(self):
        return _lib.ovg_forest_size(self._h)

    def graph_begin(self):
        """Capture the device work of the following calls (ovg_graph_begin)."""
        _check(_lib.ovg_graph_begin(self._h), "ovg_graph_begin")

    def graph_end(self):
        """Instantiate the captured graph; returns its node count."""
        n = C.c_int32(0)
        _check(_lib.ovg_graph_end(self._h, C.byref(n)), "ovg_graph_end")
        return n.value

    def graph_launch(self, repeats=1):
        _check(_lib.ovg_graph_launch(self._h, int(repeats)), "ovg_graph_launch")

    def graph_release(self):
        _check(_lib.ovg_graph_release(self._h), "ovg_graph_release")

    def velocity_at_noalloc(self, m, d_slots, t=0.0, d_vs=None, d_vels=None):
        """ovg_velocity_at_noalloc on device addresses (ints): d_slots
        int32[m], d_vs / d_vels float64[m, 3] (or None); queued on the
        context's stream, nothing allocated."""
        _check(_lib.ovg_velocity_at_noalloc(self._h, int(m), C.c_void_p(d_slots), t,
                                            C.c_void_p(d_vs) if d_vs else None,
                                            C.c_void_p(d_vels) if d_vels else None),
               "ovg_velocity_at_noalloc")

    def save(self, path, digits=15):
        """The tangle to a reference tangle file (src/utils.c:132-184)."""
        _check(_lib.ovg_save_tangle(self._h, str(path).encode(), digits), "ovg_save_tangle")

    def load(self, path):
        """Replace the tangle by a reference tangle file's (src/utils.c:186-199)."""
        _check(_lib.ovg_load_tangle(self._h, str(path).encode()), "ovg_load_tangle")
        self.n = _lib.ovg_num_slots(self._h)

    def autotune(self, repeats=2):
        """Time the list traversal over points per wave, leaf size and group
        level, then over the launch shape (stack depth, far-field flush
        tile, persistent grid), on the current tangle or forest and keep the
        fastest; returns (rows, best index), rows as dicts."""
        t = Tune()
        _check(_lib.ovg_autotune(self._h, repeats, C.byref(t)), "ovg_autotune")
        rows = [{"points_per_wave": r.points_per_wave, "leaf_size": r.leaf_size,
                 "group_level": r.group_level, "list_stack": r.list_stack,
                 "list_tile": r.list_tile, "list_blocks": r.list_blocks, "ms": r.ms,
                 "pairs": r.pairs}
                for r in t.rows[:t.n_rows]]
        return rows, t.best

    def num_points(self):
        return _lib.ovg_num_points(self._h)

    def line_length(self):
        L = C.c_double(0)
        n = C.c_int32(0)
        _check(_lib.ovg_line_length(self._h, C.byref(L), C.byref(n)), "ovg_line_length")
        return L.value, n.value

    def stats(self):
        s = Stats()
        _check(_lib.ovg_get_stats(self._h, C.byref(s)), "ovg_get_stats")
        return s

    def set_timing(self, on):
        _check(_lib.ovg_set_timing(self._h, int(bool(on))), "ovg_set_timing")

    def synchronize(self):
        _check(_lib.ovg_synchronize(self._h), "ovg_synchronize")

    def marker_record(self, mid):
        """Queue stream marker `mid` behind the work queued so far."""
        _check(_lib.ovg_marker_record(self._h, int(mid)), "ovg_marker_record")

    def marker_elapsed(self, mid0, mid1):
        """Device milliseconds between two recorded markers (waits for mid1)."""
        ms = C.c_double()
        _check(_lib.ovg_marker_elapsed(self._h, int(mid0), int(mid1), C.byref(ms)),
               "ovg_marker_elapsed")
        return ms.value


def loop_points(center, direction, r, dl_min, dl_max):
    """add_loop's points (src/vortex_injection.c:44-103), (n, 3)."""
    c = np.ascontiguousarray(center, dtype=np.float64)
    d = np.ascontiguousarray(direction, dtype=np.float64)
    n = C.c_int32(0)
    _check(_lib.ovg_loop_points(_d(c), _d(d), r, dl_min, dl_max, None, 0, C.byref(n)),
           "ovg_loop_points")
    out = np.zeros((max(n.value, 0), 3))
    _check(_lib.ovg_loop_points(_d(c), _d(d), r, dl_min, dl_max, _d(out), len(out), C.byref(n)),
           "ovg_loop_points")
    return out


def rand48_state(seed):
    """The drand48 state after srand48(seed), in erand48 form."""
    return np.array([0x330E, seed & 0xFFFF, (seed >> 16) & 0xFFFF], dtype=np.uint16)


def inject_loop_draw(t, frequency, box_lo, box_hi, rng, last_injection):
    """inject_loop's schedule (src/vortex_injection.c:711-745) with explicit
    state: rng (uint16[3], advanced in place) and the last injection time.
    Returns (due, center, direction, r, last_injection)."""
    lo = np.ascontiguousarray(box_lo, dtype=np.float64)
    hi = np.ascontiguousarray(box_hi, dtype=np.float64)
    assert rng.dtype == np.uint16 and rng.shape == (3,) and rng.flags.c_contiguous
    last = C.c_double(last_injection)
    c, d, r, due = np.zeros(3), np.zeros(3), C.c_double(0), C.c_int32(0)
    _check(_lib.ovg_inject_loop_draw(t, frequency, _d(lo), _d(hi),
                                     rng.ctypes.data_as(_P(C.c_ushort)), C.byref(last), _d(c),
                                     _d(d), C.byref(r), C.byref(due)), "ovg_inject_loop_draw")
    return bool(due.value), c, d, r.value, last.value


def inject_line_pairs_draw(t, frequency, n_pairs, polarized, box_lo, box_hi, rng,
                           last_injection):
    """inject_line_pairs' schedule (src/vortex_injection.c:760-800) with
    explicit state.  Returns (due, xy of shape (n_pairs, 4) as x1, y1, x2,
    y2, last_injection)."""
    lo = np.ascontiguousarray(box_lo, dtype=np.float64)
    hi = np.ascontiguousarray(box_hi, dtype=np.float64)
    assert rng.dtype == np.uint16 and rng.shape == (3,) and rng.flags.c_contiguous
    last = C.c_double(last_injection)
    xy, due = np.zeros((max(n_pairs, 1), 4)), C.c_int32(0)
    _check(_lib.ovg_inject_line_pairs_draw(t, frequency, n_pairs, int(polarized), _d(lo), _d(hi),
                                           rng.ctypes.data_as(_P(C.c_ushort)), C.byref(last),
                                           _d(xy), C.byref(due)), "ovg_inject_line_pairs_draw")
    return bool(due.value), xy[:n_pairs], last.value


def write_tangle_file(path, t, vels=None, tangents=None, normals=None, digits=15):
    """Reference tangle file of slot arrays (save_tangle, src/utils.c:132-184):
    digits 15 is the reference's %.15g, 17 round-trips exactly."""
    c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    vels, tangents, normals = c(vels), c(tangents), c(normals)
    _check(_lib.ovg_write_tangle_file(str(path).encode(), t.n, _d(np.ascontiguousarray(t.vnodes)),
                                      _d(vels), _d(tangents), _d(normals),
                                      _i(np.ascontiguousarray(t.conn, dtype=np.int32)),
                                      _i(np.ascontiguousarray(t.status, dtype=np.int32)), digits),
           "ovg_write_tangle_file")


def read_tangle_file(path):
    """(Tangle, vels, tangents, normals) of a reference tangle file
    (load_tangle into create_tangle(512), src/utils.c:93-130,186-199)."""
    n = C.c_int(0)
    _check(_lib.ovg_read_tangle_file(str(path).encode(), 0, C.byref(n), None, None, None, None,
                                     None, None), "ovg_read_tangle_file")
    n = n.value
    x, v, tg, nr = (np.zeros((n, 3)) for _ in range(4))
    conn = np.zeros((n, 2), dtype=np.int32)
    st = np.zeros((n, 2), dtype=np.int32)
    m = C.c_int(0)
    _check(_lib.ovg_read_tangle_file(str(path).encode(), n, C.byref(m), _d(x), _d(v), _d(tg),
                                     _d(nr), _i(conn), _i(st)), "ovg_read_tangle_file")
    return Tangle(x, conn, st), v, tg, nr
