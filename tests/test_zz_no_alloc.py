"""The allocation-free entry point ovg_velocity_at_noalloc against
ovg_velocity_at (the reference's update_velocity with tree == NULL,
src/tangle.c:600-726): the same direct sum and compose, so the outputs must
be bit-identical, on two calls in a row (the second with fewer targets, on
the scratch array the first one left behind).  The caller's device buffers
(slots, outputs) are allocated here with the HIP runtime the library is
linked against; the entry point itself must not change the library's
device-memory count."""
import ctypes as C
import sys

import numpy as np
import pytest

import openvort_amd as ov
from openvort_amd import configs

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libopenvort_gpu.so
    loaded (libamdhip64.so.7, already in the process)."""
    H2D, D2H = 1, 2

    def __init__(self):
        ov.lib()  # the library, and with it its HIP runtime, is loaded
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        self.ptrs.append(p.value)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D) == 0
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        self.h.hipDeviceSynchronize()
        for p in self.ptrs:
            self.h.hipFree(p)
        self.ptrs = []


@pytest.fixture
def hip():
    h = Hip()
    yield h
    h.free()


def device_arrays(hip, slots):
    m = len(slots)
    nan = np.full((m, 3), np.nan)
    return hip.put(slots.astype(np.int32)), hip.put(nan), hip.put(nan)


def test_noalloc_equals_velocity_at_on_two_calls(hip):
    t, p, _ = configs.make("c2", rings=60, traverse=ov.TRAVERSE_LIST, precision=ov.PREC_MIXED)
    rng = np.random.default_rng(3)
    with ov.Engine(p) as e:
        e.upload(t)
        for m in (3000, 700):
            slots = np.sort(rng.choice(t.live, m, replace=False)).astype(np.int32)
            say(f"velocity_at_noalloc: {m} targets of {len(t.live)} points")
            e.update_tangle(0.0)
            d_slots, d_vs, d_vels = device_arrays(hip, slots)
            e.synchronize()
            used = ov.device_memory_in_use()
            e.velocity_at_noalloc(m, d_slots, 0.0, d_vs, d_vels)
            e.synchronize()
            assert ov.device_memory_in_use() == used
            got_vs = hip.get(d_vs, (m, 3), np.float64)
            got_vels = hip.get(d_vels, (m, 3), np.float64)
            want_vs, want_vels = e.velocity_at(slots, 0.0)
            assert np.isfinite(got_vels).all()
            assert np.array_equal(got_vs, want_vs)
            assert np.array_equal(got_vels, want_vels)


def test_noalloc_limits_and_bad_slots(hip):
    t, p, _ = configs.make("c1", rings=6)
    with ov.Engine(p) as e:
        e.upload(t)
        e.update_tangents_normals()
        say("velocity_at_noalloc: size limit, a bad slot")
        with pytest.raises(ov.OvgError) as ei:
            e.velocity_at_noalloc(ov.NOALLOC_MAX_TARGETS + 1, 1, 0.0, None, None)
        assert ei.value.code == ov.ELIMIT
        empty = int(np.nonzero(t.status[:, 0] == ov.EMPTY)[0][0])
        slots = np.array([t.live[0], empty, t.live[5]], dtype=np.int32)
        d_slots, d_vs, d_vels = device_arrays(hip, slots)
        e.velocity_at_noalloc(3, d_slots, 0.0, d_vs, d_vels)
        with pytest.raises(ov.OvgError) as ei:
            e.synchronize()
        assert ei.value.code == ov.EINVAL
        got = hip.get(d_vels, (3, 3), np.float64)
        assert np.isnan(got[1]).all() and np.isfinite(got[[0, 2]]).all()
        want = e.velocity_at(slots[[0, 2]], 0.0)[1]
        assert np.array_equal(got[[0, 2]], want)
        e.synchronize()  # the flag was reported once and cleared
