"""The C-ABI library builds, loads and exports every declared entry point;
host-side helpers (parameter defaults, image tables, generator) behave.
No GPU needed."""
import os
import re

import ctypes as C
import numpy as np
import pytest

import openvort_amd as ov
from conftest import ROOT


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(ovg_[a-z0-9_]+)\s*\(", txt))


def test_every_declared_symbol_is_exported():
    names = declared("openvort_gpu.h")
    assert len(names) > 25
    lib = C.CDLL(ov.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_serve_symbols_exported_and_structs_match():
    from openvort_amd import serve
    txt = open(os.path.join(ROOT, "include", "openvort_serve.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(ovs_[a-z0-9_]+)\s*\(", txt))
    assert {"ovs_run", "ovs_job_bytes"} <= names
    lib = C.CDLL(ov.LIB_PATH)
    assert all(hasattr(lib, n) for n in names)
    # field layouts restated in Python: 8-byte aligned, sizes as declared
    assert C.sizeof(serve.Job) == 8 + 4 * 2 + 8 * 9 + 4 * 6 + 8 + 4 * 2
    assert C.sizeof(serve.Report) % 8 == 0


def test_c5_job_stream_is_seeded():
    from openvort_amd import serve
    a = serve.c5_jobs(7, 12)
    b = serve.c5_jobs(7, 12)
    assert [(j.seed, j.n_rings, j.steps, j.arrival_s) for j in a] == \
        [(j.seed, j.n_rings, j.steps, j.arrival_s) for j in b]
    arr = [j.arrival_s for j in a]
    assert arr == sorted(arr) and arr[0] > 0
    assert all(50 <= j.steps <= 500 and 0.1 <= j.D <= 0.2 for j in a)


def test_dropin_exports_reference_prototypes():
    if not os.path.exists(ov.DROPIN_PATH):
        pytest.skip("libopenvort_dropin.so not built")
    txt = open(os.path.join(ROOT, "include", "openvort_dropin.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"^\s*(?:void|int)\s+\**([a-z_0-9]+)\s*\(", txt, flags=re.M))
    assert {"update_velocities", "rk4_step", "update_tangents_normals", "reconnect",
            "smooth"} <= names
    out = os.popen(f"nm -D --defined-only {ov.DROPIN_PATH}").read()
    for n in names:
        assert re.search(rf"\bT {n}$", out, flags=re.M), n


def test_struct_sizes_match():
    assert ov.lib().ovg_params_sizeof() == C.sizeof(ov.Params)
    assert ov.lib().ovg_stats_sizeof() == C.sizeof(ov.Stats)
    assert ov.lib().ovg_tune_sizeof() == C.sizeof(ov.Tune)


def test_defaults_follow_reference_globals():
    p = ov.default_params(0.1)
    # src/initial_conditions.c:12-72
    assert p.kappa == 9.97e-4 and p.vortex_width == 1e-8
    assert p.alpha == 0.034 and p.alpha_p == 1.383e-2
    assert p.theta == 0.4 and p.rec_dist == 4e-4
    assert p.freq_to_cutoff == 4.0 / 5
    assert p.rec_angle == 5 * np.pi / 180.0
    assert p.n_images == 26


@pytest.mark.parametrize("n", [6, 18, 26])
def test_image_tables_are_the_reference_tables(n):
    p = ov.default_params(1.0)
    ov.set_images(p, n)
    got = np.array(p.images)[:n]
    from oracle.ovoracle import RefLib, ref_available
    if ref_available():
        assert np.array_equal(got, RefLib().image_table(n))
    # closed form: all nonzero shifts with <= 1/2/3 nonzero entries, lexicographic
    nz = {6: 1, 18: 2, 26: 3}[n]
    want = [s for s in np.ndindex(3, 3, 3)
            if 0 < sum(v != 1 for v in s) <= nz]
    assert np.array_equal(got, np.array(want) - 1)


def test_bad_image_count_is_rejected():
    p = ov.default_params(1.0)
    assert ov.lib().ovg_params_set_periodic_images(C.byref(p), 7) == ov.EINVAL
    assert b"0, 6, 18 or 26" in ov.lib().ovg_last_error()


def test_generator_is_deterministic_and_closed():
    a = ov.generate_rings(7, 5, 0.1, 0.01, 0.03, 1e-3)
    b = ov.generate_rings(7, 5, 0.1, 0.01, 0.03, 1e-3)
    assert np.array_equal(a.vnodes, b.vnodes)
    live = a.live
    assert np.all(a.vnodes[live] >= 0) and np.all(a.vnodes[live] < 0.1)
    f, r = a.conn[live, 0], a.conn[live, 1]
    assert np.array_equal(a.conn[f, 1], live) and np.array_equal(a.conn[r, 0], live)
    # empty tail
    assert np.all(a.status[len(live):, 0] == ov.EMPTY)
    assert a.n >= len(live) and (a.n & (a.n - 1)) == 0


def test_generator_point_count_rule():
    # n = floor(2 pi R / dxi) per ring (add_loop's rule with dl_min+dl_max=2 dxi)
    t = ov.generate_rings(3, 1, 1.0, 0.1, 0.1, 1e-2)
    assert len(t.live) == int(np.floor(2 * np.pi * 0.1 / 1e-2))
    # all points at distance R from their centroid
    x = t.vnodes[t.live]
    c = x.mean(axis=0)
    assert np.allclose(np.linalg.norm(x - c, axis=1), 0.1, rtol=1e-12)


def test_rings_for_points():
    k = ov.rings_for_points(11, 5000, 0.1, 0.005, 0.03, 8e-4)
    t = ov.generate_rings(11, k, 0.1, 0.005, 0.03, 8e-4)
    t2 = ov.generate_rings(11, k - 1, 0.1, 0.005, 0.03, 8e-4)
    assert len(t.live) >= 5000 > len(t2.live)


def test_create_without_gpu_fails_loudly():
    h = C.c_void_p()
    rc = ov.lib().ovg_create(C.byref(h), 0)
    if rc == ov.OK:
        ov.lib().ovg_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == ov.EDEVICE
    assert ov.lib().ovg_last_error()
