"""Tangle files (src/utils.c:68-199): the build's writer against the
reference's save_tangle byte for byte, its reader against the reference's
load_tangle, and exact round trips at 17 digits.  CPU only."""
import numpy as np

import openvort_amd as ov
from openvort_amd import configs
from oracle.ovoracle import ref_load


def ref_state(oracle_ref, t, p):
    rt = oracle_ref.tangle(t, p)
    rt.update_tangle(0.0)  # nonzero velocities, tangents, normals
    return rt


def test_writer_matches_reference_bytes(tmp_path, ref):
    t, p, _ = configs.make("c1", rings=8)
    rt = ref_state(ref, t, p)
    o = rt.get()
    rt.save(tmp_path / "ref.dat")
    ov.write_tangle_file(tmp_path / "ours.dat", ov.Tangle(o["vnodes"], o["conn"], rt.status()),
                         o["vels"], o["tangents"], o["normals"], digits=15)
    a = (tmp_path / "ref.dat").read_bytes()
    b = (tmp_path / "ours.dat").read_bytes()
    assert len(a) > 10000 and a == b


def test_reader_matches_reference_load(tmp_path, ref):
    t, p, _ = configs.make("c1", rings=8)
    rt = ref_state(ref, t, p)
    rt.save(tmp_path / "ref.dat")
    back = ref_load(ref, tmp_path / "ref.dat", p)
    want = back.get()
    got, v, tg, nr = ov.read_tangle_file(tmp_path / "ref.dat")
    assert got.n == back.n
    st = back.status()
    assert np.array_equal(got.status, st) and np.array_equal(got.conn, want["conn"])
    live = st[:, 0] != -1
    for a, b in ((got.vnodes, want["vnodes"]), (v, want["vels"]), (tg, want["tangents"]),
                 (nr, want["normals"])):
        assert np.array_equal(a[live], b[live])


def test_round_trip_17_digits_and_capacity(tmp_path):
    t, _, _ = configs.make("c1", rings=30)  # > 512 points: the slot count doubles on load
    rng = np.random.default_rng(3)
    v = rng.normal(size=(t.n, 3))
    ov.write_tangle_file(tmp_path / "t.dat", t, v, None, None, digits=17)
    got, v2, tg, nr = ov.read_tangle_file(tmp_path / "t.dat")
    n512 = 512
    while n512 < t.n:
        n512 *= 2
    assert got.n == n512
    live = t.live
    assert np.array_equal(got.vnodes[live], t.vnodes[live])
    assert np.array_equal(v2[live], v[live])
    assert np.array_equal(got.conn[:t.n], t.conn) and np.array_equal(got.status[:t.n], t.status)
    assert not tg.any() and not nr.any()
