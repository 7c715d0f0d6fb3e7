"""Drop-in check: the reference's own tangle API and step loop
(oracle/dropin_driver.c, the loop of src/main.c:117-141 without injection
and walls) linked two ways by oracle/Makefile -- against the reference
objects (CPU) and against libopenvort_dropin.so for the hot path (GPU) --
run on the same tangle.  Both links are built in the container that holds
the reference sources and travel with the tree.

Tolerances: the GPU hot path is FP64 with the reference's tree and
acceptance test (velocities within 1e-10 relative L2 per evaluation, from
summation order); topology (reconnection, small loops, remesh) bit-exact;
positions after five steps within 1e-12 cm.
"""
import os
import subprocess

import numpy as np
import pytest

import openvort_amd as ov
from openvort_amd import configs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ovref_driver")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "ovref_dropin")


def write_input(path, t, p):
    with open(path, "wb") as f:
        f.write(np.int32(t.n).tobytes())
        f.write(bytes(p))
        f.write(np.ascontiguousarray(t.vnodes).tobytes())
        f.write(np.ascontiguousarray(t.conn, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(t.status, dtype=np.int32).tobytes())


def read_output(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    o = 4
    x = np.frombuffer(raw[o:o + 24 * n], dtype=np.float64).reshape(n, 3)
    o += 24 * n
    conn = np.frombuffer(raw[o:o + 8 * n], dtype=np.int32).reshape(n, 2)
    o += 8 * n
    st = np.frombuffer(raw[o:o + 8 * n], dtype=np.int32).reshape(n, 2)
    o += 8 * n
    vels = np.frombuffer(raw[o:o + 24 * n], dtype=np.float64).reshape(n, 3)
    return x, conn, st, vels


def run(exe, inp, out, steps, dt, smooth_every):
    print(f"[gpu] {os.path.basename(exe)}: {steps} steps", flush=True)
    r = subprocess.run([exe, inp, out, str(steps), repr(dt), str(smooth_every), "5e-4",
                        "1.5e-3"], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, OMP_NUM_THREADS="8"))
    assert r.returncode == 0, (exe, r.returncode, r.stderr[-2000:])
    return read_output(out)


def test_dropin_matches_reference_driver(tmp_path):
    assert os.path.exists(DRIVER) and os.path.exists(DROPIN), \
        "oracle/_ref drivers missing: build them with make -C oracle ref"
    t, p, dt = configs.make("c1", method=ov.METHOD_BH)
    p.vn.type = ov.FLOW_SIMPLE
    p.vn.strength = 1.0
    p.vn.direction[0] = 1.0
    inp = str(tmp_path / "in.bin")
    write_input(inp, t, p)
    cpu = run(DRIVER, inp, str(tmp_path / "cpu.bin"), 5, dt, 0)
    gpu = run(DROPIN, inp, str(tmp_path / "gpu.bin"), 5, dt, 0)
    assert cpu[0].shape == gpu[0].shape
    assert np.array_equal(cpu[2], gpu[2])
    assert np.array_equal(cpu[1], gpu[1])
    live = np.nonzero(cpu[2][:, 0] != -1)[0]
    assert len(live) > 1000
    assert np.abs(cpu[0][live] - gpu[0][live]).max() < 1e-12
    v_ref = cpu[3][live]
    assert np.linalg.norm(gpu[3][live] - v_ref) / np.linalg.norm(v_ref) < 1e-10


def test_dropin_runs_with_smoothing(tmp_path):
    """smooth() exists only in the drop-in link (the reference's needs FFTW);
    the loop runs with it every second step and keeps a finite state."""
    assert os.path.exists(DROPIN)
    t, p, dt = configs.make("c1", method=ov.METHOD_BH)
    inp = str(tmp_path / "in.bin")
    write_input(inp, t, p)
    x, conn, st, vels = run(DROPIN, inp, str(tmp_path / "gpu.bin"), 4, dt, 2)
    live = st[:, 0] != -1
    assert live.sum() > 1000
    assert np.isfinite(x[live]).all() and np.isfinite(vels[live]).all()
