"""The compiled reference driver (oracle/_ref/ovref_driver, the CPU side of
the drop-in comparison) reads exactly this build's ovg_params layout: it runs
a short C1 loop, and an input written with another layout is refused instead
of handing the reference shifted links."""
import os
import subprocess

import numpy as np
import pytest

import openvort_amd as ov
from openvort_amd import configs
from test_dropin import DRIVER, read_output, write_input

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER),
                                reason="oracle/_ref not built (make -C oracle ref)")


def run_driver(inp, out, steps, dt):
    return subprocess.run([DRIVER, inp, out, str(steps), repr(dt), "0", "5e-4", "1.5e-3"],
                          capture_output=True, text=True, timeout=120,
                          env=dict(os.environ, OMP_NUM_THREADS="4"))


def test_reference_driver_runs(tmp_path):
    t, p, dt = configs.make("c1", method=ov.METHOD_BH)
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(inp, t, p)
    r = run_driver(inp, out, 1, dt)
    assert r.returncode == 0, r.stderr[-500:]
    x, conn, st, vels = read_output(out)
    live = st[:, 0] != -1
    assert live.sum() > 1000 and np.isfinite(x[live]).all()


def test_reference_driver_refuses_other_layout(tmp_path):
    t, p, dt = configs.make("c1", method=ov.METHOD_BH)
    bad = str(tmp_path / "bad.bin")
    with open(bad, "wb") as f:
        f.write(np.int32(t.n).tobytes())
        f.write(bytes(p)[:-8])  # the layout before the last field
        f.write(np.ascontiguousarray(t.vnodes).tobytes())
        f.write(np.ascontiguousarray(t.conn, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(t.status, dtype=np.int32).tobytes())
    r = run_driver(bad, str(tmp_path / "x.bin"), 1, dt)
    assert r.returncode == 2 and "expected" in r.stderr
