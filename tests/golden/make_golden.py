"""Write the golden vectors of tests/golden/*.npz from the compiled reference.

Run in the container that holds the reference sources, after
``make -C oracle ref`` has compiled them into oracle/_ref/libovref.so:

    python tests/golden/make_golden.py

Every fixture stores the generator arguments, a SHA-256 of the generated
input state (so a test can tell when the generator changed), the
parameters, and the reference's outputs (strict IEEE build, -O2
-ffp-contract=off, zero-fill malloc).  Inputs come from the build's seeded
ring generator (ovg_generate_rings); the reference's own RNG is wall-clock
seeded (src/main.c:80) and its tangle files are written with %.15g
(src/utils.c:70), so neither can supply reproducible inputs.
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import openvort_amd as ov  # noqa: E402
from openvort_amd import configs  # noqa: E402
from oracle.ovoracle import RefLib  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def state_hash(t):
    h = hashlib.sha256()
    for a in (t.vnodes, t.conn, t.status):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def params_array(p):
    return np.frombuffer(bytes(p), dtype=np.uint8).copy()


def save(name, **arrays):
    path = os.path.join(OUT, name)
    np.savez_compressed(path, **arrays)
    print(f"{name}: {os.path.getsize(path) / 1024:.1f} KiB")


def remesh_fixture(R):
    # C1 tangle after a reconnect pass with rec_dist 6e-3 and small-loop
    # removal, then remesh with dl_min = 1.05 dxi, dl_max = 1.5 dxi: both
    # removals and insertions (with slot reuse and capacity doubling)
    t, p, dt = configs.make("c1")
    p.rec_dist = 6e-3
    rt = R.tangle(t, p)
    rt.reconnect(0.0)
    rt.eliminate_small_loops(5)
    before = rt.get()
    st_before = rt.status()
    rt.remesh(1.05e-3, 1.5e-3)
    after = rt.get()
    st_after = rt.status()
    save("c1_remesh.npz", config=np.array("c1"), rings=16, sha=np.array(state_hash(t)),
         params=params_array(p), min_dist=1.05e-3, max_dist=1.5e-3,
         x_before=before["vnodes"], conn_before=before["conn"], status_before=st_before,
         n_after=rt.n, x_after=after["vnodes"], conn_after=after["conn"], status_after=st_after)


def line_length(x, conn, st, D):
    """Sum over live slots of |seg_pwrap(x_k, x_fwd(k))| in a cube of side D."""
    live = np.nonzero((st[:, 0] != -1) & (conn[:, 0] >= 0))[0]
    d = x[conn[live, 0]] - x[live]
    d -= D * np.round(d / D)
    return float(np.linalg.norm(d, axis=1).sum())


def c1_100_fixture(R):
    # C1, direct sum with 26 images, 100 steps of the driver order
    # (src/main.c:127-138) without smoothing (src/vortex_smooth.c needs
    # FFTW): line length, live points and reconnections every 10 steps
    t, p, dt = configs.make("c1")
    D = float(p.box_hi[0] - p.box_lo[0])
    dl_min, dl_max = configs.remesh_bounds("c1")
    rt = R.tangle(t, p)
    tm, recs = 0.0, 0
    L, npts, nrec = [], [], []
    for step in range(100):
        recs += rt.reconnect(tm)
        rt.eliminate_small_loops(5)
        rt.enforce_periodic()
        rt.remesh(dl_min, dl_max)
        rt.eliminate_small_loops(5)
        rt.update_tangle(tm)
        rt.rk4_step(tm, dt)
        rt.enforce_periodic()
        tm += dt
        if (step + 1) % 10 == 0:
            o = rt.get()
            st = rt.status()
            L.append(line_length(o["vnodes"], o["conn"], st, D))
            npts.append(int((st[:, 0] != -1).sum()))
            nrec.append(recs)
            print(f"step {step + 1}: L {L[-1]:.10f} points {npts[-1]} reconnections {recs}",
                  flush=True)
    o = rt.get()
    save("c1_100steps.npz", config=np.array("c1"), rings=16, sha=np.array(state_hash(t)),
         params=params_array(p), dt=dt, dl_min=dl_min, dl_max=dl_max,
         line_length=np.array(L), points=np.array(npts), reconnections=np.array(nrec),
         x_final=o["vnodes"], conn_final=o["conn"], status_final=rt.status())


def main():
    R = RefLib()
    if len(sys.argv) > 1:
        for name in sys.argv[1:]:
            globals()[name + "_fixture"](R)
        return
    remesh_fixture(R)
    c1_100_fixture(R)
    # C1 (16 rings, ~2k points): direct sum with 26 images, full outputs
    t, p, dt = configs.make("c1")
    live = t.live
    rt = R.tangle(t, p)
    rt.update_tangle(0.0)
    o = rt.get()
    rt.rk4_step(0.0, dt)
    rt.enforce_periodic()
    o2 = rt.get()
    save("c1_direct.npz", config=np.array("c1"), rings=16, sha=np.array(state_hash(t)),
         params=params_array(p), live=live, dt=dt,
         vs=o["vs"][live], vels=o["vels"][live], tangents=o["tangents"][live],
         normals=o["normals"][live], x_after_rk4=o2["vnodes"][live])

    # C1 with Barnes-Hut theta = 0.4, counterflow on, two steps
    t, p, dt = configs.make("c1", method=ov.METHOD_BH)
    p.vn.type = ov.FLOW_SIMPLE
    p.vn.strength = 1.0
    p.vn.direction[0] = 1.0
    rt = R.tangle(t, p)
    rt.update_tangle(0.0)
    o = rt.get()
    xs = []
    tm = 0.0
    for _ in range(2):
        rt.update_tangle(tm)
        rt.rk4_step(tm, dt)
        rt.enforce_periodic()
        tm += dt
        xs.append(rt.get()["vnodes"][live])
    save("c1_bh.npz", config=np.array("c1"), rings=16, sha=np.array(state_hash(t)),
         params=params_array(p), live=live, dt=dt, vs=o["vs"][live], vels=o["vels"][live],
         x_after_1=xs[0], x_after_2=xs[1])

    # C2 (200 rings, ~27.5k points): BH theta 0.4 velocities of a 2048-point
    # subsample, and the FP64 direct sum (the oracle of record) at the same
    # points through update_velocity(tree = NULL)
    t, p, dt = configs.make("c2", method=ov.METHOD_BH)
    live = t.live
    rng = np.random.default_rng(2024)
    sub = np.sort(rng.choice(live, size=2048, replace=False))
    rt = R.tangle(t, p)
    rt.update_tangle(0.0)
    o = rt.get()
    dvs = np.zeros((len(sub), 3))
    dvels = np.zeros((len(sub), 3))
    for i, k in enumerate(sub):
        dvs[i], dvels[i] = rt.velocity_one(int(k), 0.0)
    save("c2_sub.npz", config=np.array("c2"), rings=200, sha=np.array(state_hash(t)),
         params=params_array(p), sub=sub, bh_vs=o["vs"][sub], bh_vels=o["vels"][sub],
         direct_vs=dvs, direct_vels=dvels)

    # Reconnection: C1 tangle with rec_dist 4e-3 (many candidates), the
    # candidate list and the links after one reconnect pass
    t, p, dt = configs.make("c1")
    p.rec_dist = 4e-3
    rt = R.tangle(t, p)
    nrec = rt.reconnect(0.0)
    conn_after = rt.get()["conn"]
    save("c1_reconnect.npz", config=np.array("c1"), rings=16, sha=np.array(state_hash(t)),
         params=params_array(p), reconnections=nrec, conn_after=conn_after)


if __name__ == "__main__":
    main()
