#!/usr/bin/env python3
"""RK4 point-steps/s of the vortex-filament hot path on one MI355X.

Workload (default "c3"): 4000 seeded rings in the periodic cube D = 0.2 cm,
radii U[0.005, 0.05] cm, point spacing 5e-4 cm (1.39M points), T = 1.65 K
(alpha 0.11, alpha' 0.014), counterflow v_ns = 1.5 cm/s, Barnes-Hut theta 0.4
with the 26 periodic images, dt = 5e-5 s.  One step follows the reference
driver's order (src/main.c:127-138): reconnect, small-loop removal,
smooth (every 10th step), periodic wrap, remesh (dl_min = dxi/2,
dl_max = 1.5 dxi), small-loop removal, update_tangle, rk4_step (3 more
evaluations), periodic wrap.  State stays on the device between steps; the timed region
starts with the tangle resident in HBM.

The fast path is mixed precision (FP32 far-field pairs, FP64 sums,
positions, LIA and RK4) with the interaction-list traversal: 16 targets
per wavefront, groups bounded by one octree cell, no buckets.  The plain
path is the reference's FP64 split-until-one tree with per-point
traversal, timed on the same tangle beside it.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK4 point-steps/s at 1.4M pts θ=0.4; p50/p99 step latency vs plain FP64"
FLOPS_PER_PAIR = 44  # 41 add/sub/mul + 2 sqrt + 1 div per segment or monopole pair
PEAK_FP32 = 157.3e12  # MI355X vector FP32, FLOP/s
PEAK_FP64 = 78.6e12   # MI355X vector FP64, FLOP/s


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_steps(e, n_steps, t0, dt, step0, smooth_every, reconnect, remesh=None):
    """Run n_steps device-resident steps in the order of src/main.c:127-138
    without synchronising between them (the host queues a step while the GPU
    finishes the previous one, as a driver loop would); returns per-step
    seconds: device time between the completions of consecutive steps, from
    stream markers (the first from a marker queued before it)."""
    tm = t0
    e.marker_record(0)
    for s in range(n_steps):
        k = step0 + s
        # no count pointers: with device topology the passes queue without
        # waiting for the host
        if reconnect:
            e.reconnect(tm, counts=False)
            e.eliminate_small_loops(5, counts=False)
        if smooth_every and k % smooth_every == 0:
            e.smooth()
        e.enforce_periodic()
        if remesh:
            e.remesh(*remesh, counts=False)
            e.eliminate_small_loops(5, counts=False)
        e.update_tangle(tm)
        e.rk4_step(tm, dt)
        e.enforce_periodic()
        e.marker_record(s + 1)
        tm += dt
    lat = [e.marker_elapsed(s, s + 1) * 1e-3 for s in range(n_steps)]
    return lat, tm


def accuracy(ov, t, p_fast, p_plain, n_sample, seed=7):
    """Velocity error on the initial tangle: both paths against the FP64
    direct sum (ovg_velocity_at, the reference's update_velocity with
    tree == NULL) at a seeded sample of points (relative L2 over the
    sample), and the fast path's precision error alone: mixed against FP64
    with the same tree and acceptance, over every point."""
    live = t.live
    rng = np.random.default_rng(seed)
    sub = np.sort(rng.choice(live, size=min(n_sample, len(live)), replace=False))
    p64 = ov.Params.from_buffer_copy(p_fast)
    p64.precision = ov.PREC_FP64
    out = {}
    direct = None
    for name, q in (("fast", p_fast), ("fast_fp64", p64), ("plain_fp64", p_plain)):
        e = ov.Engine(q)
        e.upload(t)
        e.update_tangle(0.0)
        out[name] = e.download()["vs"]
        if direct is None:
            direct = e.velocity_at(sub, 0.0)[0]
        e.close()
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return {"sample_points": int(len(sub)), "field": "vs (Biot-Savart + LIA + imposed v_s)",
            "fast_vs_direct": rel(out["fast"][sub], direct),
            "fast_fp64_vs_direct": rel(out["fast_fp64"][sub], direct),
            "plain_fp64_vs_direct": rel(out["plain_fp64"][sub], direct),
            "fast_vs_plain_fp64": rel(out["fast"][sub], out["plain_fp64"][sub]),
            "mixed_vs_fp64_same_tree": rel(out["fast"][live], out["fast_fp64"][live]),
            "mixed_vs_fp64_points": int(len(live))}


def serve_bench(args):
    """The serving stream (or the batched ensemble) through the request loop:
    fast path with several workers against the plain FP64 path one job at a
    time (FIFO), both under saturation on the same seeded jobs; then the fast
    path once more with the jobs' Poisson arrival times for latencies."""
    from openvort_amd import serve
    make = (lambda fast: serve.c5_jobs(n=args.jobs, steps_scale=args.steps_scale, fast=fast)) \
        if args.workload == "c5" else \
        (lambda fast: serve.c4_jobs(n=args.jobs, steps=max(1, int(round(args.steps_scale * 100))),
                                    fast=fast))
    fast_jobs = make(True)
    import openvort_amd as ov
    for j in fast_jobs:
        j.topology = ov.TOPO_DEVICE if args.topology == "device" else ov.TOPO_HOST
    pts = sum(ov_points(j) for j in fast_jobs)
    log(f"{args.workload}: {len(fast_jobs)} jobs, {pts} points at admission")
    rep, res = serve.run(fast_jobs, workers=args.workers, max_active=args.max_active,
                         cap_bytes=args.hbm_cap, realtime=False)
    log(f"fast: {rep['point_steps_per_s']:.4g} point-steps/s, {rep['wall_s']:.1f} s")
    plain = None
    if args.plain_steps > 0:
        prep, _ = serve.run(make(False), workers=1, max_active=1, cap_bytes=args.hbm_cap,
                            realtime=False)
        plain = {"value": prep["point_steps_per_s"], "wall_s": prep["wall_s"],
                 "p50_step_ms": prep["step_p50_ms"], "p99_step_ms": prep["step_p99_ms"],
                 "path": "FP64, reference split-until-one tree, per-point traversal, "
                         "one job at a time in arrival order"}
        log(f"plain: {plain['value']:.4g} point-steps/s, {prep['wall_s']:.1f} s")
    live = None
    if args.workload == "c5":
        lrep, _ = serve.run(fast_jobs, workers=args.workers, max_active=args.max_active,
                            cap_bytes=args.hbm_cap, realtime=True)
        live = {"arrivals": "Poisson, 20 jobs/s", "point_steps_per_s": lrep["point_steps_per_s"],
                "p50_step_ms": lrep["step_p50_ms"], "p99_step_ms": lrep["step_p99_ms"],
                "p50_job_s": lrep["job_p50_s"], "p99_job_s": lrep["job_p99_s"],
                "wall_s": lrep["wall_s"]}
    out = {
        "metric": f"RK4 point-steps/s, {args.workload} job stream, batched mixed-precision BH "
                  "serving vs plain FP64 one job at a time",
        "value": rep["point_steps_per_s"], "unit": "point-steps/s", "n_gpus": 1,
        "steps": rep["steps_total"], "warmup": 0, "ms_per_step": rep["step_p50_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": (f"{args.workload}: first {len(fast_jobs)} seeded jobs, steps "
                                f"scaled by {args.steps_scale}, all admitted at t = 0"),
                   "topology": args.topology,
                   "workers": args.workers, "max_active": args.max_active,
                   "hbm_cap_bytes": args.hbm_cap},
        "p50_step_ms": rep["step_p50_ms"], "p99_step_ms": rep["step_p99_ms"],
        "p50_job_s": rep["job_p50_s"], "p99_job_s": rep["job_p99_s"],
        "jobs_done": rep["jobs_done"], "jobs_failed": rep["jobs_failed"],
        "peak_active": rep["peak_active"],
        "plain_fp64": plain,
        "speedup_vs_plain": rep["point_steps_per_s"] / plain["value"] if plain else None,
        "with_arrivals": live,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def forest_members(n, seed=0xC0FFEE04):
    """The batched ensemble's tangles and parameters (openvort_amd.serve
    c4_jobs: sizes log-uniform 2k-100k, D = 0.1 cm, dxi = 8e-4 cm, v_ns
    U[0.5, 2], 1.6 K or 1.9 K), mixed precision list traversal."""
    import openvort_amd as ov
    from openvort_amd import configs, serve
    out = []
    for j in serve.c4_jobs(seed=seed, n=n, steps=1):
        t = ov.generate_rings(j.seed, j.n_rings, j.D, j.r_min, j.r_max, j.dxi)
        p = configs.params(j.D, j.dxi, (j.alpha, j.alpha_p), j.v_ns, precision=ov.PREC_MIXED,
                           leaf_size=1, traverse=ov.TRAVERSE_LIST)
        p.points_per_wave = 16
        out.append((t, p))
    return out


def forest_bench(args):
    """The batched ensemble as one forest: every step's update_tangle +
    rk4_step + wrap of all tangles in one captured HIP graph (one tree build
    and one list launch per evaluation for all of them), replayed; against
    the plain FP64 path (reference tree, per-point traversal) one tangle at
    a time, and the same mixed path one tangle at a time, on the same
    steps (no topology passes: the ensemble's step is the evaluation and
    the RK4 update)."""
    import openvort_amd as ov
    n_jobs = args.jobs if args.jobs != 24 else 64
    ms = forest_members(n_jobs)
    n_pts = sum(len(t.live) for t, _ in ms)
    dt = 5e-5
    log(f"c4 forest: {len(ms)} tangles, {n_pts} points")

    def step(e, tm):
        e.update_tangle(tm)
        e.rk4_step(tm, dt)
        e.enforce_periodic()

    e = ov.Engine(ms[0][1])
    e.upload_forest([t for t, _ in ms], [p for _, p in ms])
    tune = None
    if not args.no_autotune:
        rows, best = e.autotune(repeats=2)
        tune = dict(rows[best], candidates=len(rows),
                    shapes=[[r["list_stack"], r["list_tile"], r["list_blocks"], round(r["ms"], 3)]
                            for r in rows[-18:]])
        log(f"forest autotune: {tune}")
    step(e, 0.0)
    # the traversal's launch time and pairs: two eager steps with timing
    e.stats()
    e.set_timing(True)
    for s in range(2):
        step(e, (1 + s) * dt)
    e.synchronize()
    st = e.stats()
    e.set_timing(False)
    launches = max(st.bs_launches, 1)
    bs_ms = st.ms_bs / launches
    pairs = st.interactions / launches
    # eager fused steps
    e.synchronize()
    a = time.perf_counter()
    for s in range(args.steps):
        step(e, (3 + s) * dt)
    e.synchronize()
    eager_ms = (time.perf_counter() - a) / args.steps * 1e3
    # one step captured, replayed
    e.graph_begin()
    step(e, 0.0)
    nodes = e.graph_end()
    e.graph_launch(args.warmup)
    e.synchronize()
    e.marker_record(0)
    a = time.perf_counter()
    for s in range(args.steps):
        e.graph_launch(1)
        e.marker_record(s + 1)
    e.synchronize()
    wall = time.perf_counter() - a
    lat = [e.marker_elapsed(s, s + 1) * 1e-3 for s in range(args.steps)]
    e.close()
    value = n_pts * args.steps / wall
    log(f"forest graph: {wall / args.steps * 1e3:.1f} ms/step ({nodes} nodes), eager "
        f"{eager_ms:.1f} ms/step, {value:.4g} point-steps/s")

    def one_at_a_time(fp64, steps):
        total, pts_steps = 0.0, 0.0
        for t, p in ms:
            q = ov.Params.from_buffer_copy(p)
            if fp64:
                q.precision, q.traverse, q.leaf_size = ov.PREC_FP64, ov.TRAVERSE_POINT, 1
            f = ov.Engine(q)
            f.upload(t)
            step(f, 0.0)
            f.synchronize()
            a = time.perf_counter()
            for s in range(steps):
                step(f, (1 + s) * dt)
            f.synchronize()
            total += time.perf_counter() - a
            pts_steps += len(t.live) * steps
            f.close()
        return pts_steps / total, total

    plain_v, plain_s = one_at_a_time(True, args.plain_steps)
    log(f"plain FP64 one at a time: {plain_v:.4g} point-steps/s")
    sep_v, sep_s = one_at_a_time(False, args.plain_steps)
    log(f"mixed one at a time: {sep_v:.4g} point-steps/s")
    achieved = pairs * FLOPS_PER_PAIR / (bs_ms * 1e-3) if bs_ms > 0 else 0.0
    out = {
        "metric": "RK4 point-steps/s, c4 batched ensemble as one forest (one tree and one "
                  "list launch per evaluation, HIP-graph step) vs plain FP64 one tangle at a time",
        "value": value, "unit": "point-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": (f"c4: {len(ms)} seeded tangles, {n_pts} points, sizes "
                                f"log-uniform 2k-100k, BH theta 0.4, 26 periodic images, "
                                f"dt {dt}; step = update_tangle + rk4_step + wrap"),
                   "tangles": len(ms), "points": n_pts, "autotune": tune, "graph_nodes": nodes},
        "p50_step_ms": pct(lat, 50), "p99_step_ms": pct(lat, 99),
        "eager_ms_per_step": eager_ms,
        "plain_fp64": {"value": plain_v, "seconds": plain_s, "steps": args.plain_steps,
                       "path": "FP64, reference split-until-one tree, per-point traversal, "
                               "one tangle at a time"},
        "mixed_one_at_a_time": {"value": sep_v, "seconds": sep_s, "steps": args.plain_steps},
        "speedup_vs_plain": value / plain_v,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "traffic": None,
                     "kernel": "k_bh_list<mixed> (forest)", "avg_launch_ms": bs_ms,
                     "pairs_per_launch": pairs},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def ov_points(job):
    import openvort_amd as ov
    return ov.lib().ovg_generate_rings(job.seed, job.n_rings, job.D, job.r_min, job.r_max,
                                       job.dxi, 0, None, None, None)


def pct(v, q):
    return float(np.percentile(np.asarray(v) * 1e3, q))


def ref_bench(t, p, dt, threads, steps=1, reconnect=0, mode="steps", timeout=1200):
    """One run of the reference timer (oracle/_ref/ovref_bench, the
    reference built with its Makefile's flags) on tangle t; its JSON line,
    or None."""
    from oracle.ovoracle import REF_BENCH
    if not os.path.exists(REF_BENCH):
        return None
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        f.write(np.int32(t.n).tobytes())
        f.write(bytes(p))
        f.write(t.vnodes.tobytes())
        f.write(t.conn.tobytes())
        f.write(t.status.tobytes())
        path = f.name
    env = dict(os.environ, LSAN_OPTIONS="detect_leaks=0", OMP_NUM_THREADS=str(threads))
    try:
        out = subprocess.run([REF_BENCH, path, str(threads), str(steps), repr(dt),
                              str(reconnect), mode],
                             capture_output=True, text=True, env=env, timeout=timeout)
    finally:
        os.unlink(path)
    if out.returncode != 0:
        log(f"cpu baseline ({mode}) failed: {out.stderr[-400:]}")
        return None
    return json.loads(out.stdout.strip().splitlines()[-1])


def sub_box(ov, t, side):
    """The points of t inside the corner cube [lo, lo + side)^3 of its box,
    as closed loops of the same slots' positions: each maximal run of
    consecutive points inside the cube is closed on itself (runs of fewer
    than 8 points are dropped), followed by empty slots in t's ratio of
    slots to points.  Line density and spacing are the full tangle's, so
    the candidate pairs per point of the reconnection scan are too, away
    from the cube's faces."""
    inside = np.all(t.vnodes < side, axis=1) & (t.status[:, 0] != ov.EMPTY)
    fwd, rev = t.conn[:, 0], t.conn[:, 1]
    seen = np.zeros(t.n, dtype=bool)
    runs = []
    for k in np.nonzero(inside)[0]:
        if seen[k]:
            continue
        a = k  # walk back to the run's first point (a loop wholly inside stops at k)
        while inside[rev[a]] and rev[a] != k:
            a = rev[a]
        run, b = [], a
        while inside[b] and not seen[b]:
            seen[b] = True
            run.append(b)
            b = fwd[b]
        if len(run) >= 8:
            runs.append(run)
    m = sum(len(r) for r in runs)
    # the full tangle's ratio of slots to points: the reference scans every
    # slot, empty ones included
    n_cap = max(m, int(round(m * t.n / len(t.live))))
    v = np.zeros((n_cap, 3))
    c = np.full((n_cap, 2), -1, dtype=np.int32)
    s = np.zeros((n_cap, 2), dtype=np.int32)
    s[:, 0] = ov.EMPTY
    o = 0
    for run in runs:
        L = len(run)
        v[o:o + L] = t.vnodes[run]
        j = np.arange(L)
        c[o:o + L, 0] = o + (j + 1) % L
        c[o:o + L, 1] = o + (j - 1) % L
        s[o:o + L] = (ov.FREE, -1)
        o += L
    return ov.Tangle(v, c, s)


def cpu_baseline(workload, threads, sub_points, sample_rings, sample_steps, full_step):
    """The reference on the host cores, per phase, on the benchmark tangle:
    one octree build and one update_tangle (tangents, tree, the velocity of
    every point) at full size; rk4_step is three more such evaluations, so
    it is taken as 3 x update_tangle unless full_step asks to time it too
    (it takes that long); one reconnect on a corner sub-box of sub_points
    points at the same line density, scaled by (N / N_sub)^2 (the
    reference's reconnect is a serial all-pairs scan whose candidates per
    point are fixed by the density, each costing two direct sums over every
    segment: both parts grow as N^2).  A step is reconnect + update_tangle +
    rk4_step + wrap (FFTW, the reference's smoothing library, is absent, so
    smoothing is not timed).  Beside it, the bounded sample of earlier
    versions: `sample_steps` full steps of the first `sample_rings` rings."""
    import openvort_amd as ov
    from openvort_amd import configs
    t, p, dt = configs.make(workload)
    n = len(t.live)
    ph = ref_bench(t, p, dt, threads, steps=1 if full_step else 0, mode="phases",
                   timeout=3000)
    if not ph:
        return None
    rk4_measured = ph["rk4_step_s"] >= 0
    if not rk4_measured:
        ph["rk4_step_s"] = 3 * ph["update_tangle_s"]
    D = p.box_hi[0] - p.box_lo[0]
    st = sub_box(ov, t, D * (sub_points / n) ** (1.0 / 3.0))
    rc = ref_bench(st, p, dt, threads, reconnect=1, mode="reconnect")
    if not rc:
        return None
    scale = (n / rc["points"]) ** 2
    rec_s = rc["reconnect_s"] * scale
    step_s = rec_s + ph["update_tangle_s"] + ph["rk4_step_s"] + ph["wrap_s"]
    out = {"value": n / step_s, "unit": "point-steps/s", "cores": threads, "kind": "reference",
           "sample": (f"{workload} at full size ({n} points): one octree_build and one "
                      f"update_tangle timed, rk4_step "
                      + ("timed" if rk4_measured else "taken as 3 x update_tangle (its three "
                         "evaluations)")
                      + f"; reconnect timed on a corner "
                      f"sub-box of {rc['points']} points at the same line density and scaled "
                      f"by (N/N_sub)^2 = {scale:.4g}; step = reconnect + update_tangle + "
                      f"rk4_step + wrap; smoothing not timed (FFTW absent); reference "
                      f"src/ built with its Makefile flags (-march=x86-64-v3)"),
           "step_s": step_s,
           "phases_s": {"tree_build": ph["tree_build_s"], "update_tangle": ph["update_tangle_s"],
                        "rk4_step": ph["rk4_step_s"], "rk4_step_measured": rk4_measured,
                        "wrap": ph["wrap_s"], "reconnect_extrapolated": rec_s},
           "reconnect_sub_box": {"points": rc["points"], "seconds": rc["reconnect_s"],
                                 "reconnections": rc["reconnections"]},
           "nproc": os.cpu_count()}
    if sample_rings > 0 and sample_steps > 0:
        ts, ps, dts = configs.make(workload, rings=sample_rings)
        r = ref_bench(ts, ps, dts, threads, steps=sample_steps, reconnect=1, mode="steps")
        if r:
            out["bounded_sample"] = {
                "value": r["point_steps_per_s"], "points": r["points"], "steps": r["steps"],
                "seconds": r["seconds"], "reconnect_s": r["reconnect_s"],
                "sample": (f"first {sample_rings} rings of the {workload} stream, "
                           f"{r['steps']} full steps incl. the O(N^2) reconnect")}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--leaf", type=int, default=1)
    ap.add_argument("--traverse", default="list", choices=["list", "group", "point"])
    ap.add_argument("--ppw", type=int, default=16, help="list traversal: points per wavefront")
    ap.add_argument("--group-level", type=int, default=0,
                    help="list traversal: octree level that bounds a group (0 = auto)")
    ap.add_argument("--stack", type=int, default=0,
                    help="list traversal stack depth (0 = 384; the autotuner may change it)")
    ap.add_argument("--tile", type=int, default=0,
                    help="list traversal far-field flush tile, 96 or 128 (0 = 128)")
    ap.add_argument("--blocks", type=int, default=0,
                    help="list traversal persistent grid, wavefronts per CU (0 = one per group)")
    ap.add_argument("--rec-velocity", default="tree", choices=["direct", "tree"])
    ap.add_argument("--topology", default="device", choices=["device", "host"],
                    help="where reconnection, small-loop removal and remesh run")
    ap.add_argument("--other-topo-steps", type=int, default=3,
                    help="also time this many steps with the other topology mode (0 = skip)")
    ap.add_argument("--smooth-every", type=int, default=10)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--reuse-steps", type=int, default=5,
                    help="also time this many steps with reuse_lists = 1 (0 = skip)")
    ap.add_argument("--rec-direct-steps", type=int, default=3,
                    help="also time this many steps with the reference's direct-sum "
                         "reconnection velocities (0 = skip)")
    ap.add_argument("--cpu-rings", type=int, default=62,
                    help="CPU baseline: rings of the secondary bounded sample (0 = skip)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-full-step", action="store_true",
                    help="CPU baseline: also time rk4_step at full size (minutes)")
    ap.add_argument("--cpu-sub-points", type=int, default=25000,
                    help="CPU baseline: points of the sub-box the reconnect is timed on")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0 = the host cores this process may use, "
                         "at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep --ppw/--leaf/--group-level instead of tuning them first")
    ap.add_argument("--accuracy-sample", type=int, default=2048,
                    help="points of the direct-sum error check (0 = skip)")
    ap.add_argument("--jobs", type=int, default=24,
                    help="c5 (and c4 with --serve): number of jobs; c4 forest: 64 unless set")
    ap.add_argument("--serve", action="store_true",
                    help="c4: run the ensemble through the request loop (full steps with "
                         "topology, concurrent contexts) instead of as one forest")
    ap.add_argument("--steps-scale", type=float, default=0.05,
                    help="c4/c5: fraction of each job's requested steps")
    ap.add_argument("--workers", type=int, default=8, help="c4/c5: stepping threads")
    ap.add_argument("--max-active", type=int, default=16, help="c4/c5: admitted jobs")
    ap.add_argument("--hbm-cap", type=float, default=200e9, help="c4/c5: admission cap, bytes")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (5 launches) for "
                         "calibrating the FETCH_SIZE / WRITE_SIZE counters, then exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path "
              f"(asked for --gpus {args.gpus}); the traversal would shard by target "
              f"groups with the tree replicated on every GPU.",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            torch.neg(a, out=b)
        torch.cuda.synchronize()
        log(f"copy probe: 5 x {args.copy_probe} MiB read + {args.copy_probe} MiB written")
        return

    if args.workload == "c4" and not args.serve:
        forest_bench(args)
        return
    if args.workload in ("c4", "c5"):
        serve_bench(args)
        return

    import openvort_amd as ov
    from openvort_amd import configs

    prec = ov.PREC_MIXED if args.precision == "mixed" else ov.PREC_FP64
    trav = {"list": ov.TRAVERSE_LIST, "group": ov.TRAVERSE_GROUP,
            "point": ov.TRAVERSE_POINT}[args.traverse]
    recv = ov.RECVEL_DIRECT if args.rec_velocity == "direct" else ov.RECVEL_TREE
    t, p, dt = configs.make(args.workload, precision=prec, traverse=trav, leaf_size=args.leaf)
    p.rec_velocity = recv
    p.topology = ov.TOPO_DEVICE if args.topology == "device" else ov.TOPO_HOST
    p.points_per_wave = args.ppw
    p.group_level = args.group_level
    p.list_stack, p.list_tile, p.list_blocks = args.stack, args.tile, args.blocks
    n_pts = len(t.live)
    log(f"workload {args.workload}: {n_pts} points, {t.n} slots")

    q_plain = configs.make(args.workload)[1]
    q_plain.method, q_plain.precision, q_plain.traverse, q_plain.leaf_size = (
        ov.METHOD_BH, ov.PREC_FP64, ov.TRAVERSE_POINT, 1)
    q_plain.rec_velocity = recv

    # ---- fast path ------------------------------------------------------
    e = ov.Engine(p)
    e.upload(t)
    e.synchronize()
    tune = None
    if trav == ov.TRAVERSE_LIST and not args.no_autotune:
        a = time.perf_counter()
        rows, best = e.autotune(repeats=2)
        tune = dict(rows[best], seconds=time.perf_counter() - a, candidates=len(rows),
                    shapes=[[r["list_stack"], r["list_tile"], r["list_blocks"], round(r["ms"], 3)]
                            for r in rows[-18:]])
        log(f"autotune: {tune}")
    p_run = e.get_params()  # the tuned setting that the timed steps use

    acc = None
    if args.accuracy_sample > 0:
        acc = accuracy(ov, t, p_run, q_plain, args.accuracy_sample)
        log(f"accuracy vs direct: fast {acc['fast_vs_direct']:.3e}, "
            f"plain {acc['plain_fp64_vs_direct']:.3e}; mixed vs FP64 same tree "
            f"{acc['mixed_vs_fp64_same_tree']:.3e}")
    rb = configs.remesh_bounds(args.workload)
    lat_w, tm = run_steps(e, args.warmup, 0.0, dt, 0, args.smooth_every, True, rb)
    e.stats()  # reset counters
    e.set_timing(True)
    e.synchronize()
    t0 = time.perf_counter()
    lat, tm = run_steps(e, args.steps, tm, dt, args.warmup, args.smooth_every, True, rb)
    e.synchronize()
    wall = time.perf_counter() - t0
    st = e.stats()
    L, n_after = e.line_length()
    e.close()
    value = n_pts * args.steps / wall
    log(f"fast: {wall / args.steps * 1e3:.1f} ms/step, {value:.4g} point-steps/s")

    launches = max(st.bs_launches, 1)
    bs_ms = st.ms_bs / launches
    inter_per_launch = st.interactions / launches
    achieved = inter_per_launch * FLOPS_PER_PAIR / (bs_ms * 1e-3) if bs_ms > 0 else 0.0
    peak = PEAK_FP32 if prec == ov.PREC_MIXED else PEAK_FP64
    interactions_per_s = st.interactions / wall

    # ---- the same path with the step's first lists reused ----------------
    reuse = None
    if trav == ov.TRAVERSE_LIST and args.reuse_steps > 0:
        qr = ov.Params.from_buffer_copy(p_run)
        qr.reuse_lists = 1
        e3 = ov.Engine(qr)
        e3.upload(t)
        e3.synchronize()
        _, tm3 = run_steps(e3, args.warmup, 0.0, dt, 0, args.smooth_every, True, rb)
        a = time.perf_counter()
        lat3, _ = run_steps(e3, args.reuse_steps, tm3, dt, args.warmup, args.smooth_every, True,
                            rb)
        e3.synchronize()
        w3 = time.perf_counter() - a
        e3.close()
        reuse = {"value": n_pts * args.reuse_steps / w3, "steps": args.reuse_steps,
                 "ms_per_step": w3 / args.reuse_steps * 1e3,
                 "p50_step_ms": pct(lat3, 50), "p99_step_ms": pct(lat3, 99),
                 "path": "as the headline path, with reuse_lists = 1: the interaction "
                         "lists of each step's first evaluation serve its RK4 stages "
                         "(acceptance decided at the step's first stage; not the "
                         "headline value)"}
        log(f"reuse_lists: {reuse['ms_per_step']:.1f} ms/step")

    # ---- the same path with the topology passes in the other mode -------
    other_topo = None
    if args.other_topo_steps > 0:
        qh = ov.Params.from_buffer_copy(p_run)
        dev = p_run.topology != ov.TOPO_DEVICE
        qh.topology = ov.TOPO_DEVICE if dev else ov.TOPO_HOST
        e5 = ov.Engine(qh)
        e5.upload(t)
        e5.synchronize()
        _, tm5 = run_steps(e5, args.warmup, 0.0, dt, 0, args.smooth_every, True, rb)
        e5.synchronize()
        a = time.perf_counter()
        lat5, _ = run_steps(e5, args.other_topo_steps, tm5, dt, args.warmup, args.smooth_every,
                            True, rb)
        e5.synchronize()
        w5 = time.perf_counter() - a
        e5.close()
        other_topo = {"value": n_pts * args.other_topo_steps / w5,
                      "steps": args.other_topo_steps,
                      "ms_per_step": w5 / args.other_topo_steps * 1e3,
                      "p50_step_ms": pct(lat5, 50), "p99_step_ms": pct(lat5, 99),
                      "topology": "device" if dev else "host",
                      "path": "as the headline path, with reconnection, small-loop removal and "
                              + ("remesh as device kernels that queue without waiting for the "
                                 "host (topology = device)" if dev else
                                 "remesh as sequential passes on host mirrors (topology = host)")}
        log(f"{other_topo['topology']} topology: {other_topo['ms_per_step']:.1f} ms/step")

    # ---- the same path with the reference's reconnection velocities ------
    rec_direct = None
    if args.rec_direct_steps > 0 and p_run.rec_velocity != ov.RECVEL_DIRECT:
        qd = ov.Params.from_buffer_copy(p_run)
        qd.rec_velocity = ov.RECVEL_DIRECT
        e4 = ov.Engine(qd)
        e4.upload(t)
        e4.synchronize()
        _, tm4 = run_steps(e4, 1, 0.0, dt, 0, args.smooth_every, True, rb)
        e4.synchronize()
        a = time.perf_counter()
        lat4, _ = run_steps(e4, args.rec_direct_steps, tm4, dt, 1, args.smooth_every, True, rb)
        e4.synchronize()
        w4 = time.perf_counter() - a
        # one pass alone: its endpoints and wall time
        pairs = e4.reconnect_candidates()
        ends = len(np.unique(pairs)) if len(pairs) else 0
        e4.synchronize()
        a = time.perf_counter()
        e4.reconnect(tm4)
        e4.synchronize()
        w_pass = time.perf_counter() - a
        e4.close()
        rec_direct = {"value": n_pts * args.rec_direct_steps / w4,
                      "steps": args.rec_direct_steps,
                      "ms_per_step": w4 / args.rec_direct_steps * 1e3,
                      "p50_step_ms": pct(lat4, 50), "p99_step_ms": pct(lat4, 99),
                      "reconnect_pass_ms": w_pass * 1e3, "candidate_pairs": int(len(pairs)),
                      "candidate_endpoints": int(ends),
                      "direct_pairs_per_pass": float(ends) * n_pts * (1 + p_run.n_images),
                      "path": "as the headline path, with the reference's reconnection "
                              "velocities: the FP64 direct sum at every candidate endpoint "
                              "(update_velocity with tree == NULL, "
                              "src/vortex_dynamics.c:250-251)"}
        log(f"rec_velocity direct: {rec_direct['ms_per_step']:.1f} ms/step, pass "
            f"{w_pass * 1e3:.1f} ms over {ends} endpoints")

    # ---- plain FP64 path ------------------------------------------------
    plain = None
    if args.plain_steps > 0:
        e2 = ov.Engine(q_plain)
        e2.upload(t)
        e2.synchronize()
        run_steps(e2, 1, 0.0, dt, 0, args.smooth_every, True, rb)
        a = time.perf_counter()
        lat2, _ = run_steps(e2, args.plain_steps, dt, dt, 1, args.smooth_every, True, rb)
        w2 = time.perf_counter() - a
        e2.close()
        plain = {"value": n_pts * args.plain_steps / w2, "steps": args.plain_steps,
                 "ms_per_step": w2 / args.plain_steps * 1e3,
                 "p50_step_ms": pct(lat2, 50), "p99_step_ms": pct(lat2, 99),
                 "path": "FP64, reference split-until-one tree, per-point traversal"}
        log(f"plain FP64: {plain['ms_per_step']:.1f} ms/step")

    # ---- CPU baseline ---------------------------------------------------
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        cpu = cpu_baseline(args.workload, threads, args.cpu_sub_points, args.cpu_rings,
                           args.cpu_steps, args.cpu_full_step)
        if cpu:
            log(f"cpu: {cpu['value']:.4g} point-steps/s ({cpu['step_s']:.0f} s per step, "
                f"phases {cpu['phases_s']})")

    # HBM bytes per launch of the traversal: measured with rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes (copy-kernel calibrated) for this
    # workload and setting, committed beside the kernel-trace summary
    traffic, traffic_src = None, None
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         "k_bh_list_mixed_c3_traffic.json")
    cfg_now = {"workload": args.workload, "points_per_wave": p_run.points_per_wave,
               "leaf_size": p_run.leaf_size, "group_level": p_run.group_level,
               "list_stack": p_run.list_stack or 384, "list_tile": p_run.list_tile or 128,
               "list_blocks": p_run.list_blocks}
    if (os.path.exists(tfile) and trav == ov.TRAVERSE_LIST and prec == ov.PREC_MIXED):
        with open(tfile) as f:
            tj = json.load(f)
        if all(tj.get(k) == v for k, v in cfg_now.items()):
            traffic, traffic_src = tj["traffic_bytes_per_launch"], tj["source"]

    out = {
        "metric": METRIC, "value": value, "unit": "point-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if prec == ov.PREC_MIXED else "f64", "data": "synthetic",
        "config": {"workload": (f"{args.workload}: seeded ring tangle, {n_pts} points, "
                                f"BH theta {p.theta}, {p.n_images} periodic images, "
                                f"dt {dt}, reconnect every step, smooth every "
                                f"{args.smooth_every} steps"),
                   "points": n_pts, "precision": args.precision, "traverse": args.traverse,
                   "leaf_size": p_run.leaf_size, "points_per_wave": p_run.points_per_wave,
                   "group_level": p_run.group_level, "list_stack": p_run.list_stack or 384,
                   "list_tile": p_run.list_tile or 128, "list_blocks": p_run.list_blocks,
                   "rec_velocity": args.rec_velocity, "topology": args.topology,
                   "autotune": tune},
        "p50_step_ms": pct(lat, 50), "p99_step_ms": pct(lat, 99),
        "interactions_per_s": interactions_per_s,
        "direct_equivalent_interactions_per_s": 27.0 * n_pts * n_pts * 4 * args.steps / wall,
        "line_length_cm": L,
        "plain_fp64": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "reuse_lists": reuse,
        "other_topology": other_topo,
        "rec_velocity_direct": rec_direct,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                     "traffic_unit": "bytes per launch", "traffic_source": traffic_src,
                     "kernel": {ov.TRAVERSE_LIST: "k_bh_list", ov.TRAVERSE_GROUP: "k_bh_group",
                                ov.TRAVERSE_POINT: "k_bh_point"}[trav]
                     + ("<mixed>" if prec == ov.PREC_MIXED else "<fp64>"),
                     "avg_launch_ms": bs_ms,
                     "pairs_per_launch": inter_per_launch,
                     "fp32_pair_fraction": st.interactions_fp32 / max(st.interactions, 1)},
        "cpu_baseline": cpu,
        "accuracy": acc,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
