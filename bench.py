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
        if reconnect:
            e.reconnect(tm)
            e.eliminate_small_loops(5)
        if smooth_every and k % smooth_every == 0:
            e.smooth()
        e.enforce_periodic()
        if remesh:
            e.remesh(*remesh)
            e.eliminate_small_loops(5)
        e.update_tangle(tm)
        e.rk4_step(tm, dt)
        e.enforce_periodic()
        e.marker_record(s + 1)
        tm += dt
    lat = [e.marker_elapsed(s, s + 1) * 1e-3 for s in range(n_steps)]
    return lat, tm


def accuracy(ov, t, p_fast, p_plain, n_sample, seed=7):
    """Velocity error of both paths on the initial tangle against the FP64
    direct sum (ovg_velocity_at, the reference's update_velocity with
    tree == NULL) at a seeded sample of points: relative L2 over the sample."""
    live = t.live
    rng = np.random.default_rng(seed)
    sub = np.sort(rng.choice(live, size=min(n_sample, len(live)), replace=False))
    out = {}
    direct = None
    for name, q in (("fast", p_fast), ("plain_fp64", p_plain)):
        e = ov.Engine(q)
        e.upload(t)
        e.update_tangle(0.0)
        v = e.download()["vels"][sub]
        if direct is None:
            direct = e.velocity_at(sub, 0.0)[1]
        e.close()
        out[name] = v
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return {"sample_points": int(len(sub)), "field": "vels",
            "fast_vs_direct": rel(out["fast"], direct),
            "plain_fp64_vs_direct": rel(out["plain_fp64"], direct),
            "fast_vs_plain_fp64": rel(out["fast"], out["plain_fp64"])}


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


def ov_points(job):
    import openvort_amd as ov
    return ov.lib().ovg_generate_rings(job.seed, job.n_rings, job.D, job.r_min, job.r_max,
                                       job.dxi, 0, None, None, None)


def pct(v, q):
    return float(np.percentile(np.asarray(v) * 1e3, q))


def cpu_baseline(workload, rings, threads, steps):
    """The reference (oracle/_ref/ovref_bench, built with the reference
    Makefile's flags) on a bounded sample of the same workload."""
    from oracle.ovoracle import REF_BENCH
    from openvort_amd import configs
    if not os.path.exists(REF_BENCH):
        return None
    t, p, dt = configs.make(workload, rings=rings)
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        f.write(np.int32(t.n).tobytes())
        f.write(bytes(p))
        f.write(t.vnodes.tobytes())
        f.write(t.conn.tobytes())
        f.write(t.status.tobytes())
        path = f.name
    env = dict(os.environ, LSAN_OPTIONS="detect_leaks=0", OMP_NUM_THREADS=str(threads))
    try:
        out = subprocess.run([REF_BENCH, path, str(threads), str(steps), repr(dt), "1"],
                             capture_output=True, text=True, env=env, timeout=600)
    finally:
        os.unlink(path)
    if out.returncode != 0:
        log(f"cpu baseline failed: {out.stderr[-400:]}")
        return None
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return r


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
    ap.add_argument("--rec-velocity", default="tree", choices=["direct", "tree"])
    ap.add_argument("--smooth-every", type=int, default=10)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--reuse-steps", type=int, default=5,
                    help="also time this many steps with reuse_lists = 1 (0 = skip)")
    ap.add_argument("--cpu-rings", type=int, default=62)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep --ppw/--leaf/--group-level instead of tuning them first")
    ap.add_argument("--accuracy-sample", type=int, default=2048,
                    help="points of the direct-sum error check (0 = skip)")
    ap.add_argument("--jobs", type=int, default=24, help="c4/c5: number of jobs")
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
    p.points_per_wave = args.ppw
    p.group_level = args.group_level
    n_pts = len(t.live)
    log(f"workload {args.workload}: {n_pts} points, {t.n} slots")

    q_plain = configs.make(args.workload)[1]
    q_plain.method, q_plain.precision, q_plain.traverse, q_plain.leaf_size = (
        ov.METHOD_BH, ov.PREC_FP64, ov.TRAVERSE_POINT, 1)
    q_plain.rec_velocity = recv

    acc = None
    if args.accuracy_sample > 0:
        acc = accuracy(ov, t, p, q_plain, args.accuracy_sample)
        log(f"accuracy vs direct: fast {acc['fast_vs_direct']:.3e}, "
            f"plain {acc['plain_fp64_vs_direct']:.3e}")

    # ---- fast path ------------------------------------------------------
    e = ov.Engine(p)
    e.upload(t)
    e.synchronize()
    tune = None
    if trav == ov.TRAVERSE_LIST and not args.no_autotune:
        a = time.perf_counter()
        rows, best = e.autotune(repeats=2)
        tune = dict(rows[best], seconds=time.perf_counter() - a, candidates=len(rows))
        log(f"autotune: {tune}")
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
        qr = ov.Params.from_buffer_copy(p)
        if tune:
            qr.points_per_wave, qr.leaf_size, qr.group_level = (
                tune["points_per_wave"], tune["leaf_size"], tune["group_level"])
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
        r = cpu_baseline(args.workload, args.cpu_rings, args.cpu_threads, args.cpu_steps)
        if r:
            cpu = {"value": r["point_steps_per_s"], "unit": "point-steps/s",
                   "cores": r["threads"], "kind": "reference",
                   "sample": (f"first {args.cpu_rings} rings of the {args.workload} stream "
                              f"({r['points']} points, same box and parameters), "
                              f"{r['steps']} full steps incl. the O(N^2) reconnect "
                              f"({r['reconnect_s']:.1f} s of {r['seconds']:.1f} s); "
                              f"reference src/ built with its Makefile flags "
                              f"(-march=x86-64-v3)"),
                   "seconds": r["seconds"]}
            log(f"cpu: {cpu['value']:.4g} point-steps/s on {r['points']} points")

    # HBM bytes per launch of the traversal: measured with rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes (copy-kernel calibrated) for this
    # workload and setting, committed beside the kernel-trace summary
    traffic, traffic_src = None, None
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         "k_bh_list_mixed_c3_traffic.json")
    cfg_now = {"workload": args.workload,
               "points_per_wave": tune["points_per_wave"] if tune else args.ppw,
               "leaf_size": tune["leaf_size"] if tune else args.leaf,
               "group_level": tune["group_level"] if tune else args.group_level}
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
                   "leaf_size": tune["leaf_size"] if tune else args.leaf,
                   "points_per_wave": tune["points_per_wave"] if tune else args.ppw,
                   "group_level": tune["group_level"] if tune else args.group_level,
                   "rec_velocity": args.rec_velocity, "autotune": tune},
        "p50_step_ms": pct(lat, 50), "p99_step_ms": pct(lat, 99),
        "interactions_per_s": interactions_per_s,
        "direct_equivalent_interactions_per_s": 27.0 * n_pts * n_pts * 4 * args.steps / wall,
        "line_length_cm": L,
        "plain_fp64": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "reuse_lists": reuse,
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
