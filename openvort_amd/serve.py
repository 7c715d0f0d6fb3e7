"""ctypes binding of the request loop (include/openvort_serve.h) and the
seeded job streams of the batched-ensemble and serving workloads.

Batched ensemble: 64 tangles of log-uniform 2k-100k points (ring count set
by the point target), radii U[0.005, 0.03] cm, point spacing 8e-4 cm,
D = 0.1 cm, v_ns U[0.5, 2] cm/s, T = 1.6 K or 1.9 K, all present at t = 0.
Serving stream: Poisson arrivals at 20 jobs/s, sizes log-uniform 5k-500k
points, D U[0.1, 0.2] cm, point spacing U[5e-4, 1e-3] cm, the same flows
and temperatures, U{50..500} steps per job.
"""
import ctypes as C

import numpy as np

import openvort_amd as ov

T16 = (0.098, 0.016)  # 1.6 K: alpha, alpha'
T19 = (0.206, 0.009)  # 1.9 K


class Job(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_rings", C.c_int32), ("steps", C.c_int32),
                ("D", C.c_double), ("r_min", C.c_double), ("r_max", C.c_double),
                ("dxi", C.c_double), ("alpha", C.c_double), ("alpha_p", C.c_double),
                ("v_ns", C.c_double), ("dt", C.c_double), ("arrival_s", C.c_double),
                ("smooth_every", C.c_int32), ("precision", C.c_int32),
                ("traverse", C.c_int32), ("leaf_size", C.c_int32),
                ("points_per_wave", C.c_int32), ("rec_velocity", C.c_int32),
                ("theta", C.c_double), ("topology", C.c_int32), ("pad_", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("points_initial", C.c_int32), ("points_final", C.c_int32),
                ("steps_done", C.c_int32), ("failed", C.c_int32), ("admit_s", C.c_double),
                ("done_s", C.c_double), ("latency_s", C.c_double),
                ("line_length", C.c_double), ("checksum", C.c_uint64),
                ("point_steps", C.c_double)]


class Report(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("point_steps", C.c_double),
                ("point_steps_per_s", C.c_double), ("step_p50_ms", C.c_double),
                ("step_p99_ms", C.c_double), ("job_p50_s", C.c_double),
                ("job_p99_s", C.c_double), ("peak_bytes", C.c_double),
                ("jobs_done", C.c_int32), ("jobs_failed", C.c_int32),
                ("peak_active", C.c_int32), ("steps_total", C.c_int32),
                ("pad_", C.c_int32)]


_lib = ov.lib()
_lib.ovs_job_bytes.restype = C.c_double
_lib.ovs_job_bytes.argtypes = [C.c_int]
_lib.ovs_run.restype = C.c_int
_lib.ovs_run.argtypes = [C.POINTER(Job), C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                         C.POINTER(Report), C.POINTER(Result)]


def set_path(job, fast):
    """fast: mixed precision, list traversal (16 points per wave, no
    buckets), tree velocities for reconnection; plain: the reference's FP64
    split-until-one tree with per-point traversal."""
    if fast:
        job.precision, job.traverse, job.leaf_size = ov.PREC_MIXED, ov.TRAVERSE_LIST, 1
        job.points_per_wave = 16
    else:
        job.precision, job.traverse, job.leaf_size = ov.PREC_FP64, ov.TRAVERSE_POINT, 1
        job.points_per_wave = 0
    job.rec_velocity = ov.RECVEL_TREE
    job.theta = 0.4
    return job


def _job(seed, target, D, rmin, rmax, dxi, temp, v_ns, steps, arrival, fast):
    j = Job()
    j.seed = seed
    j.n_rings = ov.rings_for_points(seed, int(target), D, rmin, rmax, dxi)
    j.steps = int(steps)
    j.D, j.r_min, j.r_max, j.dxi = D, rmin, rmax, dxi
    j.alpha, j.alpha_p = temp
    j.v_ns = v_ns
    j.dt = 5e-5
    j.arrival_s = arrival
    j.smooth_every = 10
    return set_path(j, fast)


def c4_jobs(seed=0xC0FFEE04, n=64, steps=10, fast=True):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        target = np.exp(rng.uniform(np.log(2e3), np.log(1e5)))
        temp = T16 if rng.random() < 0.5 else T19
        out.append(_job(seed * 1000 + i, target, 0.1, 0.005, 0.03, 8e-4, temp,
                        rng.uniform(0.5, 2.0), steps, 0.0, fast))
    return out


def c5_jobs(seed=0xC0FFEE05, n=1000, rate=20.0, steps_scale=1.0, fast=True):
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        target = np.exp(rng.uniform(np.log(5e3), np.log(5e5)))
        D = rng.uniform(0.1, 0.2)
        dxi = rng.uniform(5e-4, 1e-3)
        temp = T16 if rng.random() < 0.5 else T19
        v_ns = rng.uniform(0.5, 2.0)
        steps = max(1, int(round(rng.integers(50, 501) * steps_scale)))
        out.append(_job(seed * 1000 + i, target, D, 0.005, 0.03, dxi, temp, v_ns, steps, t,
                        fast))
    return out


def run(jobs, workers=4, max_active=16, cap_bytes=200e9, realtime=True):
    """Run the jobs through the request loop; returns (report dict, results)."""
    n = len(jobs)
    arr = (Job * n)(*jobs)
    res = (Result * n)()
    rep = Report()
    rc = _lib.ovs_run(arr, n, workers, max_active, cap_bytes, 1 if realtime else 0,
                      C.byref(rep), res)
    if rc != 0:
        raise ov.OvgError(f"ovs_run failed with code {rc}")
    report = {f: getattr(rep, f) for f, _ in Report._fields_ if f != "pad_"}
    return report, list(res)


def job_bytes(n_slots):
    return _lib.ovs_job_bytes(n_slots)
