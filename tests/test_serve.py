"""Request loop (include/openvort_serve.h): jobs run concurrently on their
own contexts give the same result as alone, admission respects the HBM cap,
arrivals are honoured.  Results are compared bit for bit (checksum of final
positions, links and status of the live slots)."""
import sys

import pytest

import openvort_amd as ov

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def small_jobs(n=5, steps=6):
    from openvort_amd import serve
    jobs = serve.c4_jobs(seed=11, n=n, steps=steps)
    for j, pts in zip(jobs, (2500, 4000, 6000, 3000, 5000, 3500, 4500, 2000)):
        j.n_rings = ov.rings_for_points(j.seed, pts, j.D, j.r_min, j.r_max, j.dxi)
        j.smooth_every = 3
    return jobs


def test_mix_equals_alone():
    from openvort_amd import serve
    jobs = small_jobs()
    say(f"{len(jobs)} jobs alone, then together on 3 workers")
    alone = []
    for j in jobs:
        rep, res = serve.run([j], workers=1, max_active=1, realtime=False)
        assert rep["jobs_done"] == 1 and res[0].failed == 0
        alone.append(res[0])
    rep, res = serve.run(jobs, workers=3, max_active=5, realtime=False)
    assert rep["jobs_done"] == len(jobs) and rep["jobs_failed"] == 0
    assert rep["peak_active"] > 1
    for a, b in zip(alone, res):
        assert b.steps_done == a.steps_done == jobs[0].steps
        assert b.checksum == a.checksum and b.points_final == a.points_final
        assert b.line_length == a.line_length
    assert rep["point_steps"] == sum(r.point_steps for r in res)
    assert rep["step_p99_ms"] >= rep["step_p50_ms"] > 0


def test_device_topology_jobs_equal_host_topology_jobs():
    """The same jobs with the topology passes on the device (queued, no
    host wait inside a step) and on the host, concurrently: every job's
    final positions, links and status are the same bit for bit."""
    from openvort_amd import serve
    jobs = small_jobs(n=6, steps=8)
    res = {}
    for mode in (ov.TOPO_HOST, ov.TOPO_DEVICE):
        for j in jobs:
            j.topology = mode
        say(f"{len(jobs)} jobs on 3 workers, topology {mode}")
        rep, res[mode] = serve.run(jobs, workers=3, max_active=6, realtime=False)
        assert rep["jobs_done"] == len(jobs) and rep["jobs_failed"] == 0
    for a, b in zip(res[ov.TOPO_HOST], res[ov.TOPO_DEVICE]):
        assert a.steps_done == b.steps_done == jobs[0].steps
        assert a.checksum == b.checksum and a.points_final == b.points_final
        assert a.line_length == b.line_length


def test_admission_respects_cap():
    from openvort_amd import serve
    jobs = small_jobs()
    # every job here has 8192 slots: room for two at a time
    cap = 2.5 * serve.job_bytes(8192)
    say("admission under a cap of two jobs")
    rep, res = serve.run(jobs, workers=4, max_active=8, cap_bytes=cap, realtime=False)
    assert rep["jobs_done"] == len(jobs)
    assert rep["peak_active"] <= 2 and rep["peak_bytes"] <= cap


def test_arrivals_honoured():
    from openvort_amd import serve
    jobs = small_jobs(n=3, steps=3)
    for i, j in enumerate(jobs):
        j.arrival_s = 0.3 * i
    say("three jobs arriving 0.3 s apart")
    rep, res = serve.run(jobs, workers=2, max_active=4, realtime=True)
    assert rep["jobs_done"] == 3
    for j, r in zip(jobs, res):
        assert r.admit_s >= j.arrival_s - 1e-3
        assert r.latency_s > 0 and r.done_s >= r.admit_s
