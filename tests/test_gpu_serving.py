"""Serving through one resident launch (dms_server): requests join the running
trajectory launch as slots.

Bar: every request's results equal the oracle's Generate_Data run of its own
descriptor (FP64) and the results of dms_job_submit on the same descriptor
(both precisions): counters, visit counts and evaporation events identical,
histogram sums within 1e-12 relative (the device sums each sync round's rows
with atomics, in a varying order).  The server's counters show that requests
submitted while the launch runs joined it instead of starting launches of their
own, and that the launch ends when idle and starts again for the next request.
"""
import time

import numpy as np
import pytest

import dmsun
import oracle_lib
from test_gpu_parity import RK_BUDGET, assert_job_equal

pytestmark = pytest.mark.gpu

# the serving workload's request ranges: SI points with m ~ U_log(0.1, 5) GeV
# and sigma_p ~ U_log(1e-38, 1e-35) cm^2, targets of captured particles
BINS = 200
MAX_SCAT = 500
RK = 200000


def launch_desc(mode=dmsun.MODE_PARAMETER_POINT, precision=dmsun.FP64, bins=BINS):
    return dmsun.make_job(mode=mode, max_trajectories=0, capture_target=0, max_scatterings=MAX_SCAT,
                          rk_attempt_budget=RK, n_bins=bins, precision=precision, record=0)


def draw_requests(n, seed, target_lo, target_hi, max_traj=40000, mode=dmsun.MODE_PARAMETER_POINT,
                  precision=dmsun.FP64):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        m = float(np.exp(rng.uniform(np.log(0.1), np.log(5.0))))
        s = float(np.exp(rng.uniform(np.log(1e-38), np.log(1e-35))))
        target = int(np.exp(rng.uniform(np.log(target_lo), np.log(target_hi))))
        pdesc = dmsun.make_point(m, s, interpolation_points=300)
        job = dmsun.make_job(mode=mode, max_trajectories=max_traj, capture_target=target, seed=seed + k, point_id=k,
                             max_scatterings=MAX_SCAT, rk_attempt_budget=RK, n_bins=BINS, precision=precision,
                             record=0)
        out.append((pdesc, job))
    return out


def serve_all(server, points, jobs):
    handles = [server.submit(pt, j) for pt, j in zip(points, jobs)]
    return [server.collect(h) for h in handles]


def test_served_requests_match_the_oracle(gpu_ctx, oracle_solar):
    reqs = draw_requests(4, 101, 300, 800)
    points = [gpu_ctx.point(p) for p, _ in reqs]
    jobs = [j for _, j in reqs]
    srv = dmsun.Server(gpu_ctx, launch_desc())
    try:
        print("serving 4 requests through one resident launch", flush=True)
        res = serve_all(srv, points, jobs)
        st = srv.stats()
    finally:
        srv.close()
    print(f"server stats {st}", flush=True)
    assert st["launches"] == 1 and st["requests_done"] == 4
    for (pdesc, job), r in zip(reqs, res):
        print(f"request point_id {job.point_id}: m={pdesc.mass_GeV:.4g} sigma_p={pdesc.sigma_p_cm2:.4g} "
              f"target {job.capture_target}: {r['counters'].trajectories} trajectories", flush=True)
        o = oracle_lib.Point(oracle_solar, pdesc, threads=16).run_job(job, threads=16, want_records=False)
        assert_job_equal(r, o)
        assert r["counters"].captured >= job.capture_target or r["counters"].early_stop_reason != 0


def test_capture_mode_requests_match_the_oracle(gpu_ctx, oracle_solar):
    reqs = draw_requests(3, 202, 200, 500, mode=dmsun.MODE_CAPTURE)
    points = [gpu_ctx.point(p) for p, _ in reqs]
    jobs = [j for _, j in reqs]
    srv = dmsun.Server(gpu_ctx, launch_desc(mode=dmsun.MODE_CAPTURE))
    try:
        print("serving 3 capture-mode requests", flush=True)
        res = serve_all(srv, points, jobs)
    finally:
        srv.close()
    for (pdesc, job), r in zip(reqs, res):
        o = oracle_lib.Point(oracle_solar, pdesc, threads=16).run_job(job, threads=16, want_records=False)
        assert_job_equal(r, o)


def test_fixed_count_requests_small_ring_match_the_oracle(gpu_ctx, oracle_solar):
    """Fixed trajectory counts (every sync round counts) through a ring of 1024
    slots and 64 round buffers, so that ring positions and round buffers are
    reused many times within each request."""
    pdescs = [dmsun.make_point(0.3, 1e-36, interpolation_points=300), dmsun.make_point(1.5, 4e-37, interpolation_points=300),
              dmsun.make_point(0.15, 1e-35, interpolation_points=300)]
    jobs = [dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=n, capture_target=0, seed=7, point_id=k,
                           max_scatterings=MAX_SCAT, rk_attempt_budget=RK, n_bins=BINS, record=0)
            for k, n in enumerate([3000, 2500, 1700])]
    points = [gpu_ctx.point(p) for p in pdescs]
    srv = dmsun.Server(gpu_ctx, launch_desc(), slot_ring=1024, max_rounds=64)
    try:
        print("fixed-count requests, ring 1024, 64 round buffers", flush=True)
        res = serve_all(srv, points, jobs)
    finally:
        srv.close()
    for pdesc, job, r in zip(pdescs, jobs, res):
        o = oracle_lib.Point(oracle_solar, pdesc, threads=16).run_job(job, threads=16, want_records=False)
        assert_job_equal(r, o)
        assert r["counters"].trajectories == job.max_trajectories


@pytest.mark.parametrize("precision", [dmsun.FP64, dmsun.MIXED])
def test_requests_join_the_running_launch(gpu_ctx, precision):
    """A request submitted while another one runs joins the same launch (one
    launch in the server's counters) and gets the results it gets alone."""
    reqs = draw_requests(3, 303, 400, 900, precision=precision)
    points = [gpu_ctx.point(p) for p, _ in reqs]
    jobs = [j for _, j in reqs]
    srv = dmsun.Server(gpu_ctx, launch_desc(precision=precision))
    try:
        first = srv.submit(points[0], jobs[0])
        t0 = time.time()
        while srv.poll(first).trajectories_done == 0 and time.time() - t0 < 30:
            time.sleep(0.002)
        running = srv.stats()
        print(f"first request running: {running}", flush=True)
        later = [srv.submit(pt, j) for pt, j in zip(points[1:], jobs[1:])]
        res = [srv.collect(first)] + [srv.collect(h) for h in later]
        st = srv.stats()
    finally:
        srv.close()
    print(f"server stats {st}", flush=True)
    assert running["running"] == 1
    assert st["launches"] == 1 and st["requests_done"] == 3
    for pt, job, r in zip(points, jobs, res):
        assert_job_equal(r, pt.run_job(job, want_records=False))


def test_point_created_beside_the_running_launch(gpu_ctx):
    """A point created while the serving launch runs has its rate table built
    beside the launch (the table kernel fits in the registers the resident
    launch leaves free on each SIMD), so it is ready before the running request
    ends and a request for it joins the same launch."""
    pt = gpu_ctx.point(dmsun.make_point(0.3, 1e-36, interpolation_points=300))
    long_job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=300000, capture_target=0,
                              max_scatterings=MAX_SCAT, rk_attempt_budget=RK, n_bins=BINS, record=0)
    srv = dmsun.Server(gpu_ctx, launch_desc())
    try:
        first = srv.submit(pt, long_job)
        t0 = time.time()
        while srv.poll(first).trajectories_done == 0 and time.time() - t0 < 30:
            time.sleep(0.002)
        t1 = time.time()
        pdesc2, job2 = draw_requests(1, 606, 100, 101)[0]
        pt2 = gpu_ctx.point(pdesc2)
        built = time.time() - t1
        during = srv.stats()
        second = srv.submit(pt2, job2)
        res = [srv.collect(first), srv.collect(second)]
        st = srv.stats()
    finally:
        srv.close()
    print(f"point built in {built * 1e3:.1f} ms while the launch ran: {during}; after {st}", flush=True)
    assert during["running"] == 1 and during["requests_done"] == 0
    assert st["launches"] == 1 and st["requests_done"] == 2
    assert_job_equal(res[1], pt2.run_job(job2, want_records=False))


def test_idle_launch_ends_and_restarts(gpu_ctx):
    pdesc, job = draw_requests(1, 404, 100, 101)[0]
    pt = gpu_ctx.point(pdesc)
    srv = dmsun.Server(gpu_ctx, launch_desc())
    try:
        a = srv.collect(srv.submit(pt, job))
        t0 = time.time()
        while srv.stats()["running"] and time.time() - t0 < 5:
            time.sleep(0.005)
        idle = srv.stats()
        b = srv.collect(srv.submit(pt, job))
        st = srv.stats()
    finally:
        srv.close()
    print(f"idle {idle} after {st}", flush=True)
    assert idle["running"] == 0 and idle["launches"] == 1
    assert st["launches"] == 2
    assert_job_equal(a, b)
    assert_job_equal(a, pt.run_job(job, want_records=False))


def test_server_rejects_requests_outside_its_launch(gpu_ctx):
    pdesc, job = draw_requests(1, 505, 10, 11)[0]
    pt = gpu_ctx.point(pdesc)
    srv = dmsun.Server(gpu_ctx, launch_desc())
    try:
        other = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=100, capture_target=0,
                               max_scatterings=MAX_SCAT, rk_attempt_budget=RK, n_bins=BINS + 1, record=0)
        with pytest.raises(dmsun.DmsError):
            srv.submit(pt, other)
        rec = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=100, capture_target=0,
                             max_scatterings=MAX_SCAT, rk_attempt_budget=RK, n_bins=BINS, record=1)
        with pytest.raises(dmsun.DmsError):
            srv.submit(pt, rec)
        assert srv.stats()["launches"] == 0
    finally:
        srv.close()
    with pytest.raises(dmsun.DmsError):
        dmsun.Server(gpu_ctx, launch_desc(), slot_ring=3000)
