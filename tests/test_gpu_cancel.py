"""dms_job_cancel: a running launch ends within milliseconds on the host's
request (every lane reads the launch's cancel word), the job reports
DMS_ERR_CANCELLED, and the device and the context stay usable.  Every job here
is bounded (it would end by itself within tens of seconds), so a cancel that
did not work fails the elapsed-time checks instead of running on."""
import time

import pytest

import dmsun
from test_gpu_parity import assert_job_equal

pytestmark = pytest.mark.gpu


def bench_job(n, **kw):
    return dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=n, capture_target=0, seed=33,
                          max_scatterings=10**14, n_bins=1000, record=0, halo=dmsun.default_halo((0.0, 248.0, 0.0)), **kw)


@pytest.fixture(scope="module")
def bench_point(gpu_ctx):
    return gpu_ctx.point(dmsun.make_point(0.3, 1e-36, interpolation_points=1000))


def wait_running(pt, h, limit=20.0):
    t0 = time.time()
    st = dmsun.JobStatus()
    while time.time() - t0 < limit:
        dmsun.lib().dms_job_poll(h[0], st)
        if st.trajectories_done > 0:
            return
        time.sleep(0.005)
    raise AssertionError("job did not start")


def test_cancel_ends_a_long_launch(bench_point):
    """3 x 10^6 trajectories (about 6-10 s, its longest trajectory alone takes
    seconds) cancelled once running: the wait returns DMS_ERR_CANCELLED within
    a second, and a job run afterwards on the same context is correct."""
    pt = bench_point
    h = pt.submit(bench_job(3_000_000))
    wait_running(pt, h)
    t0 = time.time()
    pt.cancel(h)
    rc = pt.wait_status(h)
    dt = time.time() - t0
    dmsun.lib().dms_job_destroy(h[0])
    print(f"cancelled after start: rc {rc}, {dt * 1e3:.1f} ms", flush=True)
    assert rc == dmsun.DMS_ERR_CANCELLED
    assert dt < 1.0
    small = bench_job(2000, first_trajectory_id=5)
    assert_job_equal(pt.run_job(small, want_records=False), pt.run_job(small, want_records=False))


def test_cancel_a_capture_target_job_between_chunks(bench_point):
    """A capture-target job (chunked launches) stops, whether the cancel lands
    in a launch or between launches."""
    pt = bench_point
    job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=2_000_000, capture_target=10**9, seed=34,
                         max_scatterings=10**14, n_bins=1000, record=0, rk_attempt_budget=200000)
    h = pt.submit(job)
    wait_running(pt, h)
    time.sleep(0.3)
    t0 = time.time()
    pt.cancel(h)
    rc = pt.wait_status(h)
    dt = time.time() - t0
    dmsun.lib().dms_job_destroy(h[0])
    print(f"capture-target job cancelled: rc {rc}, {dt * 1e3:.1f} ms", flush=True)
    assert rc == dmsun.DMS_ERR_CANCELLED and dt < 1.0


def test_cancel_a_batch_ends_every_job_of_the_launch(bench_point):
    pt = bench_point
    jobs = [bench_job(1_000_000, first_trajectory_id=k * 1_000_000) for k in range(3)]
    descs = (dmsun.JobDesc * 3)(*jobs)
    import ctypes as C
    hs = (C.c_void_p * 3)()
    pt.ctx.check(dmsun.lib().dms_job_submit_batch(pt.h, descs, 3, hs), "dms_job_submit_batch")
    handles = [(C.c_void_p(hs[k]), jobs[k]) for k in range(3)]
    wait_running(pt, handles[0])
    t0 = time.time()
    pt.cancel(handles[1])
    rcs = [pt.wait_status(h) for h in handles]
    dt = time.time() - t0
    for h in handles:
        dmsun.lib().dms_job_destroy(h[0])
    print(f"batch cancelled: {rcs}, {dt * 1e3:.1f} ms", flush=True)
    assert rcs == [dmsun.DMS_ERR_CANCELLED] * 3 and dt < 1.0


def test_cancel_a_finished_job_changes_nothing(bench_point):
    pt = bench_point
    job = bench_job(1000, first_trajectory_id=7)
    h = pt.submit(job)
    assert pt.wait_status(h) == dmsun.DMS_OK
    pt.cancel(h)
    res = pt.collect(h)
    assert res["counters"].trajectories == 1000


def test_cancel_a_served_request(gpu_ctx):
    """A request of the resident launch stops where it is (its unstarted
    slots are handed back, its running trajectories finish); a request beside
    it completes with the results it gets alone."""
    launch = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=0, capture_target=0, max_scatterings=500,
                            rk_attempt_budget=200000, n_bins=200, record=0)
    big_pt = gpu_ctx.point(dmsun.make_point(0.5, 3e-38, interpolation_points=300))
    small_pt = gpu_ctx.point(dmsun.make_point(1.2, 2e-36, interpolation_points=300))
    big = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=2_000_000, capture_target=100_000, seed=41,
                         max_scatterings=500, rk_attempt_budget=200000, n_bins=200, record=0)
    small = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=20_000, capture_target=300, seed=42,
                           point_id=1, max_scatterings=500, rk_attempt_budget=200000, n_bins=200, record=0)
    srv = dmsun.Server(gpu_ctx, launch)
    try:
        hb = srv.submit(big_pt, big)
        hs = srv.submit(small_pt, small)
        t0 = time.time()
        while srv.poll(hb).trajectories_done == 0 and time.time() - t0 < 20:
            time.sleep(0.005)
        t0 = time.time()
        srv.cancel(hb)
        rc = srv.wait_status(hb)
        dt = time.time() - t0
        dmsun.lib().dms_job_destroy(hb[0])
        res = srv.collect(hs)
        st = srv.stats()
    finally:
        srv.close()
    print(f"served request cancelled: rc {rc}, {dt * 1e3:.1f} ms, stats {st}", flush=True)
    assert rc == dmsun.DMS_ERR_CANCELLED and dt < 5.0
    assert st["slots_cancelled"] > 0
    assert_job_equal(res, small_pt.run_job(small, want_records=False))
