"""Many parameter points on one GPU: multi-point launches (dms_job_submit_multi),
concurrent jobs on their own streams, and the serving loop (dmsun_serve).

Bar: every job's results equal the oracle's Generate_Data run on the same
descriptor (FP64) and the job's results when it runs alone (both precisions):
counters, visit counts and evaporation events identical, histogram sums within
1e-12 relative (device atomics sum in a varying order), per-trajectory records
bit-identical.
"""
import json
import pathlib
import subprocess

import numpy as np
import pytest

import dmsun
import oracle_lib
from test_gpu_parity import RK_BUDGET, assert_job_equal, rec_bytes

pytestmark = pytest.mark.gpu

REPO = pathlib.Path(__file__).resolve().parents[1]
SERVE = REPO / "damascus-earth-evap_amd" / "dmsun_serve"

# (point, seed, point_id, trajectories): SI points whose trajectories differ
# in length (from a few to thousands of scatterings), a dark-photon point and
# one without a rate table; sized so the oracle finishes each in about a second
MULTI = [
    (dict(mass_GeV=0.3, sigma_p_cm2=1e-36, interpolation_points=1000), 11, 0, 900),
    (dict(mass_GeV=0.2, sigma_p_cm2=1e-35, interpolation_points=1000), 11, 1, 300),
    (dict(mass_GeV=0.8, sigma_p_cm2=3e-37, interpolation_points=600), 12, 2, 200),
    (dict(mass_GeV=0.1, sigma_p_cm2=0.0, sigma_e_cm2=1e-37, interaction=dmsun.DARK_PHOTON, form_factor=dmsun.FF_GENERAL,
          mediator_mass_MeV=0.01, interpolation_points=300), 13, 3, 20),
    (dict(mass_GeV=0.5, sigma_p_cm2=1e-37, interpolation_points=0), 14, 4, 60),
]


def multi_jobs(mode, precision=dmsun.FP64, record=1):
    return [dmsun.make_job(mode=mode, max_trajectories=n, capture_target=0, seed=seed, point_id=pid,
                           max_scatterings=10**14, n_bins=400, precision=precision, record=record,
                           rk_attempt_budget=RK_BUDGET) for _, seed, pid, n in MULTI]


@pytest.fixture(scope="module")
def multi_points(gpu_ctx):
    return [gpu_ctx.point(dmsun.make_point(**d)) for d, *_ in MULTI]


@pytest.mark.parametrize("mode", [dmsun.MODE_PARAMETER_POINT, dmsun.MODE_CAPTURE])
def test_multi_point_launch_matches_oracle_and_alone(gpu_ctx, oracle_solar, multi_points, mode):
    jobs = multi_jobs(mode)
    print(f"multi-point launch mode={mode} jobs={[(j.point_id, j.max_trajectories) for j in jobs]}", flush=True)
    res = dmsun.run_multi(multi_points, jobs, want_records=True)
    for (d, *_), pt, job, r in zip(MULTI, multi_points, jobs, res):
        ores = oracle_lib.Point(oracle_solar, dmsun.make_point(**d)).run_job(job, record_capacity=job.max_trajectories)
        assert_job_equal(r, ores)
        assert np.array_equal(rec_bytes(r["records"]), rec_bytes(ores["records"])), d
        assert_job_equal(pt.run_job(job, want_records=False), r)


def test_batched_uneven_all_65_points_match_the_oracle(gpu_ctx, oracle_solar):
    """The batched-uneven workload's 65 points (64 SI on m = logspace(0.1, 5
    GeV, 8) x sigma_p = logspace(1e-39, 1e-35 cm^2, 8), plus the dark-photon
    point) in one multi-point launch, with its limits (1e5 scatterings and 1e5
    RKF45 attempts per trajectory): each job equals the oracle's run of its own
    point and descriptor.  150 trajectories per point and 300 x 300 rate tables
    keep the oracle's share to seconds."""
    masses = np.logspace(np.log10(0.1), np.log10(5.0), 8)
    sigmas = np.logspace(-39.0, -35.0, 8)
    descs = [dmsun.make_point(float(m), float(s), interpolation_points=300) for m in masses for s in sigmas]
    descs.append(dmsun.make_point(0.1, 0.0, sigma_e_cm2=1e-37, interaction=dmsun.DARK_PHOTON,
                                  form_factor=dmsun.FF_GENERAL, mediator_mass_MeV=0.01, interpolation_points=300))
    jobs = [dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=150, capture_target=0, seed=20260710,
                           point_id=k, max_scatterings=100_000, rk_attempt_budget=100_000, n_bins=400, record=0)
            for k in range(len(descs))]
    pts = [gpu_ctx.point(d) for d in descs]
    print(f"65-point launch: {sum(j.max_trajectories for j in jobs)} trajectories", flush=True)
    res = dmsun.run_multi(pts, jobs)
    for k, (d, job, r) in enumerate(zip(descs, jobs, res)):
        o = oracle_lib.Point(oracle_solar, d, threads=16).run_job(job, threads=16, want_records=False)
        assert_job_equal(r, o), k
    for p in pts:
        p.close()


def test_refill_batch_knob_keeps_results_and_autotunes(gpu_ctx, multi_points):
    """Waves that wait for 16 free lanes before refilling give every job the
    results of immediate refills (bit-identical records); the knob autotuner
    times the calibration launch per candidate and keeps the fastest."""
    jobs = multi_jobs(dmsun.MODE_PARAMETER_POINT)
    base = dmsun.run_multi(multi_points, jobs, want_records=True)
    try:
        gpu_ctx.set_refill_batch(16)
        batched = dmsun.run_multi(multi_points, jobs, want_records=True)
        best, secs = dmsun.autotune_knobs(multi_points, jobs, refill_candidates=(1, 8, 32), repeats=2)
        print(f"refill_batch candidates (1, 8, 32): {np.round(secs, 4).tolist()} s, best {best}", flush=True)
        assert best in (1, 8, 32) and np.all(secs > 0) and gpu_ctx.refill_batch() == best
        # the first candidate stays unless another is more than 10% faster
        k = [1, 8, 32].index(best)
        assert (k == 0 and np.all(secs[1:] >= 0.90 * secs[0])) or (secs[k] < 0.90 * secs[0] and secs[k] == secs.min())
    finally:
        gpu_ctx.set_refill_batch(1)
    for a, b in zip(base, batched):
        assert_job_equal(b, a)
        assert np.array_equal(rec_bytes(a["records"]), rec_bytes(b["records"]))
    with pytest.raises(dmsun.DmsError):
        gpu_ctx.set_refill_batch(0)


def test_multi_point_mixed_equals_alone(multi_points):
    jobs = multi_jobs(dmsun.MODE_PARAMETER_POINT, precision=dmsun.MIXED)
    print("multi-point launch, mixed precision", flush=True)
    res = dmsun.run_multi(multi_points, jobs, want_records=True)
    for pt, job, r in zip(multi_points, jobs, res):
        alone = pt.run_job(job, want_records=True)
        assert_job_equal(alone, r)
        assert np.array_equal(rec_bytes(alone["records"]), rec_bytes(r["records"]))


def test_multi_rejects_disagreeing_jobs(multi_points):
    a, b = multi_jobs(dmsun.MODE_PARAMETER_POINT)[:2]
    b.n_bins = 300
    with pytest.raises(dmsun.DmsError):
        dmsun.run_multi(multi_points[:2], [a, b])
    c = multi_jobs(dmsun.MODE_PARAMETER_POINT, precision=dmsun.MIXED)[1]
    with pytest.raises(dmsun.DmsError):
        dmsun.run_multi(multi_points[:2], [a, c])


def test_concurrent_jobs_equal_alone(multi_points):
    """Jobs submitted back to back (each on its own streams, capture-target
    jobs with their synchronisation rounds among them) give the results each
    gives alone."""
    jobs = [dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=20000, capture_target=25, seed=21,
                           point_id=k, max_scatterings=10**14, n_bins=400, record=0, rk_attempt_budget=RK_BUDGET)
            for k in range(3)]
    jobs += multi_jobs(dmsun.MODE_CAPTURE, record=0)[:3]
    pts = multi_points[:3] + multi_points[:3]
    print(f"concurrent jobs: {len(jobs)}", flush=True)
    handles = [pt.submit(j) for pt, j in zip(pts, jobs)]
    together = [pt.collect(h) for pt, h in zip(pts, handles)]
    for pt, job, r in zip(pts, jobs, together):
        assert_job_equal(pt.run_job(job, want_records=False), r)


def test_job_memory_budget_chunks_give_same_results(gpu_ctx, multi_points):
    """A small per-job memory budget splits a capture-target job (whose
    trajectories keep per-trajectory rows until their round is decided) into
    more launches; the results stay those of the unsplit job."""
    pt = multi_points[0]
    # 2000 bins: about 40 KB of rows per trajectory, so 64 MiB holds ~1600
    job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=0, capture_target=400, seed=5,
                         max_scatterings=10**14, n_bins=2000, record=0, rk_attempt_budget=RK_BUDGET)
    full = pt.run_job(job, want_records=False)
    gpu_ctx.set_job_memory_budget(64 << 20)
    try:
        small = pt.run_job(job, want_records=False)
    finally:
        gpu_ctx.set_job_memory_budget(32 << 30)
    assert_job_equal(small, full)
    assert small["launches"] > full["launches"]
    with pytest.raises(dmsun.DmsError):
        gpu_ctx.set_job_memory_budget(1 << 20)


def test_fixed_count_job_over_the_memory_budget_runs_in_chunks(gpu_ctx, multi_points):
    """A fixed-count job whose batched launch needs more device memory than the
    job budget (2000 bins: 2.9 GB of per-lane scratch against 64 MiB) runs as
    chunked launches within it, with the results of the unrestricted run."""
    pt = multi_points[0]
    job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=4000, capture_target=0, seed=6,
                         max_scatterings=10**14, n_bins=2000, record=0, rk_attempt_budget=RK_BUDGET)
    full = pt.run_job(job, want_records=False)
    gpu_ctx.set_job_memory_budget(64 << 20)
    try:
        small = pt.run_job(job, want_records=False)
    finally:
        gpu_ctx.set_job_memory_budget(32 << 30)
    print(f"launches: unrestricted {full['launches']}, 64 MiB budget {small['launches']}", flush=True)
    assert_job_equal(small, full)
    assert full["launches"] == 1 and small["launches"] > 1


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_serving_results_same_alone_and_in_the_mix(tmp_path, precision):
    out = tmp_path / "serve.json"
    cmd = [str(SERVE), "--requests", "10", "--rate", "40", "--sample-size", "3", "--sample-size-max", "8",
           "--max-trajectories", "4000", "--max-scatterings", "3000", "--interp", "300", "--bins", "200",
           "--precision", precision, "--check-alone", "10", "--json", str(out)]
    print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, cwd=str(REPO), capture_output=True, text=True, timeout=110)
    assert p.returncode == 0, p.stdout + p.stderr
    s = json.loads(out.read_text())
    assert s["failed"] == 0 and s["alone_mismatches"] == 0 and s["checked_alone"] == 10
    assert 0 < s["latency_p50_sec"] <= s["latency_p99_sec"] <= s["latency_max_sec"]
    assert s["trajectories"] > 0


def test_autotuned_order_keeps_results(multi_points):
    """The autotuner's calibration orders the jobs by measured cost; the
    launch in that order gives every job the results of the given order."""
    jobs = multi_jobs(dmsun.MODE_PARAMETER_POINT, record=0)
    order, cost = dmsun.autotune_multi(multi_points, jobs, calibration_trajectories=64, calibration_rk_attempts=3000)
    print(f"autotuned order {order}, cost per trajectory {np.round(cost, 1).tolist()}", flush=True)
    assert sorted(order) == list(range(len(jobs)))
    assert np.all(cost > 0) and all(cost[order[i]] <= cost[order[i + 1]] for i in range(len(order) - 1))
    tuned = dmsun.run_multi_tuned(multi_points, jobs, calibration_trajectories=64, calibration_rk_attempts=3000)
    plain = dmsun.run_multi(multi_points, jobs)
    for a, b in zip(tuned, plain):
        assert_job_equal(a, b)


def test_serving_requests_match_the_oracle(tmp_path, oracle_solar):
    """The serving workload's request ranges (SI m ~ U_log(0.1, 5) GeV, sigma_p
    ~ U_log(1e-38, 1e-35) cm^2, captured-particle targets 1e3 - 2e3) served
    concurrently by dmsun_serve: every request's counters, visit counts and
    evaporation events equal the oracle's Generate_Data run of the request's
    own descriptor, its histogram sums to 1e-12."""
    res_path = tmp_path / "results.jsonl"
    cmd = [str(SERVE), "--requests", "4", "--rate", "20", "--sample-size", "1000", "--sample-size-max", "2000",
           "--max-trajectories", "40000", "--max-scatterings", "500", "--rk-budget", "200000", "--interp", "300",
           "--bins", "200", "--concurrency", "4", "--seed", "17", "--results", str(res_path), "--quiet"]
    print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, cwd=str(REPO), capture_output=True, text=True, timeout=200)
    assert p.returncode == 0, p.stdout + p.stderr
    lines = [json.loads(l) for l in res_path.read_text().splitlines()]
    assert len(lines) == 4
    for r in lines:
        assert r["status"] == 0
        desc = dmsun.make_point(r["mass_GeV"], r["sigma_p_cm2"], sigma_e_cm2=r["sigma_e_cm2"], interpolation_points=r["interp"])
        h = r["halo"]
        halo = dmsun.default_halo(tuple(h["vobs"]))
        halo.rho_GeV_cm3, halo.v0_kms, halo.vesc_kms = h["rho"], h["v0"], h["vesc"]
        job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=r["max_trajectories"],
                             capture_target=r["sample_size"], seed=r["seed"], point_id=r["point_id"],
                             max_scatterings=r["max_scatterings"], rk_attempt_budget=r["rk_budget"], n_bins=r["bins"],
                             record=0, precision=r["precision"], halo=halo)
        print(f"request {r['id']}: m={r['mass_GeV']:.4g} GeV sigma_p={r['sigma_p_cm2']:.4g} target {r['sample_size']}",
              flush=True)
        o = oracle_lib.Point(oracle_solar, desc, threads=16).run_job(job, threads=16, want_records=False)
        oc = o["counters"]
        for f, v in r["counters"].items():
            got = list(getattr(oc, f)) if f == "reason_counts" else getattr(oc, f)
            assert got == v, (r["id"], f, got, v)
        assert r["counters"]["captured"] >= min(r["sample_size"], 1) or r["counters"]["early_stop_reason"] != 0
        assert np.array_equal(np.array(r["counts"], dtype=np.uint64), o["counts"].ravel())
        hist = np.array(r["hist"])
        assert np.all(np.abs(hist - o["hist"].ravel()) <= 1e-12 * np.maximum(np.abs(o["hist"].ravel()), 1e-300))
        assert [(int(a), float(b)) for a, b in r["events"]] == [(e.trajectory_id, e.lifetime_unbinding_sec)
                                                                for e in o["events"]]
