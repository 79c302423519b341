"""The C++ host API and driver (dmsun_run) on the GPU against the oracle's
Generate_Data restatement: same counters, histograms and evaporation times in
the reference's output files."""
import pathlib
import subprocess

import numpy as np
import pytest

import dmsun
import oracle_lib

pytestmark = pytest.mark.gpu

REPO = pathlib.Path(__file__).resolve().parent.parent
RUN = REPO / "damascus-earth-evap_amd" / "dmsun_run"

CFG = """ID = "host_api_{mode}";
run_mode = "{run_mode}";
sample_size = {sample};
max_trajectories = {maxt};
fixed_seed = 20260710;
interpolation_points = 300;
output_dir = "{out}";
DM_mass = 0.3;
DM_spin = 0.5;
DM_fraction = 1.0;
DM_light = true;
DM_interaction = "SI";
DM_isospin_conserved = true;
DM_cross_section_nucleon = 1.0e-36;
DM_cross_section_electron = 0.0;
DM_form_factor = "Contact";
DM_distribution = "SHM";
DM_local_density = 0.4;
SHM_v0 = 220.0;
SHM_vObserver = (0.0, 248.0, 0.0);
SHM_vEscape = 544.0;
"""


def header(path):
    return {l[2:].split(" = ")[0]: l.split(" = ")[1] for l in path.read_text().splitlines()
            if l.startswith("# ") and " = " in l}


def oracle_job(oracle_solar, mode, sample, maxt, nbins):
    desc = dmsun.make_point(0.3, 1e-36, interpolation_points=300)
    job = dmsun.make_job(mode=mode, max_trajectories=maxt, capture_target=sample, max_scatterings=10**14,
                         n_bins=nbins, halo=dmsun.default_halo((0.0, 248.0, 0.0)))
    return oracle_lib.Point(oracle_solar, desc).run_job(job, want_records=False)


def test_parameter_point_run_matches_oracle(oracle_solar, tmp_path):
    cfg = tmp_path / "pp.cfg"
    cfg.write_text(CFG.format(mode="pp", run_mode="Parameter point", sample=60, maxt=1500, out=tmp_path))
    p = subprocess.run([str(RUN), str(cfg), "--bins", "400"], cwd=str(REPO), capture_output=True, text=True, timeout=120)
    print(p.stdout[-2000:], p.stderr[-2000:], flush=True)
    assert p.returncode == 0
    res_dir = next(tmp_path.glob("results_*"))
    h = header(res_dir / "bincount.txt")
    o = oracle_job(oracle_solar, dmsun.MODE_PARAMETER_POINT, 60, 1500, 400)
    c = o["counters"]
    assert int(h["total_trajectories"]) == c.trajectories and int(h["captured_particles"]) == c.captured
    assert int(h["completed_outward_escapes"]) == c.completed_outward_escapes
    assert int(h["numerical_failures"]) == c.numerical_failures and int(h["total_scatterings"]) == c.total_scatterings
    assert int(h["capture_target_overshoot"]) == c.capture_target_overshoot
    rows = np.array([[float(x) for x in l.split("\t")] for l in (res_dir / "bincount.txt").read_text().splitlines()
                     if not l.startswith("#")])
    assert rows.shape == (400, 9)
    for col, q in [(1, 0), (2, 1), (5, 4), (6, 5)]:
        ref = o["hist"][q]
        assert np.allclose(rows[:, col], ref, rtol=1e-9, atol=0.0)  # 10 printed digits
    ev = [l.split("\t") for l in (res_dir / "evaporation_times.txt").read_text().splitlines() if not l.startswith("#")]
    want = sorted((e.lifetime_unbinding_sec, e.rank, e.trajectory_id) for e in o["events"]
                  if e.lifetime_unbinding_sec >= 0)
    assert [(int(r), int(t)) for r, t, _ in ev] == [(r, t) for _, r, t in want]


def test_capture_run_matches_oracle(oracle_solar, tmp_path):
    cfg = tmp_path / "cap.cfg"
    cfg.write_text(CFG.format(mode="cap", run_mode="Capture", sample=40, maxt=0, out=tmp_path))
    p = subprocess.run([str(RUN), str(cfg)], cwd=str(REPO), capture_output=True, text=True, timeout=120)
    print(p.stdout[-2000:], p.stderr[-2000:], flush=True)
    assert p.returncode == 0
    lines = dict(l.split(":", 1) for l in p.stdout.splitlines() if ":" in l)
    o = oracle_job(oracle_solar, dmsun.MODE_CAPTURE, 40, 0, 2000)
    assert int(lines["Simulated trajectories"].strip()) == o["counters"].trajectories
    assert int(lines["Captured count"].strip()) == o["counters"].captured
    assert not list(tmp_path.glob("results_*"))  # capture mode writes no files (src/main.cpp:93-94)


def test_progress_is_live_and_monotone(gpu_ctx):
    """dms_job_poll reports finished trajectories while the launch runs
    (device counters read on a second stream), ending at the job's total."""
    import ctypes as C
    import time
    L = dmsun.lib()
    pt = gpu_ctx.point(dmsun.make_point(0.3, 1e-36, interpolation_points=1000))
    job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=60000, capture_target=0,
                         max_scatterings=10**14, n_bins=1000, record=0)
    h = C.c_void_p()
    assert L.dms_job_submit(pt.h, C.byref(job), C.byref(h)) == 0
    seen = []
    st = dmsun.JobStatus()
    t0 = time.time()
    while True:
        assert L.dms_job_poll(h, C.byref(st)) == 0
        seen.append(st.trajectories_done)
        if st.done or time.time() - t0 > 100:
            break
        time.sleep(0.01)
    assert L.dms_job_wait(h, C.byref(st)) == 0
    L.dms_job_destroy(h)
    print(f"{len(seen)} polls, distinct values {len(set(seen))}", flush=True)
    assert all(a <= b for a, b in zip(seen, seen[1:]))
    assert st.trajectories_done == 60000
    assert len(set(v for v in seen if 0 < v < 60000)) >= 2  # intermediate progress was visible


def snapshot_reports(directory):
    """snapshot_<t>s.txt reports in checkpoint order: (label, header keys, histogram rows)."""
    out = []
    for f in directory.glob("snapshot_*s.txt"):
        label = int(f.name[len("snapshot_"):-len("s.txt")])
        lines = f.read_text().splitlines()
        kv = {l[2:].split(" = ")[0]: l.split(" = ", 1)[1] for l in lines if l.startswith("# ") and " = " in l}
        rows = [list(map(float, l.split("\t"))) for l in lines if l and not l.startswith("#")]
        out.append((label, kv, np.array(rows)))
    return sorted(out, key=lambda t: t[0])


def wilson(n, k):
    if n == 0:
        return 0.0, 0.0, 0.0, 0.0
    p = min(k, n) / n
    z = 1.96
    den = 1 + z * z / n
    c = p + z * z / (2 * n)
    h = z * np.sqrt(p * (1 - p) / n + z * z / (4 * n * n))
    return p, np.sqrt(p * (1 - p) / n), max(0.0, (c - h) / den), min(1.0, (c + h) / den)


def run_with_snapshots(tmp_path, sample, nb):
    cfg = tmp_path / "snap.cfg"
    cfg.write_text(CFG.format(mode="snap", run_mode="Parameter point", sample=sample, maxt=0, out=tmp_path)
                   + "snapshot_enabled = true;\nsnapshot_interval = 1;\n")
    p = subprocess.run([str(RUN), str(cfg), "--bins", str(nb)], cwd=str(REPO), capture_output=True, text=True, timeout=200)
    assert p.returncode == 0, p.stderr
    return snapshot_reports(tmp_path / "snapshots")


def test_final_snapshot_report_matches_the_oracle(oracle_solar, tmp_path):
    """The snapshot report at the first checkpoint after the job (the
    reference's report keys, src/Snapshot_IO.cpp:508-646) carries the job's
    final counters, capture-rate estimates with Wilson intervals and
    histograms, which equal the oracle's Generate_Data run."""
    nb = 100
    reps = run_with_snapshots(tmp_path, 5000, nb)
    label, kv, rows = reps[-1]
    o = oracle_job(oracle_solar, dmsun.MODE_PARAMETER_POINT, 5000, 0, nb)
    c = o["counters"]
    print(f"{len(reps)} snapshot reports; final {label} s: {kv['total_trajectories']} trajectories", flush=True)
    assert kv["snapshot_status"] == "merged" and int(kv["snapshot_target_wall_time_s"]) == label
    assert int(kv["total_trajectories"]) == c.trajectories and int(kv["captured_particles"]) == c.captured
    valid = c.captured + c.completed_outward_escapes
    assert int(kv["valid_trajectories"]) == valid and int(kv["numerical_failures"]) == c.numerical_failures
    assert int(kv["unresolved_not_captured_trajectories"]) == c.trajectories - valid
    for prefix, n in (("capture_rate_raw", c.trajectories), ("capture_rate_valid", valid)):
        rate, err, lo, hi = wilson(n, c.captured)
        for suffix, want in (("", rate), ("_err", err), ("_CI_95_lower", lo), ("_CI_95_upper", hi)):
            assert abs(float(kv[prefix + suffix]) - want) <= 6e-9, (prefix + suffix, kv[prefix + suffix], want)
    assert rows.shape == (nb, 9)
    h = o["hist"]
    for col, q in ((1, 0), (2, 1), (5, 4), (6, 5)):
        assert np.allclose(rows[:, col], h[q], rtol=2e-10, atol=0.0)


def test_live_snapshot_reports_are_consistent(tmp_path):
    """While a longer job runs (10^5 captured particles), a report every second
    from the device's live counts: non-decreasing totals, captured <= valid <=
    total, no histogram section until the final report.  (Live counts include
    the trajectories a capture-target launch runs past its stop point, which
    the final round accounting leaves out, so the final report's totals can
    be below the last live ones.)"""
    reps = run_with_snapshots(tmp_path, 100000, 100)
    print(f"{len(reps)} snapshot reports at {[r[0] for r in reps]} s", flush=True)
    assert len(reps) >= 2
    keys = ["total_trajectories", "captured_particles", "valid_trajectories", "numerical_failures"]
    prev = [0] * 4
    for label, kv, rows in reps[:-1]:
        cur = [int(kv[k]) for k in keys]
        assert all(a <= b for a, b in zip(prev, cur)), (label, prev, cur)
        assert cur[1] <= cur[2] <= cur[0] and int(kv["snapshot_target_wall_time_s"]) == label
        prev = cur
    assert all(r[2].size == 0 for r in reps[:-1]) and reps[-1][2].shape == (100, 9)
    final = reps[-1][1]
    assert int(final["captured_particles"]) >= 100000 and reps[-1][0] > reps[-2][0]
    assert int(final["captured_particles"]) <= int(final["valid_trajectories"]) <= int(final["total_trajectories"])
