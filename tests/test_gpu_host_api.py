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


def test_driver_writes_progress_snapshots(tmp_path):
    cfg = tmp_path / "snap.cfg"
    cfg.write_text(CFG.format(mode="snap", run_mode="Parameter point", sample=30, maxt=800, out=tmp_path)
                   + "snapshot_enabled = true;\nsnapshot_interval = 1;\n")
    p = subprocess.run([str(RUN), str(cfg), "--bins", "100"], cwd=str(REPO), capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    snap = (tmp_path / "snapshots" / "snapshot_progress.txt").read_text()
    kv = dict(l.split(" = ") for l in snap.splitlines() if " = " in l)
    assert kv["finished"] == "1" and int(kv["trajectories_done"]) > 0


def test_solar_model_non_square_grid():
    """tests/test_Solar_Model.cpp:192-224 through the C++ Solar_Model: a 31 x 29
    grid is exact at its nodes and reports both counts; (1, 29) turns the grid
    off (0 x 0) and the total rate is the directly computed one."""
    exe = REPO / "tests" / "harness" / "host_unit"
    pkg = REPO / "damascus-earth-evap_amd"
    if not exe.exists():
        subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe), str(REPO / "tests" / "harness" / "host_unit.cpp"),
                        "-L", str(pkg), "-ldmsun_host", "-ldmsun", f"-Wl,-rpath,{pkg}"], check=True)
    p = subprocess.run([str(exe), "grid"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    out = dict(l.split("=", 1) for l in p.stdout.splitlines() if "=" in l)
    assert out["radius_points"] == "31" and out["speed_points"] == "29"
    assert float(out["node_err_over_tolerance"]) <= 1.0
    assert out["off_radius_points"] == "0" and out["off_speed_points"] == "0" and out["off_equal"] == "1"
