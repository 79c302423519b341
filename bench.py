#!/usr/bin/env python3
"""Throughput benchmark of the solar dark-matter trajectory path on one MI355X.

Workload (one "step"): a Parameter-point job of the reference
(Simulation_Data::Generate_Data with run_mode "Parameter point"): SI contact
DM-nucleon, m = 0.3 GeV, sigma_p = 1e-36 cm^2, DM_light, 10^6 trajectories
from the Standard Halo Model (v0 = 220, v_esc = 544, |v_sun| = 248 km/s,
rho = 0.4 GeV/cm^3) injected at 1000 AU and shifted to 2 R_sun, AGSS09 solar
model, 1000 x 1000 rate grid, 1000 time-weighted radial bins over [0, 2 R_sun),
evaporation times of valid unbinding events.  Every step simulates a new set of
10^6 trajectory ids (Philox counter blocks), so the work is never cached.

value = trajectories/s of the whole job stream (initial conditions, trajectory
kernel, run accounting and histogram sums), device-resident inputs.  The K timed
steps are K such jobs (distinct trajectory ids) submitted together with
dms_job_submit_batch, so they share one persistent launch and the long bound
orbits of one job overlap the bulk of the others.  The latency of one job on its
own (one launch of 10^6 trajectories, bounded by its longest trajectory) is
reported beside it as single_job.

roofline: the trajectory kernel is FP64-VALU bound (no dense contraction, tiny
HBM traffic).  achieved = algorithmic FP64 operations of the launch (frozen
per-event weights x the kernel's own event counters) / average launch time from
HIP events on the kernel's stream.

cpu_baseline: the CPU restatement (oracle/liboracle.so, one trajectory stream
per host thread) on a bounded sample of the same workload.

Beside the headline line: "mixed" (the same steps on the mixed-precision
kernel), "mixed_validation" (the mixed-precision workload, m = 0.5 GeV and
sigma_p = 1e-37 cm^2, at 10^6 trajectories in both precisions: per-output
differences and speed), "batched_uneven" (65 parameter points of very
different trajectory lengths in one autotuned launch) and "serving" (a Poisson
request stream through the serving loop, p50/p99 latency).  --no-mixed,
--no-validate, --no-uneven, --no-serving and --no-cpu leave them out.
"""
import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "damascus-earth-evap_amd"))

# Algorithmic FP64 operations per event (+, -, *, /, sqrt = 1; each
# transcendental = 1), measured with the op-counting build of the oracle on
# 2000 trajectories of each workload and frozen here: oracle/opcount/weights.txt
# (the RKF45 attempt costs less outside the Sun, where Mass() is constant, so
# the two modes' weights differ with their mix of interior and exterior steps).
FLOP_WEIGHTS = {
    "parameter": {"rk_attempt": 246.2, "step_interior": 400.6, "step_exterior": 191.6, "scatter": 2518.5,
                  "trajectory": 66.8, "segment": 136.8},
    "capture": {"rk_attempt": 233.8, "step_interior": 308.4, "step_exterior": 178.2, "scatter": 2404.4,
                "trajectory": 12.0, "segment": 131.1},
}
PEAK_FP64_TFLOPS = 78.6
# The default timed launch (20 jobs of 10^6 trajectories, the bench's default
# --steps) profiled with rocprofv3: HBM bytes from the FETCH_SIZE and WRITE_SIZE
# passes scaled by a plain 256 MiB copy (read x2.0, write x1.0), and the FP64
# wave instructions weighted by operations (ADD + MUL + TRANS + 2 FMA); see
# profiles/dms_traj_kernel_parameter_point_m0.3_20x1e6.md.  Other step counts
# are other launches: their traffic is reported as null (unmeasured).
PROFILED_LAUNCH = {"steps": 20, "trajectories": 1_000_000, "hbm_bytes": 2.0 * 321_819_444 * 1024 + 477_599_739 * 1024,
                   "f64_op_weighted_wave_instructions": 4.1058e11}


def algorithmic_flops(c, capture_mode):
    w = FLOP_WEIGHTS["capture" if capture_mode else "parameter"]
    segs = c.total_scatterings + c.trajectories
    return (w["rk_attempt"] * c.n_rk_attempts + w["step_interior"] * c.n_steps_interior
            + w["step_exterior"] * c.n_steps_exterior + w["scatter"] * c.total_scatterings
            + w["trajectory"] * c.trajectories + w["segment"] * segs)


def make_workload(dmsun, n_traj, first_id, precision=0):
    desc = dmsun.make_point(0.3, 1e-36, interpolation_points=1000)
    job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=n_traj, capture_target=0,
                         max_scatterings=10**14, n_bins=1000, first_trajectory_id=first_id, record=0,
                         precision=precision, halo=dmsun.default_halo((0.0, 248.0, 0.0)))
    return desc, job


def cpu_baseline(sample, threads):
    sys.path.insert(0, str(ROOT / "tests"))
    import dmsun
    import oracle_lib

    solar = oracle_lib.Solar()
    desc, job = make_workload(dmsun, sample, 0)
    pt = oracle_lib.Point(solar, desc, threads=threads)
    t0 = time.perf_counter()
    res = pt.run_job(job, threads=threads, want_records=False)
    dt = time.perf_counter() - t0
    c = res["counters"]
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return dict(value=c.trajectories / dt, unit="trajectories/s", cores=threads, kind="port",
                host=f"{model}; {os.cpu_count()} logical CPUs visible",
                sample=f"{sample} trajectories of the same workload (ids 0..{sample - 1}), "
                       f"{threads} threads, {dt:.2f} s; scatterings/s {c.total_scatterings / dt:.4g}",
                longest_trajectory_rk_attempts=int(c.max_rk_attempts),
                note="the sample's longest trajectory needs fewer RKF45 attempts than the longest of a 1e6 job "
                     "(single_job) or of the timed launch (tail), so trajectories/s here and on the GPU are not "
                     "like-for-like; the roofline fraction measures the kernel")


def measure(dmsun, pt, precision, args):
    """One untimed job alone (reported as single_job), then the K timed steps as
    one batch (one launch)."""
    name = "fp64" if precision == dmsun.FP64 else "mixed"
    single = None
    for w in range(args.warmup):
        _, job = make_workload(dmsun, args.trajectories, (1000 + w) * args.trajectories, precision)
        t = time.perf_counter()
        res = pt.run_job(job, want_records=False)
        dt = time.perf_counter() - t
        single = dict(trajectories_per_s=res["counters"].trajectories / dt, seconds=dt, kernel_ms=res["kernel_ms"],
                      trajectories=res["counters"].trajectories,
                      longest_trajectory_rk_attempts=res["counters"].max_rk_attempts,
                      us_per_attempt_of_longest=1e3 * res["kernel_ms"] / max(1, res["counters"].max_rk_attempts),
                      note="bounded by its longest trajectory; cpu_baseline's sample has a shorter tail, so the "
                           "ratio of the two is not like-for-like")
        print(f"{name} warmup {w} (one job alone): {dt:.2f} s, kernel {res['kernel_ms']:.1f} ms", file=sys.stderr,
              flush=True)
    jobs = [make_workload(dmsun, args.trajectories, k * args.trajectories, precision)[1] for k in range(args.steps)]
    t0 = time.perf_counter()
    results = pt.run_batch(jobs, want_records=False)
    wall = time.perf_counter() - t0
    out = dict(traj=0, scat=0, cap=0, flops=0.0, wall=wall, single=single, kern_ms=results[0]["kernel_ms"],
               launches=results[0]["launches"], max_rk=max(r["counters"].max_rk_attempts for r in results))
    for k, res in enumerate(results):
        c = res["counters"]
        out["traj"] += c.trajectories
        out["scat"] += c.total_scatterings
        out["cap"] += c.captured
        out["flops"] += algorithmic_flops(c, False)
        print(f"{name} step {k}: {c.trajectories} trajectories, {c.total_scatterings} scatterings, {c.captured} captured",
              file=sys.stderr, flush=True)
    print(f"{name} timed: {args.steps} jobs in one launch, kernel {out['kern_ms']:.1f} ms, wall {wall:.2f} s",
          file=sys.stderr, flush=True)
    return out


def uneven_points(dmsun, ctx):
    """64 SI points, m log-spaced over [0.1, 5] GeV x sigma_p log-spaced over
    [1e-39, 1e-35] cm^2, plus a dark-photon point (m = 0.1 GeV, sigma_e =
    1e-37 cm^2, light mediator); 1000 x 1000 rate tables built on the GPU."""
    import numpy as np
    masses = np.logspace(np.log10(0.1), np.log10(5.0), 8)
    sigmas = np.logspace(-39.0, -35.0, 8)
    descs = [dmsun.make_point(float(m), float(s), interpolation_points=1000) for m in masses for s in sigmas]
    descs.append(dmsun.make_point(0.1, 0.0, sigma_e_cm2=1e-37, interaction=dmsun.DARK_PHOTON,
                                  form_factor=dmsun.FF_GENERAL, mediator_mass_MeV=0.01, interpolation_points=1000))
    t = time.perf_counter()
    pts = [ctx.point(d) for d in descs]
    return pts, time.perf_counter() - t


def measure_uneven(dmsun, ctx, args):
    """Batched uneven workload: the 65 points' jobs in one persistent launch
    with lane refill, job order from the calibration autotuner."""
    pts, t_tables = uneven_points(dmsun, ctx)
    out = {"workload": f"65 parameter points (64 SI m=0.1-5 GeV x sigma_p=1e-39-1e-35 cm^2, 1 dark photon), "
                       f"{args.uneven_trajectories} trajectories each, at most 1e5 scatterings and 1e5 RKF45 attempts "
                       f"(wall-time guard analogue) per trajectory, 1000 bins, one launch",
           "points": len(pts), "rate_tables_s": t_tables}
    for name, prec in (("fp64", dmsun.FP64), ("mixed", dmsun.MIXED)):
        if name == "mixed" and args.no_mixed:
            continue
        jobs = [dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=args.uneven_trajectories,
                               capture_target=0, seed=20260710, point_id=k, max_scatterings=100_000,
                               rk_attempt_budget=100_000, n_bins=1000, record=0, precision=prec,
                               halo=dmsun.default_halo((0.0, 248.0, 0.0))) for k in range(len(pts))]
        t = time.perf_counter()
        order, cost = dmsun.autotune_multi(pts, jobs)
        refill, refill_secs = dmsun.autotune_knobs(pts, jobs, refill_candidates=(1, 4, 16), repeats=2)
        t_cal = time.perf_counter() - t
        t = time.perf_counter()
        res = dmsun.run_multi([pts[k] for k in order], [jobs[k] for k in order])
        wall = time.perf_counter() - t
        traj = sum(r["counters"].trajectories for r in res)
        scat = sum(r["counters"].total_scatterings for r in res)
        trunc = sum(r["counters"].computational_truncations for r in res)
        leg = {"trajectories_per_s": traj / wall, "scatterings_per_s": scat / wall, "seconds": wall,
               "calibration_s": t_cal, "trajectories": traj, "scatterings": scat, "truncated": trunc,
               "kernel_ms": res[0]["kernel_ms"], "cost_per_trajectory_range": [float(cost.min()), float(cost.max())],
               "refill_batch": refill, "refill_candidates_s": {str(k): float(v) for k, v in zip((1, 4, 16), refill_secs)}}
        if name == "fp64":
            t = time.perf_counter()
            dmsun.run_multi(pts, jobs)
            leg["given_order_seconds"] = time.perf_counter() - t
        out[name] = leg
        print(f"uneven {name}: {traj} trajectories, {scat} scatterings in {wall:.2f} s (calibration {t_cal:.2f} s)",
              file=sys.stderr, flush=True)
    ctx.set_refill_batch(1)
    for p in pts:
        p.close()
    return out


def measure_serving(args):
    """Serving loop (dmsun_serve): Poisson parameter-point requests whose
    trajectories join one resident launch as slots; p50/p99 latency,
    trajectories/s, and the first requests rerun alone and compared."""
    import subprocess
    exe = ROOT / "damascus-earth-evap_amd" / "dmsun_serve"
    cmd = [str(exe), "--requests", str(args.serve_requests), "--rate", str(args.serve_rate),
           "--sample-size", "1000", "--sample-size-max", "100000", "--max-trajectories", "2000000",
           "--max-scatterings", "100000", "--rk-budget", "100000", "--interp", "1000", "--bins", "1000",
           "--check-alone", str(args.serve_check_alone), "--quiet"]
    print("serving: " + " ".join(cmd[1:]), file=sys.stderr, flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    if p.returncode != 0:
        raise RuntimeError("dmsun_serve failed: " + p.stderr[-2000:])
    s = json.loads(p.stdout.strip().splitlines()[-1])
    s["workload"] = ("Poisson requests, SI m~U_log(0.1, 5) GeV, sigma_p~U_log(1e-38, 1e-35) cm^2, "
                     "parameter-point mode with a captured-particle target ~U_log(1e3, 1e5) (at most 2e6 trajectories), "
                     "at most 1e5 scatterings and 1e5 RKF45 attempts per trajectory, 1000x1000 rate tables resident; "
                     "every request's trajectories join one resident launch as slots; the first "
                     f"{args.serve_check_alone} requests rerun alone and compared")
    return s


def measure_bulk(dmsun, pt, precision):
    """The same 10^6 trajectories with every trajectory cut at 2000 RKF45
    attempts: no long tail, all lanes busy for the whole launch.  Shows the
    kernel's throughput when it is not waiting on a few long trajectories."""
    _, job = make_workload(dmsun, 1_000_000, 0, precision)
    job.rk_attempt_budget = 2000
    res = pt.run_job(job, want_records=False)
    c = res["counters"]
    s = res["kernel_ms"] * 1e-3
    fl = algorithmic_flops(c, False)
    return {"workload": "the timed workload's first 1e6 trajectories, each cut at 2000 RKF45 attempts",
            "kernel_ms": res["kernel_ms"], "rk_attempts_per_s": c.n_rk_attempts / s,
            "achieved_tflops": fl / s / 1e12, "frac": fl / s / 1e12 / PEAK_FP64_TFLOPS,
            "truncated": c.computational_truncations}


def measure_lone(dmsun, ctx, pt, precision, tid=16020326):
    """One long trajectory of the timed workload (id 16020326: 193,570 RKF45
    attempts in FP64) launched alone, with the two-wave split off
    (lone_serial) and on: launch time per RKF45 attempt of a lone lane, and
    whether the two records are identical."""
    import numpy as np
    out, recs = {}, {}
    try:
        for serial in (1, 0):
            ctx.set_lone_serial(serial)
            _, job = make_workload(dmsun, 1, tid, precision)
            job.record_trajectories = 1
            res = pt.run_job(job, want_records=True)
            n_rk = int(res["records"][0].n_rk_attempts)
            recs[serial] = np.ctypeslib.as_array(res["records"]).view(np.uint8).copy()
            out["serial_us_per_attempt" if serial else "split_us_per_attempt"] = 1e3 * res["kernel_ms"] / max(1, n_rk)
            out["rk_attempts"] = n_rk
    finally:
        ctx.set_lone_serial(0)
    out["trajectory_id"] = tid
    out["speedup"] = out["serial_us_per_attempt"] / out["split_us_per_attempt"]
    out["records_identical"] = bool(np.array_equal(recs[0], recs[1]))
    return out


def validate_mixed(dmsun, ctx, n):
    """The mixed-precision workload at full size (SI m = 0.5 GeV, sigma_p =
    1e-37 cm^2, n trajectories, parameter point): FP64 and mixed kernels on the
    same trajectory ids, per-output differences and speed."""
    import math
    import numpy as np
    from scipy import stats

    pt = ctx.point(dmsun.make_point(0.5, 1e-37, interpolation_points=1000))
    res = {}
    for name, prec in (("fp64", dmsun.FP64), ("mixed", dmsun.MIXED)):
        job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=n, capture_target=0,
                             max_scatterings=10**14, n_bins=1000, record=0, precision=prec,
                             halo=dmsun.default_halo((0.0, 248.0, 0.0)))
        t = time.perf_counter()
        r = pt.run_job(job, want_records=False)
        r["wall"] = time.perf_counter() - t
        res[name] = r
        print(f"validate-mixed {name}: {r['wall']:.2f} s", file=sys.stderr, flush=True)
    a, b = res["fp64"], res["mixed"]
    ca, cb = a["counters"], b["counters"]
    pa, pb = ca.captured / n, cb.captured / n
    sig = math.sqrt(pa * (1 - pa) / n + pb * (1 - pb) / n)

    def bin_err(s, s2, cnt):
        if cnt <= 1:
            return np.zeros_like(s)
        m = s / cnt
        return np.sqrt(cnt * np.maximum(0.0, s2 / cnt - m * m))

    chi2, dof = 0.0, 0
    for base, na, nb in ((0, ca.captured, cb.captured), (4, ca.completed_outward_escapes, cb.completed_outward_escapes)):
        for q in (0, 1):
            ha, hb = a["hist"][base + q], b["hist"][base + q]
            var = bin_err(ha, a["hist"][base + 2 + q], na) ** 2 + bin_err(hb, b["hist"][base + 2 + q], nb) ** 2
            use = var > 0
            chi2 += float(np.sum((ha[use] - hb[use]) ** 2 / var[use]))
            dof += int(use.sum())
    la = np.array([e.lifetime_unbinding_sec for e in a["events"]])
    lb = np.array([e.lifetime_unbinding_sec for e in b["events"]])
    out = {"workload": f"parameter point SI m=0.5 GeV sigma_p=1e-37 cm^2, {n} trajectories, 1000 bins, same ids",
           "capture_fraction": [pa, pb], "capture_delta_over_sigma": abs(pa - pb) / sig if sig > 0 else 0.0,
           "hist_chi2_per_dof": chi2 / max(dof, 1), "hist_bins": dof,
           "evaporation_events": [len(la), len(lb)],
           "scatterings_mean_rel_diff": (cb.total_scatterings - ca.total_scatterings) / max(1, ca.total_scatterings)}
    if min(len(la), len(lb)) >= 20:
        out["evaporation_ks_p"] = float(stats.ks_2samp(la, lb).pvalue)
        out["evaporation_median_rel_diff"] = float(np.median(lb) / np.median(la) - 1.0)
        out["evaporation_p90_rel_diff"] = float(np.percentile(lb, 90) / np.percentile(la, 90) - 1.0)
    for name, r in res.items():
        c = r["counters"]
        out[name] = {"seconds": r["wall"], "trajectories_per_s": c.trajectories / r["wall"],
                     "scatterings_per_s": c.total_scatterings / r["wall"], "longest_trajectory_rk_attempts": c.max_rk_attempts,
                     "us_per_attempt_of_longest": 1e3 * r["kernel_ms"] / max(1, c.max_rk_attempts)}
    out["speedup_whole_job"] = a["wall"] / b["wall"]
    out["speedup_per_attempt_of_longest"] = out["fp64"]["us_per_attempt_of_longest"] / out["mixed"]["us_per_attempt_of_longest"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--trajectories", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample", type=int, default=100_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-validate", action="store_true", help="leave out the full-size mixed-precision validation")
    ap.add_argument("--uneven-trajectories", type=int, default=4000)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--serve-rate", type=float, default=1.0)
    ap.add_argument("--serve-check-alone", type=int, default=4)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run one plain device-to-device copy of BYTES bytes (scale of the HBM byte counters)")
    ap.add_argument("--validate-mixed", type=int, default=0, metavar="N",
                    help="only run the mixed-precision workload (m=0.5 GeV, sigma_p=1e-37) with N trajectories in FP64 "
                         "and mixed precision and print the per-output differences")
    args = ap.parse_args()
    if args.validate_mixed > 0:
        import dmsun
        print(json.dumps(validate_mixed(dmsun, dmsun.Context(device=0), args.validate_mixed)), flush=True)
        return
    if args.copy_probe > 0:
        import torch
        src = torch.ones(args.copy_probe, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(trajectories are independent: several GPUs would each take a disjoint trajectory-id range of the same keys).", file=sys.stderr)
        sys.exit(2)

    import dmsun

    ctx = dmsun.Context(device=0)
    desc, _ = make_workload(dmsun, args.trajectories, 0)
    pt = ctx.point(desc)
    fp64 = measure(dmsun, pt, dmsun.FP64, args)
    mixed = measure(dmsun, pt, dmsun.MIXED, args) if not args.no_mixed else None
    tot_traj, tot_scat, tot_cap = fp64["traj"], fp64["scat"], fp64["cap"]
    wall, flops, kern_ms, launches, single = fp64["wall"], fp64["flops"], fp64["kern_ms"], fp64["launches"], fp64["single"]

    avg_launch_s = (kern_ms / launches) * 1e-3
    profiled = (args.steps, args.trajectories) == (PROFILED_LAUNCH["steps"], PROFILED_LAUNCH["trajectories"])
    achieved = flops / launches / avg_launch_s / 1e12
    out = {
        "metric": "trajectories/s (FP64 GPU kernel; mixed-precision kernel and scatterings/s alongside)",
        "value": tot_traj / wall,
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Standard Halo Model initial conditions from a Philox stream; AGSS09 solar model)",
        "config": {"workload": "parameter point: SI contact m=0.3 GeV sigma_p=1e-36 cm^2, 1e6 trajectories per "
                               "step (one job), 1000 radial bins, 1000x1000 rate grid, precision fp64; "
                               "the timed steps share one persistent launch",
                   "trajectories_per_step": args.trajectories},
        "single_job": single,
        "scatterings_per_s": tot_scat / wall,
        "capture_fraction": tot_cap / max(1, tot_traj),
        "kernel_ms_per_step": kern_ms / args.steps,
        "roofline": {"bound": "valu-fp64", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": PROFILED_LAUNCH["hbm_bytes"] if profiled else None,
                     "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE of this launch, not live)",
                     "counter_bound_tflops": (64 * PROFILED_LAUNCH["f64_op_weighted_wave_instructions"] / avg_launch_s / 1e12
                                              if profiled else None),
                     "counter_bound_note": "upper bound from the profiled FP64 VALU wave-instruction counters x 64 lanes "
                                           "(counters do not see inactive lanes)",
                     "kernel": "dms_traj_kernel", "avg_launch_ms": avg_launch_s * 1e3},
    }
    if mixed is not None:
        m_launch_s = mixed["kern_ms"] / mixed["launches"] * 1e-3
        out["mixed"] = {
            "value": mixed["traj"] / mixed["wall"], "unit": "trajectories/s",
            "scatterings_per_s": mixed["scat"] / mixed["wall"],
            "capture_fraction": mixed["cap"] / max(1, mixed["traj"]),
            "speedup_vs_fp64": (mixed["traj"] / mixed["wall"]) / out["value"],
            "single_job": mixed["single"],
            "single_job_speedup_vs_fp64": (mixed["single"]["trajectories_per_s"] / single["trajectories_per_s"]
                                           if mixed["single"] and single else None),
            "kernel": "dms_traj_kernel_mixed", "avg_launch_ms": m_launch_s * 1e3,
            "achieved_tflops_fp64_equivalent": mixed["flops"] / mixed["launches"] / m_launch_s / 1e12,
            "precision": "FP32 rate table, optical depth (Kahan), target sampling, kinematics; "
                         "FP64 RKF45, energy test, time sums",
        }
    out["tail"] = {
        "note": "launch time of a parameter-point job is bounded by its longest trajectory, whose RKF45 attempts are "
                "sequential; us_per_attempt_of_longest = launch time / that trajectory's attempts",
        "longest_trajectory_rk_attempts": fp64["max_rk"],
        "us_per_attempt_of_longest": 1e3 * kern_ms / max(1, fp64["max_rk"]),
        "mixed_longest_trajectory_rk_attempts": mixed["max_rk"] if mixed else None,
        "mixed_us_per_attempt_of_longest": 1e3 * mixed["kern_ms"] / max(1, mixed["max_rk"]) if mixed else None,
        "lone_split_note": "a trajectory left alone in its wave has its next step prepared by a finished wave of "
                           "its block while its own wave post-processes the current one (on in every launch above); "
                           "lone_split times one long trajectory launched alone with that split off and on",
        "lone_split": {"fp64": measure_lone(dmsun, ctx, pt, dmsun.FP64)},
    }
    if mixed is not None:
        out["tail"]["lone_split"]["mixed"] = measure_lone(dmsun, ctx, pt, dmsun.MIXED)
    out["bulk"] = {"fp64": measure_bulk(dmsun, pt, dmsun.FP64)}
    if mixed is not None:
        out["bulk"]["mixed"] = measure_bulk(dmsun, pt, dmsun.MIXED)
        # kernel speed ratios that do not depend on which trajectory of a launch
        # happens to be the longest (the two kernels' histories separate through
        # rounding, so their longest trajectories differ)
        out["mixed"]["speedup_vs_fp64_bulk"] = (out["bulk"]["mixed"]["rk_attempts_per_s"]
                                                / out["bulk"]["fp64"]["rk_attempts_per_s"])
        out["mixed"]["speedup_vs_fp64_per_attempt_of_longest"] = (out["tail"]["us_per_attempt_of_longest"]
                                                                  / out["tail"]["mixed_us_per_attempt_of_longest"])
    if not args.no_mixed and not args.no_validate:
        out["mixed_validation"] = validate_mixed(dmsun, ctx, 1_000_000)
    if not args.no_uneven:
        out["batched_uneven"] = measure_uneven(dmsun, ctx, args)
    if not args.no_serving:
        out["serving"] = measure_serving(args)
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
