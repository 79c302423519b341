"""Per-event floating-point operation counts of the trajectory path, measured
with the op-counting build of the oracle (liboracle_opcount.so; MEASUREMENT
INFRASTRUCTURE).  Runs trajectories of a workload one at a time and prints, per
part of a trajectory, the operations per event the kernel counts:

  rk_attempt       per RKF45 attempt (n_rk_attempts)
  step_interior    per accepted step inside the Sun: preparation and
                   post-processing, the attempts excluded (n_steps_interior)
  step_exterior    the same outside the Sun (n_steps_exterior)
  step_other       per other step iteration: optical-depth retries and a
                   segment's last iteration (n_tau_rejections + segments)
  scatter          per scattering (total_scatterings)
  trajectory, segment  the trajectory's own bookkeeping, fitted as
                   trajectory + segment x (scatterings + 1)

Usage: python3 oracle/opcount/count_ops.py [--n N] [--mode parameter|capture]
"""
import argparse
import ctypes as C
import pathlib
import subprocess
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO / "damascus-earth-evap_amd"))
import dmsun  # noqa: E402

SO = HERE / "liboracle_opcount.so"
PARTS = ["rk", "interior", "exterior", "step_other", "scatter", "trajectory", "outside"]
_P = C.POINTER


def load():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    L = C.CDLL(str(SO))
    vp = C.c_void_p
    for name, res, args in [
        ("or_solar_create", vp, [_P(C.c_double), C.c_size_t]),
        ("or_point_create", vp, [vp, _P(dmsun.PointDesc), C.c_int]),
        ("or_initial_conditions", C.c_int, [vp, _P(dmsun.Halo), C.c_uint32, C.c_uint32, C.c_uint64, C.c_size_t,
                                           C.c_double, _P(dmsun.Event), _P(C.c_int32)]),
        ("or_simulate_events", C.c_int, [vp, _P(dmsun.JobDesc), _P(dmsun.Event), C.c_size_t, _P(dmsun.TrajRecord),
                                        _P(C.c_double), _P(C.c_uint32), C.c_int]),
        ("or_opcount_read", None, [_P(C.c_uint64), C.c_int]),
    ]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def workload(mode):
    """The bench's parameter-point workload, or the capture workload (config 1)."""
    if mode == "parameter":
        pdesc = dmsun.make_point(0.3, 1e-36, interpolation_points=1000)
        job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=1, capture_target=0,
                             max_scatterings=10**14, n_bins=1000, record=1, halo=dmsun.default_halo((0.0, 248.0, 0.0)))
    else:
        pdesc = dmsun.make_point(1.0, 1e-35, interpolation_points=1000)
        job = dmsun.make_job(mode=dmsun.MODE_CAPTURE, max_trajectories=1, capture_target=0, max_scatterings=1000,
                             n_bins=1000, record=1, halo=dmsun.default_halo((0.0, 248.0, 0.0)))
    return pdesc, job


def count(mode="parameter", n=2000, L=None):
    """Run trajectory ids 0..n-1 of the workload one at a time through the
    counting build; returns (weights, details)."""
    L = L or load()
    rows = dmsun.load_solar_rows()
    S = L.or_solar_create(rows.ctypes.data_as(_P(C.c_double)), rows.shape[0])
    pdesc, job = workload(mode)
    P = L.or_point_create(S, C.byref(pdesc), 8)
    ev = (dmsun.Event * n)()
    ok = np.zeros(n, dtype=np.int32)
    assert L.or_initial_conditions(S, C.byref(job.halo), job.seed, job.point_id, 0, n, job.r_max_rsun, ev,
                                   ok.ctypes.data_as(_P(C.c_int32))) == 0
    rec = dmsun.TrajRecord()
    nb = job.n_bins
    hist = np.zeros(2 * nb)
    cnt = np.zeros(nb, dtype=np.uint32)
    buf = (C.c_uint64 * 14)()
    L.or_opcount_read(buf, 1)
    per, records = [], []
    for i in range(n):
        job.first_trajectory_id = i
        e1 = (dmsun.Event * 1)(ev[i])
        assert L.or_simulate_events(P, C.byref(job), e1, 1, C.byref(rec), hist.ctypes.data_as(_P(C.c_double)),
                                    cnt.ctypes.data_as(_P(C.c_uint32)), 1) == 0
        L.or_opcount_read(buf, 1)
        records.append(bytes(rec))
        per.append(list(buf) + [rec.n_rk_attempts, rec.n_steps_interior, rec.n_steps_exterior, rec.n_tau_rejections,
                                rec.n_scatterings])
    d = np.array(per, dtype=np.float64)
    ops, iters = d[:, :7], d[:, 7:14]
    n_rk, n_int, n_ext, n_rej, n_scat = (d[:, 14 + k] for k in range(5))
    segs = n_scat + 1.0
    tot = lambda x: float(np.sum(x))
    w = {}
    w["rk_attempt"] = tot(ops[:, 0]) / tot(n_rk)
    w["step_interior"] = tot(ops[:, 1]) / tot(iters[:, 1])
    w["step_exterior"] = tot(ops[:, 2]) / tot(iters[:, 2])
    w["step_other"] = tot(ops[:, 3]) / max(1.0, tot(iters[:, 3]))
    w["scatter"] = tot(ops[:, 4]) / max(1.0, tot(n_scat))
    A = np.stack([np.ones_like(segs), segs], axis=1)
    coef, *_ = np.linalg.lstsq(A, ops[:, 5], rcond=None)
    w["trajectory"], w["segment"] = float(coef[0]), float(coef[1])
    resid = ops[:, 5] - A @ coef
    details = dict(point=pdesc, n=n, n_rk=tot(n_rk), n_int=tot(n_int), n_ext=tot(n_ext), n_rej=tot(n_rej),
                   n_scat=tot(n_scat), segments=tot(segs), iterations=[tot(iters[:, k]) for k in range(7)],
                   ops=[tot(ops[:, k]) for k in range(7)], fit_rms=float(np.sqrt(np.mean(resid ** 2))),
                   records=records)
    return w, details


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=2000)
    ap.add_argument("--mode", default="parameter", choices=["parameter", "capture"])
    a = ap.parse_args()
    w, d = count(a.mode, a.n)
    p = d["point"]
    it = d["iterations"]
    print(f"# op-counting oracle build, {a.mode} workload ({p.mass_GeV} GeV, sigma_p {p.sigma_p_cm2} cm^2), "
          f"trajectory ids 0..{a.n - 1}")
    print(f"# events: {d['n_rk']:.0f} RKF45 attempts, {d['n_int']:.0f} interior and {d['n_ext']:.0f} exterior steps, "
          f"{d['n_rej']:.0f} optical-depth retries, {d['n_scat']:.0f} scatterings, {a.n} trajectories")
    print(f"# step iterations by class: interior {it[1]:.0f}, exterior {it[2]:.0f}, other {it[3]:.0f}")
    print("# total operations by part: " + ", ".join(f"{q} {d['ops'][k]:.4g}" for k, q in enumerate(PARTS)))
    print(f"# trajectory fit: rms residual {d['fit_rms']:.1f} ops per trajectory")
    for k, v in w.items():
        print(f"{k} {v:.1f}")


if __name__ == "__main__":
    main()
