"""ctypes binding of the dmsun C ABI (include/dms.h).

This module is the Python view of the boundary: struct layouts identical to
include/dms.h and thin wrappers around the exported entry points of
libdmsun.so (built in-tree next to this file).  Loading fails loudly when the
library is missing; there is no fallback implementation.
"""
import ctypes as C
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
REPO = HERE.parent
LIB_PATH = HERE / "libdmsun.so"
DATA_FILE = REPO / "data" / "model_agss09.dat"

# dms_status
DMS_OK = 0
DMS_ERR_CANCELLED = -6

# dms_reason (TrajectoryTerminationReason)
REASONS = ["Unknown", "OutwardEscape", "Scatter", "WallTimeLimit", "MaxFreeSteps", "MaxScatterings",
           "NonFiniteState", "SpeedLimit", "NumericalFailure", "CaptureMode", "EnergyDriftEscape"]
(UNKNOWN, OUTWARD_ESCAPE, SCATTER, WALL_TIME_LIMIT, MAX_FREE_STEPS, MAX_SCATTERINGS, NON_FINITE_STATE,
 SPEED_LIMIT, NUMERICAL_FAILURE, CAPTURE_MODE, ENERGY_DRIFT_ESCAPE) = range(11)

SI, SD, DARK_PHOTON = 0, 1, 2
FF_CONTACT, FF_GENERAL = 0, 1
MODE_CAPTURE, MODE_PARAMETER_POINT = 0, 1
FP64, MIXED = 0, 1

F_CAPTURED = 0x1
F_EVENT_OBSERVED = 0x2
F_BOUNDARY_ESCAPE_OBSERVED = 0x4
F_SURVIVAL_VALID = 0x8
F_NUMERICALLY_INVALID_ESCAPE = 0x10
F_TRUNCATED = 0x20
F_SHIFT_FAILED = 0x40
F_FREE = 0x80
F_REFLECTED = 0x100
F_REFLECTION_SHIFT_FAILED = 0x200


class Event(C.Structure):
    _fields_ = [("t", C.c_double), ("r", C.c_double * 3), ("v", C.c_double * 3)]


class Halo(C.Structure):
    _fields_ = [("rho_GeV_cm3", C.c_double), ("v0_kms", C.c_double), ("vobs_kms", C.c_double * 3),
                ("vesc_kms", C.c_double)]


class PointDesc(C.Structure):
    _fields_ = [("mass_GeV", C.c_double), ("sigma_p_cm2", C.c_double), ("sigma_e_cm2", C.c_double),
                ("interaction", C.c_int32), ("form_factor", C.c_int32), ("mediator_mass_MeV", C.c_double),
                ("fn_over_fp", C.c_double), ("low_mass", C.c_int32), ("interpolation_points", C.c_int32),
                ("speed_points", C.c_int32), ("host_rate_grid", C.POINTER(C.c_double))]


class JobDesc(C.Structure):
    _fields_ = [("mode", C.c_int32), ("precision", C.c_int32), ("max_trajectories", C.c_uint64),
                ("capture_target", C.c_uint64), ("seed", C.c_uint32), ("point_id", C.c_uint32),
                ("first_trajectory_id", C.c_uint64), ("max_scatterings", C.c_uint64),
                ("max_free_steps", C.c_uint64), ("rk_attempt_budget", C.c_uint64), ("r_max_rsun", C.c_double),
                ("n_bins", C.c_int32), ("record_trajectories", C.c_int32), ("halo", Halo)]


class TrajRecord(C.Structure):
    _fields_ = [("initial", Event), ("final", Event), ("trajectory_id", C.c_uint64),
                ("n_scatterings", C.c_uint64), ("reason", C.c_int32), ("flags", C.c_uint32),
                ("t_capture_sec", C.c_double), ("t_last_bound_sec", C.c_double),
                ("t_final_unbinding_scatter_sec", C.c_double), ("t_boundary_escape_sec", C.c_double),
                ("t_termination_sec", C.c_double), ("r_first_negative_km", C.c_double),
                ("E_first_negative_eV", C.c_double), ("dE_first_negative_from_prev_eV", C.c_double),
                ("max_free_energy_drift_eV", C.c_double), ("max_free_energy_drift_rel", C.c_double),
                ("n_rk_attempts", C.c_uint64), ("n_steps_interior", C.c_uint64),
                ("n_steps_exterior", C.c_uint64), ("n_tau_rejections", C.c_uint64),
                ("n_tv_iterations", C.c_uint64)]


class Counters(C.Structure):
    _fields_ = [("trajectories", C.c_uint64), ("captured", C.c_uint64),
                ("completed_outward_escapes", C.c_uint64), ("free_particles", C.c_uint64),
                ("reflected_particles", C.c_uint64), ("complete_evaporation", C.c_uint64),
                ("invalid_survival_captured", C.c_uint64), ("initial_shift_failures", C.c_uint64),
                ("final_reflection_shift_failures", C.c_uint64), ("numerical_failures", C.c_uint64),
                ("computational_truncations", C.c_uint64), ("total_scatterings", C.c_uint64),
                ("sync_rounds", C.c_uint64), ("final_round_trajectories", C.c_uint64),
                ("capture_target_overshoot", C.c_uint64), ("early_stop_reason", C.c_int32),
                ("sync_interval", C.c_int32), ("computing_time_sec", C.c_double),
                ("reason_counts", C.c_uint64 * 11), ("n_rk_attempts", C.c_uint64),
                ("n_steps_interior", C.c_uint64), ("n_steps_exterior", C.c_uint64),
                ("n_tau_rejections", C.c_uint64), ("n_tv_iterations", C.c_uint64),
                ("simulated_trajectories", C.c_uint64), ("max_rk_attempts", C.c_uint64)]


class EvapEvent(C.Structure):
    _fields_ = [("rank", C.c_int32), ("pad", C.c_int32), ("trajectory_id", C.c_uint64),
                ("completion_wall_time_sec", C.c_double), ("lifetime_unbinding_sec", C.c_double)]


class LaunchKnobs(C.Structure):
    _fields_ = [("refill_batch", C.c_uint32), ("lone_serial", C.c_uint32)]


class ServerDesc(C.Structure):
    _fields_ = [("launch", JobDesc), ("slot_ring", C.c_uint32), ("max_requests", C.c_uint32),
                ("max_rounds", C.c_uint32), ("pad", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("slots_published", C.c_uint64), ("slots_cancelled", C.c_uint64),
                ("requests_done", C.c_uint64), ("requests_active", C.c_uint32), ("running", C.c_uint32)]


class JobStatus(C.Structure):
    _fields_ = [("trajectories_done", C.c_uint64), ("captured", C.c_uint64), ("scatterings", C.c_uint64),
                ("done", C.c_int32), ("pad", C.c_int32), ("completed_outward_escapes", C.c_uint64),
                ("numerical_failures", C.c_uint64)]


def record_array(n):
    return (TrajRecord * n)()


def records_to_numpy(recs):
    """Structured numpy view (copy) of a ctypes TrajRecord array."""
    if isinstance(recs, list):
        recs = (TrajRecord * len(recs))(*recs)
    return np.ctypeslib.as_array(recs).copy()


def default_halo(vobs=(0.0, 248.0, 0.0)):
    """SHM v0 = 220, v_esc = 544 km/s, rho = 0.4 GeV/cm^3 and the Sun's speed
    through the halo along +y (only |v_obs| enters the trajectory statistics)."""
    h = Halo()
    h.rho_GeV_cm3 = 0.4
    h.v0_kms = 220.0
    h.vobs_kms[:] = list(vobs)
    h.vesc_kms = 544.0
    return h


def make_point(mass_GeV, sigma_p_cm2, sigma_e_cm2=0.0, interaction=SI, form_factor=FF_CONTACT,
               mediator_mass_MeV=0.0, interpolation_points=1000, fn_over_fp=1.0, speed_points=0):
    """interpolation_points x speed_points rate grid (radius-major); speed_points
    0 is the square grid, either count below 2 no grid (direct rates)."""
    p = PointDesc()
    p.mass_GeV = mass_GeV
    p.sigma_p_cm2 = sigma_p_cm2
    p.sigma_e_cm2 = sigma_e_cm2
    p.interaction = interaction
    p.form_factor = form_factor
    p.mediator_mass_MeV = mediator_mass_MeV
    p.fn_over_fp = fn_over_fp
    p.low_mass = 1
    p.interpolation_points = interpolation_points
    p.speed_points = speed_points
    p.host_rate_grid = None
    return p


def make_job(mode=MODE_CAPTURE, max_trajectories=1000, capture_target=0, seed=20260710, point_id=0,
             max_scatterings=1000, max_free_steps=10**12, rk_attempt_budget=0, n_bins=2000,
             precision=FP64, first_trajectory_id=0, record=1, halo=None):
    j = JobDesc()
    j.mode = mode
    j.precision = precision
    j.max_trajectories = max_trajectories
    j.capture_target = capture_target
    j.seed = seed
    j.point_id = point_id
    j.first_trajectory_id = first_trajectory_id
    j.max_scatterings = max_scatterings
    j.max_free_steps = max_free_steps
    j.rk_attempt_budget = rk_attempt_budget
    j.r_max_rsun = 2.0
    j.n_bins = n_bins
    j.record_trajectories = record
    j.halo = halo if halo is not None else default_halo()
    return j


def load_solar_rows(path=DATA_FILE):
    """Data rows of model_agss09.dat (20 header lines) as an (n, 35) array."""
    rows = []
    with open(path) as f:
        for i, line in enumerate(f):
            if i < 20:
                continue
            parts = line.split()
            if len(parts) == 35:
                rows.append([float(x) for x in parts])
    return np.ascontiguousarray(np.array(rows, dtype=np.float64))


_P = C.POINTER


def _declare(lib):
    vp = C.c_void_p
    sig = {
        "dms_abi_version": (C.c_int, []),
        "dms_create": (C.c_int, [_P(vp), C.c_int, _P(C.c_double), C.c_size_t]),
        "dms_destroy": (None, [vp]),
        "dms_last_error": (C.c_char_p, [vp]),
        "dms_load_solar_file": (C.c_int, [C.c_char_p, _P(C.c_double), C.c_size_t, _P(C.c_size_t)]),
        "dms_set_job_memory_budget": (C.c_int, [vp, C.c_size_t]),
        "dms_solar_eval": (C.c_int, [vp, C.c_int, _P(C.c_double), _P(C.c_double), C.c_size_t]),
        "dms_isotopes": (C.c_int, [vp, _P(C.c_int32), _P(C.c_int32), _P(C.c_double), _P(C.c_double),
                                   C.c_size_t, _P(C.c_size_t)]),
        "dms_constant": (C.c_double, [C.c_char_p]),
        "dms_point_create": (C.c_int, [vp, _P(PointDesc), _P(vp)]),
        "dms_point_destroy": (None, [vp]),
        "dms_point_rate_grid": (C.c_int, [vp, _P(C.c_double), C.c_size_t]),
        "dms_point_cross_sections": (C.c_int, [vp, _P(C.c_double), _P(C.c_double)]),
        "dms_point_grid_shape": (C.c_int, [vp, _P(C.c_int32), _P(C.c_int32)]),
        "dms_point_rates": (C.c_int, [vp, C.c_int, _P(C.c_double), _P(C.c_double), _P(C.c_double),
                                      C.c_size_t]),
        "dms_initial_conditions": (C.c_int, [vp, _P(Halo), C.c_uint32, C.c_uint32, C.c_uint64, C.c_size_t,
                                             C.c_double, _P(Event), _P(C.c_int32)]),
        "dms_simulate_events_noalloc": (C.c_int, [vp, _P(JobDesc), vp, C.c_size_t, vp, vp, vp]),
        "dms_simulate_events": (C.c_int, [vp, _P(JobDesc), _P(Event), C.c_size_t, _P(TrajRecord),
                                          _P(C.c_double)]),
        "dms_job_submit": (C.c_int, [vp, _P(JobDesc), _P(vp)]),
        "dms_job_submit_batch": (C.c_int, [vp, _P(JobDesc), C.c_size_t, _P(vp)]),
        "dms_job_submit_multi": (C.c_int, [_P(vp), _P(JobDesc), C.c_size_t, _P(vp)]),
        "dms_multi_autotune": (C.c_int, [_P(vp), _P(JobDesc), C.c_size_t, C.c_uint64, C.c_uint64, _P(C.c_uint32),
                                         _P(C.c_double)]),
        "dms_server_create": (C.c_int, [vp, _P(ServerDesc), _P(vp)]),
        "dms_server_submit": (C.c_int, [vp, vp, _P(JobDesc), _P(vp)]),
        "dms_server_get_stats": (C.c_int, [vp, _P(ServerStats)]),
        "dms_server_destroy": (None, [vp]),
        "dms_job_cancel": (C.c_int, [vp]),
        "dms_set_launch_knobs": (C.c_int, [vp, _P(LaunchKnobs)]),
        "dms_get_launch_knobs": (C.c_int, [vp, _P(LaunchKnobs)]),
        "dms_autotune_knobs": (C.c_int, [_P(vp), _P(JobDesc), C.c_size_t, C.c_uint64, C.c_uint64, _P(C.c_uint32),
                                         C.c_size_t, C.c_int, _P(C.c_double), _P(LaunchKnobs), C.c_int]),
        "dms_job_poll": (C.c_int, [vp, _P(JobStatus)]),
        "dms_job_wait": (C.c_int, [vp, _P(JobStatus)]),
        "dms_job_read": (C.c_int, [vp, _P(Counters), _P(C.c_double), _P(C.c_uint64), C.c_size_t,
                                   _P(EvapEvent), C.c_size_t, _P(C.c_size_t), _P(TrajRecord), C.c_size_t,
                                   _P(C.c_size_t)]),
        "dms_job_timing": (C.c_int, [vp, _P(C.c_double), _P(C.c_double), _P(C.c_double)]),
        "dms_job_destroy": (None, [vp]),
        "dms_probe_math": (C.c_int, [vp, C.c_int, _P(C.c_double), _P(C.c_double), _P(C.c_double),
                                     C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


EXPORTED = ["dms_abi_version", "dms_create", "dms_destroy", "dms_last_error", "dms_set_job_memory_budget", "dms_load_solar_file",
            "dms_solar_eval", "dms_isotopes", "dms_constant", "dms_point_create", "dms_point_destroy",
            "dms_point_rate_grid", "dms_point_grid_shape", "dms_point_cross_sections", "dms_point_rates", "dms_initial_conditions",
            "dms_simulate_events", "dms_simulate_events_noalloc", "dms_job_submit", "dms_job_submit_batch", "dms_job_submit_multi", "dms_multi_autotune", "dms_server_create", "dms_server_submit", "dms_server_get_stats", "dms_server_destroy", "dms_job_cancel", "dms_set_launch_knobs", "dms_get_launch_knobs",
            "dms_autotune_knobs", "dms_job_poll", "dms_job_wait", "dms_job_read",
            "dms_job_timing", "dms_job_destroy", "dms_probe_math"]

_lib = None


def lib():
    """The loaded libdmsun.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        _lib = _declare(C.CDLL(str(LIB_PATH)))
    return _lib


def dptr(a):
    return a.ctypes.data_as(_P(C.c_double))


class DmsError(RuntimeError):
    pass


class Context:
    """One dms_ctx bound to a GPU (device >= 0) or host-only (device < 0)."""

    def __init__(self, device=0, rows=None):
        L = lib()
        if rows is None:
            rows = load_solar_rows()
        self._rows = np.ascontiguousarray(rows, dtype=np.float64)
        h = C.c_void_p()
        st = L.dms_create(C.byref(h), device, dptr(self._rows), self._rows.shape[0])
        if st != DMS_OK:
            raise DmsError(f"dms_create failed ({st}): {L.dms_last_error(None)}")
        self.h = h

    def check(self, st, what):
        if st != DMS_OK:
            raise DmsError(f"{what} failed ({st}): {lib().dms_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            lib().dms_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solar_eval(self, which, r):
        r = np.ascontiguousarray(np.atleast_1d(r), dtype=np.float64)
        out = np.empty_like(r)
        self.check(lib().dms_solar_eval(self.h, which, dptr(r), dptr(out), r.size), "dms_solar_eval")
        return out

    def point(self, desc):
        return Point(self, desc)

    def knobs(self):
        k = LaunchKnobs()
        self.check(lib().dms_get_launch_knobs(self.h, C.byref(k)), "dms_get_launch_knobs")
        return k

    def set_refill_batch(self, k):
        """dms_set_launch_knobs: free lanes a wave waits for before it refills."""
        kn = self.knobs()
        kn.refill_batch = k
        self.check(lib().dms_set_launch_knobs(self.h, C.byref(kn)), "dms_set_launch_knobs")

    def refill_batch(self):
        return self.knobs().refill_batch

    def set_lone_serial(self, on):
        """dms_set_launch_knobs: 1 runs a trajectory left alone in its wave on
        that wave only; 0 (default) splits its steps with a finished wave."""
        kn = self.knobs()
        kn.lone_serial = 1 if on else 0
        self.check(lib().dms_set_launch_knobs(self.h, C.byref(kn)), "dms_set_launch_knobs")

    def set_job_memory_budget(self, nbytes):
        """Device bytes one job's trajectory buffers may hold (default 32 GiB);
        a job larger than that runs as several launches."""
        self.check(lib().dms_set_job_memory_budget(self.h, nbytes), "dms_set_job_memory_budget")

    def probe_math(self, fn, x, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y if y is not None else x, dtype=np.float64)
        out = np.empty_like(x)
        self.check(lib().dms_probe_math(self.h, fn, dptr(x), dptr(y), dptr(out), x.size), "dms_probe_math")
        return out

    def initial_conditions(self, halo, seed, point_id, first_id, n, r_final_rsun=2.0):
        ev = (Event * n)()
        ok = np.zeros(n, dtype=np.int32)
        self.check(lib().dms_initial_conditions(self.h, C.byref(halo), seed, point_id, first_id, n, r_final_rsun,
                                                ev, ok.ctypes.data_as(_P(C.c_int32))), "dms_initial_conditions")
        return ev, ok


def run_multi(points, jobs, want_records=False):
    """Fixed-count jobs of several points in one trajectory launch
    (dms_job_submit_multi); points[k] runs jobs[k]."""
    L = lib()
    n = len(jobs)
    assert len(points) == n and n > 0
    ctx = points[0].ctx
    descs = (JobDesc * n)(*jobs)
    hp = (C.c_void_p * n)(*[p.h.value for p in points])
    hs = (C.c_void_p * n)()
    ctx.check(L.dms_job_submit_multi(hp, descs, n, hs), "dms_job_submit_multi")
    try:
        return [points[k]._collect(C.c_void_p(hs[k]), jobs[k], want_records) for k in range(n)]
    finally:
        for k in range(n):
            L.dms_job_destroy(C.c_void_p(hs[k]))


def autotune_multi(points, jobs, calibration_trajectories=128, calibration_rk_attempts=5000):
    """Job order for a multi-point launch (dms_multi_autotune): returns
    (order, cost_per_trajectory); run_multi([points[k] for k in order], ...)."""
    L = lib()
    n = len(jobs)
    descs = (JobDesc * n)(*jobs)
    hp = (C.c_void_p * n)(*[p.h.value for p in points])
    order = (C.c_uint32 * n)()
    cost = np.zeros(n, dtype=np.float64)
    points[0].ctx.check(L.dms_multi_autotune(hp, descs, n, calibration_trajectories, calibration_rk_attempts, order, dptr(cost)),
                        "dms_multi_autotune")
    return list(order), cost


def autotune_knobs(points, jobs, refill_candidates=(1, 4, 16), repeats=2, apply=True, calibration_trajectories=128,
                   calibration_rk_attempts=5000):
    """dms_autotune_knobs: the calibration launch once per candidate
    refill_batch (repeated); returns (best refill_batch, seconds per
    candidate) and, with apply, sets it on the context."""
    L = lib()
    n = len(jobs)
    descs = (JobDesc * n)(*jobs)
    hp = (C.c_void_p * n)(*[p.h.value for p in points])
    cand = (C.c_uint32 * len(refill_candidates))(*refill_candidates)
    secs = np.zeros(len(refill_candidates), dtype=np.float64)
    best = LaunchKnobs()
    points[0].ctx.check(L.dms_autotune_knobs(hp, descs, n, calibration_trajectories, calibration_rk_attempts, cand,
                                             len(refill_candidates), repeats, dptr(secs), C.byref(best), 1 if apply else 0),
                        "dms_autotune_knobs")
    return best.refill_batch, secs


def run_multi_tuned(points, jobs, want_records=False, **calibration):
    """run_multi in the autotuned job order; results come back in the
    caller's order."""
    order, _ = autotune_multi(points, jobs, **calibration)
    res = run_multi([points[k] for k in order], [jobs[k] for k in order], want_records)
    out = [None] * len(jobs)
    for pos, k in enumerate(order):
        out[k] = res[pos]
    return out


class Server:
    """dms_server: one resident trajectory launch that requests join as slots
    while it runs.  launch: a JobDesc whose mode, precision, limits, r_max and
    n_bins every request shares."""

    def __init__(self, ctx, launch, slot_ring=0, max_requests=0, max_rounds=0):
        self.ctx = ctx
        d = ServerDesc()
        d.launch = launch
        d.slot_ring = slot_ring
        d.max_requests = max_requests
        d.max_rounds = max_rounds
        h = C.c_void_p()
        ctx.check(lib().dms_server_create(ctx.h, C.byref(d), C.byref(h)), "dms_server_create")
        self.h = h

    def submit(self, point, job):
        """Queue one request; returns a handle for collect()."""
        h = C.c_void_p()
        self.ctx.check(lib().dms_server_submit(self.h, point.h, C.byref(job), C.byref(h)), "dms_server_submit")
        return (h, job, point)

    def collect(self, handle, want_records=False):
        h, job, point = handle
        try:
            return point._collect(h, job, False)
        finally:
            lib().dms_job_destroy(h)

    def poll(self, handle):
        st = JobStatus()
        self.ctx.check(lib().dms_job_poll(handle[0], C.byref(st)), "dms_job_poll")
        return st

    def cancel(self, handle):
        """dms_job_cancel on a request: it stops where it is."""
        self.ctx.check(lib().dms_job_cancel(handle[0]), "dms_job_cancel")

    def wait_status(self, handle):
        st = JobStatus()
        return lib().dms_job_wait(handle[0], C.byref(st))

    def stats(self):
        st = ServerStats()
        self.ctx.check(lib().dms_server_get_stats(self.h, C.byref(st)), "dms_server_get_stats")
        return dict(launches=st.launches, slots_published=st.slots_published, slots_cancelled=st.slots_cancelled,
                    requests_done=st.requests_done, requests_active=st.requests_active, running=st.running)

    def close(self):
        if self.h:
            lib().dms_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Point:
    def __init__(self, ctx, desc):
        self.ctx = ctx
        self.desc = desc
        h = C.c_void_p()
        ctx.check(lib().dms_point_create(ctx.h, C.byref(desc), C.byref(h)), "dms_point_create")
        self.h = h

    def close(self):
        if self.h:
            lib().dms_point_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid_shape(self):
        nr, nv = C.c_int32(), C.c_int32()
        self.ctx.check(lib().dms_point_grid_shape(self.h, C.byref(nr), C.byref(nv)), "dms_point_grid_shape")
        return nr.value, nv.value

    def rate_grid(self):
        nr, nv = self.grid_shape()
        out = np.empty(nr * nv, dtype=np.float64)
        self.ctx.check(lib().dms_point_rate_grid(self.h, dptr(out), out.size), "dms_point_rate_grid")
        return out.reshape(nr, nv)

    def rates(self, which, r, v):
        r = np.ascontiguousarray(r, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.empty_like(r)
        self.ctx.check(lib().dms_point_rates(self.h, which, dptr(r), dptr(v), dptr(out), r.size), "dms_point_rates")
        return out

    def simulate_events_noalloc(self, job, d_events, n, d_records, d_rows=None, stream=None):
        """dms_simulate_events_noalloc on device buffers (integer device
        addresses): nothing is allocated; stream None runs synchronously on the
        context's stream, otherwise the kernels are only enqueued on it."""
        self.ctx.check(lib().dms_simulate_events_noalloc(self.h, C.byref(job), C.c_void_p(d_events), n,
                                                          C.c_void_p(d_records), C.c_void_p(d_rows) if d_rows else None,
                                                          C.c_void_p(stream) if stream else None),
                       "dms_simulate_events_noalloc")

    def simulate_events(self, job, events, n_bins=None, want_rows=False):
        n = len(events)
        recs = record_array(n)
        rows = None
        rp = None
        if want_rows:
            rows = np.zeros((n, 2, job.n_bins), dtype=np.float64)
            rp = dptr(rows)
        self.ctx.check(lib().dms_simulate_events(self.h, C.byref(job), events, n, recs, rp), "dms_simulate_events")
        return recs, rows

    def _collect(self, h, job, want_records):
        L = lib()
        st = JobStatus()
        self.ctx.check(L.dms_job_wait(h, C.byref(st)), "dms_job_wait")
        nb = job.n_bins
        hist = np.zeros((8, nb), dtype=np.float64)
        counts = np.zeros((2, nb), dtype=np.uint64)
        cnt = Counters()
        n_ev = C.c_size_t(0)
        n_rec = C.c_size_t(0)
        L.dms_job_read(h, C.byref(cnt), None, None, 0, None, 0, C.byref(n_ev), None, 0, C.byref(n_rec))
        evs = (EvapEvent * max(1, n_ev.value))()
        recs = record_array(max(1, n_rec.value)) if want_records else None
        self.ctx.check(L.dms_job_read(h, C.byref(cnt), dptr(hist), counts.ctypes.data_as(_P(C.c_uint64)), nb,
                                      evs, n_ev.value, C.byref(n_ev), recs,
                                      n_rec.value if want_records else 0, C.byref(n_rec)), "dms_job_read")
        kt = C.c_double(0)
        kl = C.c_double(0)
        ot = C.c_double(0)
        L.dms_job_timing(h, C.byref(kt), C.byref(kl), C.byref(ot))
        return dict(counters=cnt, hist=hist, counts=counts, events=list(evs)[:n_ev.value],
                    records=(recs if n_rec.value == len(recs) else recs[:n_rec.value]) if want_records else None,
                    kernel_ms=kt.value, launches=kl.value, other_ms=ot.value)

    def run_job(self, job, want_records=True):
        """One Generate_Data job (dms_job_submit + wait + read)."""
        L = lib()
        h = C.c_void_p()
        self.ctx.check(L.dms_job_submit(self.h, C.byref(job), C.byref(h)), "dms_job_submit")
        try:
            return self._collect(h, job, want_records)
        finally:
            L.dms_job_destroy(h)

    def submit(self, job):
        """Start one job (dms_job_submit) without waiting: jobs submitted
        back to back run concurrently on their own streams.  Returns a
        handle for collect()."""
        h = C.c_void_p()
        self.ctx.check(lib().dms_job_submit(self.h, C.byref(job), C.byref(h)), "dms_job_submit")
        return (h, job)

    def cancel(self, handle):
        """dms_job_cancel on a submitted job (its launch ends within a few ms)."""
        self.ctx.check(lib().dms_job_cancel(handle[0]), "dms_job_cancel")

    def wait_status(self, handle):
        """dms_job_wait's return code (DMS_OK, DMS_ERR_CANCELLED, ...) without raising."""
        st = JobStatus()
        return lib().dms_job_wait(handle[0], C.byref(st))

    def collect(self, handle, want_records=False):
        """Wait for a submitted job, read its results and release it."""
        h, job = handle
        try:
            return self._collect(h, job, want_records)
        finally:
            lib().dms_job_destroy(h)

    def run_batch(self, jobs, want_records=False):
        """Fixed-count jobs of this point in one trajectory launch
        (dms_job_submit_batch); returns one result dict per job."""
        L = lib()
        n = len(jobs)
        descs = (JobDesc * n)(*jobs)
        hs = (C.c_void_p * n)()
        self.ctx.check(L.dms_job_submit_batch(self.h, descs, n, hs), "dms_job_submit_batch")
        try:
            return [self._collect(C.c_void_p(hs[k]), jobs[k], want_records) for k in range(n)]
        finally:
            for k in range(n):
                L.dms_job_destroy(C.c_void_p(hs[k]))
