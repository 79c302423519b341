"""GPU parity: the HIP kernels, called through the C ABI, against the CPU oracle.

Bar: integer, index and classification results identical; FP64 results
bit-identical (the kernels and the oracle evaluate the same IEEE operation
sequence); ensemble histograms, which the device sums in a different but fixed
order, within 1e-12 relative.
"""
import ctypes as C
import math

import numpy as np
import pytest

import dmsun
import oracle_lib

pytestmark = pytest.mark.gpu

HIST_RTOL = 1e-12
# Per-trajectory RKF45-attempt budget (the reference's wall-time guard analogue);
# far above the longest trajectory of these cases, it only bounds a runaway lane.
RK_BUDGET = 20_000_000
O_rSun = oracle_lib.const("rSun")


def rec_bytes(recs):
    if isinstance(recs, list):
        recs = (dmsun.TrajRecord * len(recs))(*recs)
    return np.ctypeslib.as_array(recs).view(np.uint8).reshape(len(recs), -1)


def first_difference(a, b):
    for i in range(len(a)):
        for f, _ in dmsun.TrajRecord._fields_:
            x, y = getattr(a[i], f), getattr(b[i], f)
            if hasattr(x, "_fields_"):
                x = (x.t, tuple(x.r), tuple(x.v))
                y = (y.t, tuple(y.r), tuple(y.v))
            if str(x) != str(y):
                return i, f, x, y
    return None


def test_device_math_matches_host(gpu_ctx):
    rng = np.random.default_rng(7)
    n = 200000
    cases = {
        0: (rng.uniform(0, 1e30, n), None, "sqrt"),
        1: (rng.uniform(-1e3, 1e3, n), rng.uniform(0.1, 1e3, n), "div"),
        2: (rng.uniform(-740, 709, n), None, "exp"),
        3: (np.exp(rng.uniform(-700, 700, n)), None, "log"),
        4: (rng.uniform(-1e4, 1e4, n), None, "sin"),
        5: (rng.uniform(-1e4, 1e4, n), None, "cos"),
        6: (rng.uniform(-8, 8, n), None, "erf"),
        7: (rng.uniform(-1, 1, n), None, "acos"),
    }
    for fn, (x, y, name) in cases.items():
        dev = gpu_ctx.probe_math(fn, x, y)
        host = host_detmath(fn, x, y)
        assert np.array_equal(dev.view(np.uint64), host.view(np.uint64)), name


def test_shared_reciprocal_division_is_ieee(gpu_ctx):
    """dm_div_shared (one refined reciprocal per divisor, three operations per
    quotient) gives the bits of a / b over its stated domain: b in
    [2^-200, 2^200], a zero or |a| in [2^-800, 2^330], plus the divisors and
    dividends of the RKF45 stages (r in [1 km, 1e40], dividends in [1e-180, 1e100])."""
    rng = np.random.default_rng(11)
    n = 400000
    b = np.exp2(rng.uniform(-200, 200, n))
    a = np.exp2(rng.uniform(-800, 330, n)) * rng.choice([-1.0, 1.0], n)
    a[::97] = 0.0
    b2 = np.exp(rng.uniform(np.log(5.07e18), np.log(1e40), n))
    a2 = np.exp(rng.uniform(np.log(1e-180), np.log(1e100), n)) * rng.choice([-1.0, 1.0], n)
    # mantissas next to powers of two and all-ones mantissas
    edge = np.array([np.nextafter(2.0 ** k, 0.0) for k in range(-150, 150)] + [2.0 ** k for k in range(-150, 150)])
    x = np.concatenate([a, a2, np.tile(edge, 8)[: 2 * edge.size], rng.permutation(np.tile(edge, 2))])
    y = np.concatenate([b, b2, np.repeat(edge, 2)[: 2 * edge.size], np.tile(edge, 2)])
    dev = gpu_ctx.probe_math(8, x, y)
    ref = x / y
    assert np.array_equal(dev.view(np.uint64), ref.view(np.uint64))


def test_constant_divisor_division_is_ieee(gpu_ctx):
    """dm_div_const on the device (reciprocal 1/b, product, two FMAs) gives the
    bits of a / b for the unit divisors of the per-step bookkeeping (s, km,
    km/s, eV; dms_units.h operation order) over |a| in [2^-800, 2^800], zero,
    infinities and NaN."""
    cm = 1.0 / 1.973269804e-14
    meter = 100.0 * cm
    km, sec = 1000.0 * meter, 299792458.0 * meter
    rng = np.random.default_rng(12)
    n = 200000
    a = np.exp2(rng.uniform(-800, 800, n)) * rng.choice([-1.0, 1.0], n)
    a[::101] = 0.0
    edge = np.array([np.nextafter(2.0 ** k, 0.0) for k in range(-300, 300)] + [2.0 ** k for k in range(-300, 300)])
    a = np.concatenate([a, edge, -edge, [np.inf, -np.inf, np.nan]])
    for b in (sec, km, km / sec, 1.0e-9):
        dev = gpu_ctx.probe_math(9, a, np.full_like(a, b))
        ref = a / b
        nan = np.isnan(ref)
        assert np.array_equal(np.isnan(dev), nan)
        assert np.array_equal(dev[~nan].view(np.uint64), ref[~nan].view(np.uint64)), b


def host_detmath(fn, x, y):
    """Host evaluation of the same deterministic functions, via the oracle build."""
    import oracle_lib as O
    L = O.lib()
    out = np.empty_like(x)
    yy = y if y is not None else x
    L.or_probe_math(fn, dmsun.dptr(np.ascontiguousarray(x)), dmsun.dptr(np.ascontiguousarray(yy)), dmsun.dptr(out), x.size)
    return out


@pytest.mark.parametrize("desc_args", [
    dict(mass_GeV=1.0, sigma_p_cm2=1e-35, interpolation_points=200),
    dict(mass_GeV=0.01, sigma_p_cm2=1e-36, interpolation_points=31, speed_points=29),
    dict(mass_GeV=0.3, sigma_p_cm2=1e-36, interpolation_points=500, speed_points=180),
    dict(mass_GeV=0.1, sigma_p_cm2=0.0, sigma_e_cm2=1e-37, interaction=dmsun.DARK_PHOTON,
         form_factor=dmsun.FF_GENERAL, mediator_mass_MeV=0.01, interpolation_points=120),
])
def test_rate_grid_bit_identical(gpu_ctx, oracle_solar, desc_args):
    desc = dmsun.make_point(**desc_args)
    pt = gpu_ctx.point(desc)
    g_dev = pt.rate_grid()
    g_cpu = oracle_lib.Point(oracle_solar, desc).grid()
    assert g_dev.shape == g_cpu.shape == (desc.interpolation_points, desc.speed_points or desc.interpolation_points)
    assert pt.grid_shape() == g_dev.shape
    assert np.array_equal(g_dev.view(np.uint64), g_cpu.view(np.uint64))


def test_grid_turned_off_by_a_count_below_two(gpu_ctx, oracle_solar):
    """src/Solar_Model.cpp:397-406 and tests/test_Solar_Model.cpp:218-224:
    (1, 29) and (31, 1) leave no grid, so the total rate is the direct sum."""
    for nr, nv in ((1, 29), (31, 1)):
        pt = gpu_ctx.point(dmsun.make_point(0.01, 1e-36, interpolation_points=nr, speed_points=nv))
        assert pt.grid_shape() == (0, 0)
        r = np.array([0.37 * O_rSun, 0.8 * O_rSun])
        v = np.array([0.21, 1e-3])
        assert np.array_equal(pt.rates(0, r, v), pt.rates(1, r, v))


def test_initial_conditions_bit_identical(gpu_ctx, oracle_solar):
    halo = dmsun.default_halo((11.1, 232.2, 7.3))
    n = 4096
    ev_d, ok_d = gpu_ctx.initial_conditions(halo, 20260710, 3, 1000, n)
    ev_c, ok_c = oracle_lib.initial_conditions(oracle_solar, halo, 20260710, 3, 1000, n)
    a = np.ctypeslib.as_array(ev_d).view(np.float64)
    b = np.ctypeslib.as_array(ev_c).view(np.float64)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert np.array_equal(ok_d, ok_c)


CASES = [
    ("capture_si_1GeV", dmsun.MODE_CAPTURE, dict(mass_GeV=1.0, sigma_p_cm2=1e-35, interpolation_points=1000), 2048, 1000),
    ("point_si_0.3GeV", dmsun.MODE_PARAMETER_POINT, dict(mass_GeV=0.3, sigma_p_cm2=1e-36, interpolation_points=1000), 512,
     10**14),
    ("point_dp_general", dmsun.MODE_PARAMETER_POINT,
     dict(mass_GeV=0.1, sigma_p_cm2=0.0, sigma_e_cm2=1e-37, interaction=dmsun.DARK_PHOTON, form_factor=dmsun.FF_GENERAL,
          mediator_mass_MeV=0.01, interpolation_points=300), 24, 10**14),
    ("point_dp_contact", dmsun.MODE_PARAMETER_POINT,
     dict(mass_GeV=0.1, sigma_p_cm2=0.0, sigma_e_cm2=1e-37, interaction=dmsun.DARK_PHOTON, form_factor=dmsun.FF_CONTACT,
          interpolation_points=300), 24, 10**14),
    ("capture_si_0.5GeV_no_table", dmsun.MODE_CAPTURE, dict(mass_GeV=0.5, sigma_p_cm2=1e-36, interpolation_points=0), 64, 1000),
    ("point_si_0.3GeV_grid_500x180", dmsun.MODE_PARAMETER_POINT,
     dict(mass_GeV=0.3, sigma_p_cm2=1e-36, interpolation_points=500, speed_points=180), 256, 10**14),
]


@pytest.mark.parametrize("name,mode,desc_args,n,max_scat", CASES, ids=[c[0] for c in CASES])
def test_trajectories_bit_identical(gpu_ctx, oracle_solar, name, mode, desc_args, n, max_scat):
    desc = dmsun.make_point(**desc_args)
    job = dmsun.make_job(mode=mode, max_trajectories=n, max_scatterings=max_scat, n_bins=1000,
                         rk_attempt_budget=RK_BUDGET)
    ev, ok = gpu_ctx.initial_conditions(job.halo, job.seed, job.point_id, 0, n)
    print(f"case {name}: n={n}", flush=True)
    recs, rows = gpu_ctx.point(desc).simulate_events(job, ev, want_rows=True)
    orecs, orows, _ = oracle_lib.Point(oracle_solar, desc).simulate_events(job, ev, want_rows=True)
    same = (rec_bytes(recs) == rec_bytes(orecs)).all(axis=1)
    assert same.all(), f"{(~same).sum()} of {n} differ; first: {first_difference(recs, orecs)}"
    assert np.array_equal(rows.view(np.uint64), orows.view(np.uint64))


@pytest.mark.parametrize("mode,desc_args,kw", [
    (dmsun.MODE_CAPTURE, dict(mass_GeV=1.0, sigma_p_cm2=1e-35, interpolation_points=1000), dict(capture_target=300)),
    (dmsun.MODE_PARAMETER_POINT, dict(mass_GeV=0.3, sigma_p_cm2=1e-36, interpolation_points=1000),
     dict(max_trajectories=3000, max_scatterings=10**14, n_bins=1000)),
    # the reference's default histogram (NUM_BINS = 2000, include/Simulation_Trajectory.hpp:27-30)
    (dmsun.MODE_PARAMETER_POINT, dict(mass_GeV=0.3, sigma_p_cm2=1e-36, interpolation_points=1000),
     dict(max_trajectories=2000, max_scatterings=10**14, n_bins=2000, first_trajectory_id=40_000)),
])
def test_job_matches_oracle(gpu_ctx, oracle_solar, mode, desc_args, kw):
    desc = dmsun.make_point(**desc_args)
    job = dmsun.make_job(mode=mode, rk_attempt_budget=RK_BUDGET, **{"max_trajectories": 0, **kw})
    if mode == dmsun.MODE_PARAMETER_POINT:
        job.capture_target = 0
    res = gpu_ctx.point(desc).run_job(job)
    ores = oracle_lib.Point(oracle_solar, desc).run_job(job, record_capacity=200000)
    assert_job_equal(res, ores)


def test_edge_jobs(gpu_ctx, oracle_solar):
    """A single trajectory, and zero cross sections (no scattering: every
    trajectory crosses the Sun once and leaves), equal the oracle's runs; a job
    with neither a trajectory count nor a capture target and a point with a
    non-positive mass are refused with an error before any launch."""
    for desc_args, n in [(dict(mass_GeV=0.3, sigma_p_cm2=1e-36, interpolation_points=200), 1),
                         (dict(mass_GeV=0.3, sigma_p_cm2=0.0, interpolation_points=200), 500)]:
        print(f"edge job {desc_args} n={n}", flush=True)
        desc = dmsun.make_point(**desc_args)
        job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=n, capture_target=0,
                             max_scatterings=10**14, n_bins=1000, rk_attempt_budget=RK_BUDGET)
        res = gpu_ctx.point(desc).run_job(job)
        ores = oracle_lib.Point(oracle_solar, desc).run_job(job, record_capacity=n)
        assert_job_equal(res, ores)
        if desc_args["sigma_p_cm2"] == 0.0:
            assert res["counters"].total_scatterings == 0
    desc = dmsun.make_point(0.3, 1e-36, interpolation_points=200)
    empty = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=0, capture_target=0, n_bins=1000)
    with pytest.raises(RuntimeError):
        gpu_ctx.point(desc).run_job(empty)
    with pytest.raises(RuntimeError):
        gpu_ctx.point(dmsun.make_point(-1.0, 1e-36, interpolation_points=200))


def assert_job_equal(res, ores):
    c, oc = res["counters"], ores["counters"]
    for f, _ in dmsun.Counters._fields_:
        if f in ("computing_time_sec", "simulated_trajectories"):
            continue
        a, b = getattr(c, f), getattr(oc, f)
        if f == "reason_counts":
            a, b = list(a), list(b)
        assert a == b, f
    scale = np.maximum(np.abs(ores["hist"]), 1e-300)
    assert np.all(np.abs(res["hist"] - ores["hist"]) <= HIST_RTOL * scale)
    assert np.array_equal(res["counts"], ores["counts"])
    ev_g = [(e.trajectory_id, e.lifetime_unbinding_sec) for e in res["events"]]
    ev_c = [(e.trajectory_id, e.lifetime_unbinding_sec) for e in ores["events"]]
    assert ev_g == ev_c


@pytest.mark.parametrize("mode", [dmsun.MODE_PARAMETER_POINT, dmsun.MODE_CAPTURE])
def test_batch_matches_oracle_and_single(gpu_ctx, oracle_solar, mode):
    """Jobs streamed through one launch give each job's own results: equal to
    the oracle's Generate_Data run and to the job submitted alone."""
    desc = dmsun.make_point(0.3, 1e-36, interpolation_points=1000)
    sizes = [(1500, 0), (700, 50_000), (2200, 7_000_000)]
    jobs = [dmsun.make_job(mode=mode, max_trajectories=n, capture_target=0, max_scatterings=10**14, n_bins=700,
                           first_trajectory_id=first, rk_attempt_budget=RK_BUDGET) for n, first in sizes]
    print(f"batch mode={mode} sizes={sizes}", flush=True)
    pt = gpu_ctx.point(desc)
    batch = pt.run_batch(jobs, want_records=True)
    opt = oracle_lib.Point(oracle_solar, desc)
    for job, res in zip(jobs, batch):
        ores = opt.run_job(job, record_capacity=job.max_trajectories)
        assert_job_equal(res, ores)
        assert np.array_equal(rec_bytes(res["records"]), rec_bytes(ores["records"]))
        alone = pt.run_job(job, want_records=False)
        assert_job_equal(alone, res)


def test_batch_rejects_mixed_descriptors(gpu_ctx):
    desc = dmsun.make_point(0.3, 1e-36, interpolation_points=200)
    a = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=10, capture_target=0, n_bins=100)
    b = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=10, capture_target=0, n_bins=200)
    with pytest.raises(dmsun.DmsError):
        gpu_ctx.point(desc).run_batch([a, b])
    c = dmsun.make_job(mode=dmsun.MODE_CAPTURE, max_trajectories=0, capture_target=5)
    with pytest.raises(dmsun.DmsError):
        gpu_ctx.point(desc).run_batch([a, c])


def test_parameter_point_workload_full_size(gpu_ctx, oracle_solar):
    """The parameter-point workload at its full size, exactly as the bench's
    single job runs it: SI contact m = 0.3 GeV, sigma_p = 1e-36 cm^2, 10^6
    trajectories in one launch, unlimited scatterings, no RKF45 budget, 1000
    bins, 1000 x 1000 rate grid.  The job's counters, visit counts and
    evaporation events equal the oracle's Generate_Data run of the same
    descriptor (16 threads).  Its histogram sums are sums of up to 10^6
    non-negative per-trajectory terms, added by device atomics in a varying
    order against the oracle's id order: the recursive-summation bound puts
    their relative difference below n u = 10^6 x 2^-53 ~ 1.1e-10 (measured:
    ~2e-12), so that is the tolerance here (1e-12 holds for the smaller jobs
    of the other tests)."""
    desc = dmsun.make_point(0.3, 1e-36, interpolation_points=1000)
    job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=1_000_000, capture_target=0,
                         max_scatterings=10**14, n_bins=1000, record=0, halo=dmsun.default_halo((0.0, 248.0, 0.0)))
    print("parameter-point workload: 1e6 trajectories", flush=True)
    res = gpu_ctx.point(desc).run_job(job, want_records=False)
    print(f"GPU done: {res['counters'].trajectories} trajectories, {res['counters'].captured} captured", flush=True)
    ores = oracle_lib.Point(oracle_solar, desc, threads=16).run_job(job, threads=16, want_records=False,
                                                                    event_capacity=1 << 19)
    assert len(ores["events"]) < 1 << 19
    c, oc = res["counters"], ores["counters"]
    for f, _ in dmsun.Counters._fields_:
        if f in ("computing_time_sec", "simulated_trajectories"):
            continue
        a, b = getattr(c, f), getattr(oc, f)
        assert (list(a) if f == "reason_counts" else a) == (list(b) if f == "reason_counts" else b), f
    assert np.array_equal(res["counts"], ores["counts"])
    assert [(e.trajectory_id, e.lifetime_unbinding_sec) for e in res["events"]] == \
        [(e.trajectory_id, e.lifetime_unbinding_sec) for e in ores["events"]]
    rel = np.abs(res["hist"] - ores["hist"]) / np.maximum(np.abs(ores["hist"]), 1e-300)
    print(f"histogram: largest relative difference {rel.max():.3e}", flush=True)
    assert np.all(rel <= 1.0e6 * 2.0 ** -53)
    assert c.trajectories == 1_000_000 and len(res["events"]) > 1000


def test_capture_workload_full_size(gpu_ctx, oracle_solar):
    """The capture-mode workload at its full size: SI contact m = 1 GeV,
    sigma_p = 1e-35 cm^2, 10^4 trajectories, at most 1000 scatterings, 1000 x
    1000 rate grid.  Every record bit for bit, and the run's counters, equal
    the oracle's Generate_Data run."""
    desc = dmsun.make_point(1.0, 1e-35, interpolation_points=1000)
    n = 10_000
    job = dmsun.make_job(mode=dmsun.MODE_CAPTURE, max_trajectories=n, capture_target=0, max_scatterings=1000,
                         rk_attempt_budget=RK_BUDGET, record=1, halo=dmsun.default_halo((0.0, 248.0, 0.0)))
    print("capture workload: 1e4 trajectories", flush=True)
    res = gpu_ctx.point(desc).run_job(job)
    ores = oracle_lib.Point(oracle_solar, desc).run_job(job, threads=16, record_capacity=n)
    assert_job_equal(res, ores)
    assert len(res["records"]) == len(ores["records"]) == n
    same = (rec_bytes(res["records"]) == rec_bytes(ores["records"])).all(axis=1)
    assert same.all(), f"{(~same).sum()} of {n} records differ"
    assert res["counters"].captured > 100


# Long bound trajectories of the parameter-point workload (m = 0.3 GeV,
# sigma_p = 1e-36, v_obs = 248 km/s): 1.8e5 and 1.9e5 RKF45 attempts, about a
# fifth of their steps rejected on optical depth, excursions outside the Sun.
LONG_IDS = (65143, 16020326)


@pytest.mark.parametrize("precision", [dmsun.FP64, dmsun.MIXED], ids=["fp64", "mixed"])
def test_lone_trajectory_split_matches_serial(gpu_ctx, oracle_solar, precision):
    """A trajectory left alone in its wave has its next steps prepared by a
    finished wave of the block (the two-wave split, on by default); with the
    split turned off (lone_serial) its wave runs every step itself.  Records
    are bit-identical either way, and (FP64) equal the oracle's; a batch whose
    waves turn lone at different times and share the finished waves gives the
    same records as well."""
    desc = dmsun.make_point(0.3, 1e-36, interpolation_points=1000)
    pt = gpu_ctx.point(desc)
    halo = dmsun.default_halo((0.0, 248.0, 0.0))
    runs = {}
    try:
        for serial in (1, 0):
            gpu_ctx.set_lone_serial(serial)
            out = []
            for tid in LONG_IDS:
                job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=1, max_scatterings=10**14,
                                     n_bins=1000, first_trajectory_id=tid, precision=precision, halo=halo,
                                     rk_attempt_budget=RK_BUDGET)
                ev, ok = gpu_ctx.initial_conditions(job.halo, job.seed, job.point_id, tid, 1)
                recs, _ = pt.simulate_events(job, ev)
                print(f"id {tid} lone_serial {serial}: {recs[0].n_rk_attempts} attempts", flush=True)
                out.append(rec_bytes(recs))
            # 200 trajectories from 65100 (one block: four waves that finish at different times)
            job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=200, max_scatterings=10**14,
                                 n_bins=1000, first_trajectory_id=65100, precision=precision, halo=halo,
                                 rk_attempt_budget=RK_BUDGET)
            ev, ok = gpu_ctx.initial_conditions(job.halo, job.seed, job.point_id, 65100, 200)
            recs, rows = pt.simulate_events(job, ev, want_rows=True)
            out.append(rec_bytes(recs))
            out.append(rows.view(np.uint64))
            runs[serial] = out
    finally:
        gpu_ctx.set_lone_serial(0)
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)
    assert recs[43].n_rk_attempts > 100000  # id 65143 is in the batch
    if precision == dmsun.FP64:
        for tid, got in zip(LONG_IDS, runs[0]):
            job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=1, max_scatterings=10**14,
                                 n_bins=1000, first_trajectory_id=tid, halo=halo, rk_attempt_budget=RK_BUDGET)
            ev, ok = oracle_lib.initial_conditions(oracle_solar, job.halo, job.seed, job.point_id, tid, 1)
            orecs, _, _ = oracle_lib.Point(oracle_solar, desc).simulate_events(job, ev)
            assert np.array_equal(got, rec_bytes(orecs))


def test_lone_serial_knob_checked(gpu_ctx):
    kn = gpu_ctx.knobs()
    kn.lone_serial = 2
    with pytest.raises(dmsun.DmsError):
        gpu_ctx.check(dmsun.lib().dms_set_launch_knobs(gpu_ctx.h, C.byref(kn)), "dms_set_launch_knobs")
    assert gpu_ctx.knobs().lone_serial == 0
