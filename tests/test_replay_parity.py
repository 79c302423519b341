"""The kernels' per-lane phase functions (dms_engine.h, dms_halo.h,
dms_physics.h), compiled for the host and driven serially by
tests/harness/engine_replay.cpp, against the oracle: bit-identical tables,
initial conditions, trajectory records and histogram rows.  This pins on the
CPU the code the GPU runs; the GPU tests then show the device evaluates it the
same way."""
import math

import numpy as np
import pytest

import dmsun
import oracle_lib as O
import replay_lib as RL


@pytest.fixture(scope="module")
def replay():
    return RL.Replay()


def rec_bytes(recs):
    return np.ctypeslib.as_array(recs).view(np.uint8).reshape(len(recs), -1)


def test_solar_tables(replay, oracle_solar):
    rng = np.random.default_rng(1)
    rSun = O.const("rSun")
    for which in [0, 1, 2, 3, 4, 100, 105, 140]:
        for r in rng.uniform(0, 1.1 * rSun, 500):
            a, b = oracle_solar.eval(which, r), replay.eval(which, r)
            assert a == b or (math.isnan(a) and math.isnan(b))


@pytest.mark.parametrize("desc_args", [
    dict(mass_GeV=1.0, sigma_p_cm2=1e-35, interpolation_points=60),
    dict(mass_GeV=0.01, sigma_p_cm2=1e-36, interpolation_points=31, speed_points=29),
    dict(mass_GeV=0.1, sigma_p_cm2=0.0, sigma_e_cm2=1e-37, interaction=dmsun.DARK_PHOTON,
         form_factor=dmsun.FF_GENERAL, mediator_mass_MeV=0.01, interpolation_points=40),
])
def test_rate_grid(replay, oracle_solar, desc_args):
    desc = dmsun.make_point(**desc_args)
    assert np.array_equal(O.Point(oracle_solar, desc).grid(), replay.point(desc).grid())


def test_initial_conditions(replay, oracle_solar):
    halo = dmsun.default_halo((11.1, 232.2, 7.3))
    e1, ok1 = O.initial_conditions(oracle_solar, halo, 20260710, 4, 77, 1500)
    e2, ok2 = replay.initial_conditions(halo, 20260710, 4, 77, 1500)
    assert np.array_equal(rec_bytes_ev(e1), rec_bytes_ev(e2)) and np.array_equal(ok1, ok2)


def rec_bytes_ev(ev):
    return np.ctypeslib.as_array(ev).view(np.uint8)


@pytest.mark.parametrize("mode,desc_args,n,max_scat", [
    (dmsun.MODE_CAPTURE, dict(mass_GeV=1.0, sigma_p_cm2=1e-35, interpolation_points=300), 200, 1000),
    (dmsun.MODE_PARAMETER_POINT, dict(mass_GeV=0.3, sigma_p_cm2=1e-36, interpolation_points=300), 120, 10**14),
    (dmsun.MODE_PARAMETER_POINT, dict(mass_GeV=0.1, sigma_p_cm2=0.0, sigma_e_cm2=1e-37,
                                      interaction=dmsun.DARK_PHOTON, interpolation_points=100), 12, 10**14),
    # no rate table: every rate is the direct 64-target sum
    (dmsun.MODE_PARAMETER_POINT, dict(mass_GeV=0.3, sigma_p_cm2=1e-36, interpolation_points=0), 20, 10**14),
    (dmsun.MODE_CAPTURE, dict(mass_GeV=1.0, sigma_p_cm2=1e-35, interpolation_points=1), 40, 1000),
])
def test_trajectories_and_rows(replay, oracle_solar, mode, desc_args, n, max_scat):
    desc = dmsun.make_point(**desc_args)
    job = dmsun.make_job(mode=mode, max_trajectories=n, max_scatterings=max_scat, n_bins=400)
    ev, ok = O.initial_conditions(oracle_solar, job.halo, job.seed, 0, 0, n)
    r1, rows1, _ = O.Point(oracle_solar, desc).simulate_events(job, ev, want_rows=True)
    r2, rows2 = replay.point(desc).simulate_events(job, ev, want_rows=True)
    same = (rec_bytes(r1) == rec_bytes(r2)).all(axis=1)
    assert same.all(), f"{(~same).sum()} of {n} records differ"
    assert np.array_equal(rows1.view(np.uint64), rows2.view(np.uint64))


@pytest.mark.parametrize("mode,desc_args,n,max_scat", [
    (dmsun.MODE_CAPTURE, dict(mass_GeV=1.0, sigma_p_cm2=1e-35, interpolation_points=300), 120, 1000),
    (dmsun.MODE_PARAMETER_POINT, dict(mass_GeV=0.3, sigma_p_cm2=1e-36, interpolation_points=300), 80, 10**14),
    (dmsun.MODE_PARAMETER_POINT, dict(mass_GeV=0.1, sigma_p_cm2=0.0, sigma_e_cm2=1e-37,
                                      interaction=dmsun.DARK_PHOTON, interpolation_points=100), 12, 10**14),
    (dmsun.MODE_PARAMETER_POINT, dict(mass_GeV=0.3, sigma_p_cm2=1e-36, interpolation_points=0), 12, 10**14),
])
@pytest.mark.parametrize("depth", [1, 3])
def test_two_wave_split_is_the_serial_trajectory(replay, oracle_solar, mode, desc_args, n, max_scat, depth):
    """A lone trajectory's steps prepared ahead by a helper (the kernel's
    two-wave split) give the oracle's records and rows bit for bit; the helper
    starts from garbage in every field it is not given, and helper steps taken,
    steps discarded (at optical-depth rejections, scatterings, Kepler returns)
    and redo sessions (a rejected step prepared again by a fresh helper) all
    occur."""
    desc = dmsun.make_point(**desc_args)
    job = dmsun.make_job(mode=mode, max_trajectories=n, max_scatterings=max_scat, n_bins=400)
    ev, ok = O.initial_conditions(oracle_solar, job.halo, job.seed, 0, 0, n)
    r1, rows1, _ = O.Point(oracle_solar, desc).simulate_events(job, ev, want_rows=True)
    r2, rows2, taken, dropped, redos = replay.point(desc).simulate_events_pipe(job, ev, depth, want_rows=True)
    same = (rec_bytes(r1) == rec_bytes(r2)).all(axis=1)
    assert same.all(), f"{(~same).sum()} of {n} records differ"
    assert np.array_equal(rows1.view(np.uint64), rows2.view(np.uint64))
    assert taken > 0 and dropped > 0 and redos > 0


def test_two_wave_split_long_trajectory_both_precisions(replay, oracle_solar):
    """A long bound trajectory of the parameter-point workload (id 65143 of
    m = 0.3 GeV, sigma_p = 1e-36: 1.8e5 RKF45 attempts, a quarter of its steps
    rejected on optical depth, excursions outside the Sun) through the split
    equals its serial replay bit for bit, FP64 (also against the oracle) and
    mixed precision, with each rejected step redone by a fresh helper."""
    desc = dmsun.make_point(0.3, 1e-36, interpolation_points=1000)
    for precision in (dmsun.FP64, dmsun.MIXED):
        job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=1, max_scatterings=10**14, n_bins=1000,
                             first_trajectory_id=65143, precision=precision, halo=dmsun.default_halo((0.0, 248.0, 0.0)))
        ev, ok = O.initial_conditions(oracle_solar, job.halo, job.seed, 0, 65143, 1)
        rp = replay.point(desc)
        serial, rows_s = rp.simulate_events(job, ev, want_rows=True)
        split, rows_p, taken, dropped, redos = rp.simulate_events_pipe(job, ev, 2, want_rows=True)
        assert serial[0].n_rk_attempts > 100000 and serial[0].n_tau_rejections > 10000
        assert np.array_equal(rec_bytes(serial), rec_bytes(split))
        assert np.array_equal(rows_s.view(np.uint64), rows_p.view(np.uint64))
        assert taken > 50000 and dropped > 10000 and redos > 10000
        if precision == dmsun.FP64:
            r1, _, _ = O.Point(oracle_solar, desc).simulate_events(job, ev)
            assert np.array_equal(rec_bytes(r1), rec_bytes(split))


def test_per_lane_histogram_mode_equals_job_sums(replay, oracle_solar):
    """Per-lane scratch rows flushed at each trajectory end (the batched launch's
    histogram path), replayed in id order, equal the oracle job's sums bitwise;
    the scratch rows are left cleared."""
    desc = dmsun.make_point(0.3, 1e-36, interpolation_points=300)
    n = 250
    job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=n, capture_target=0,
                         max_scatterings=10**14, n_bins=300)
    res = O.Point(oracle_solar, desc).run_job(job, threads=4, want_records=False)
    ev, ok = O.initial_conditions(oracle_solar, job.halo, job.seed, job.point_id, 0, n)
    h, c = replay.point(desc).lane_hist(job, ev, ok, nlanes=5)
    assert np.array_equal(h, res["hist"]) and np.array_equal(c, res["counts"])
    assert c.sum() > 0


# Unit divisors as the kernel's headers build them (dms_units.h), same operation order.
UNIT_CM = 1.0 / 1.973269804e-14
UNIT_METER = 100.0 * UNIT_CM
UNIT_KM = 1000.0 * UNIT_METER
UNIT_SEC = 299792458.0 * UNIT_METER
UNIT_DIVISORS = (UNIT_SEC, UNIT_KM, UNIT_KM / UNIT_SEC, 1.0e-9)


def constant_division_inputs(seed=5, n=200000):
    """Dividends for dm_div_const: log-uniform over its stated domain with both
    signs, zeros, mantissas next to powers of two, infinities and NaN."""
    rng = np.random.default_rng(seed)
    a = np.exp2(rng.uniform(-800, 800, n)) * rng.choice([-1.0, 1.0], n)
    a[::101] = 0.0
    edge = np.array([np.nextafter(2.0 ** k, 0.0) for k in range(-300, 300)] + [2.0 ** k for k in range(-300, 300)])
    special = np.array([np.inf, -np.inf, np.nan, 1.0, -1.0])
    return np.concatenate([a, edge, -edge, special])


def div_const_host(a, b, y):
    out = np.empty_like(a)
    RL.lib().rp_div_const(dmsun.dptr(a), a.size, b, y, dmsun.dptr(out))
    return out


def same_bits_or_both_nan(x, ref):
    nan = np.isnan(ref)
    return np.array_equal(np.isnan(x), nan) and np.array_equal(x[~nan].view(np.uint64), ref[~nan].view(np.uint64))


def test_constant_divisor_division_is_ieee():
    """dm_div_const (product by the correctly rounded reciprocal, two FMAs) gives
    the bits of a / b for the unit divisors the per-step bookkeeping divides by
    (s, km, km/s, eV) over |a| in [2^-800, 2^800], zero, infinities and NaN."""
    a = constant_division_inputs()
    for b in UNIT_DIVISORS:
        assert same_bits_or_both_nan(div_const_host(a, b, 1.0 / b), a / b), b


def test_constant_division_check_catches_plausible_mistakes():
    """The comparison above fails for the product a y without its correction
    step and for a reciprocal rounded to single precision (one ulp of error in
    the double reciprocal, by contrast, the correction step absorbs on these
    inputs)."""
    a = constant_division_inputs(seed=6, n=50000)
    uncorrected = [same_bits_or_both_nan(a * (1.0 / b), a / b) for b in UNIT_DIVISORS]
    single = [same_bits_or_both_nan(div_const_host(a, b, float(np.float32(1.0 / b))), a / b) for b in UNIT_DIVISORS[:3]]
    assert not any(uncorrected)
    assert not any(single)
