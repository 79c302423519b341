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
