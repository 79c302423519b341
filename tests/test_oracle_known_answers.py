"""The CPU oracle pinned against the reference's own known answers, closed
forms and brute force (no GPU).

Each test names the reference test or source line whose answer it checks.
Where the reference holds no answer, the check comes from the mathematics
(closed forms, invariants, independent quadrature or Monte Carlo), chosen so
that a dropped term, a wrong sign or a transposed operand fails it.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate, stats

import dmsun
import oracle_lib as O

L = O.lib
cm, km, sec, gram, Kelvin = (O.const(n) for n in ("cm", "km", "sec", "gram", "Kelvin"))
GeV, MeV, eV, pb = O.const("GeV"), O.const("MeV"), O.const("eV"), O.const("pb")
mProton, mElectron = O.const("mProton"), O.const("mElectron")
G, mSun, rSun, AU = O.const("G_Newton"), O.const("mSun"), O.const("rSun"), O.const("AU")


def ulp_eq(a, b, ulps=4):
    """googletest's DOUBLE_EQ: within 4 units in the last place."""
    if a == b:
        return True
    ia = np.array([a]).view(np.int64)[0]
    ib = np.array([b]).view(np.int64)[0]
    ia = ia if ia >= 0 else -(ia & 0x7FFFFFFFFFFFFFFF)
    ib = ib if ib >= 0 else -(ib & 0x7FFFFFFFFFFFFFFF)
    return abs(int(ia) - int(ib)) <= ulps


def ks_uniform(u):
    """Kolmogorov-Smirnov distance of a sample from U(0,1)
    (tests/test_Physics_Validation.cpp:87-100)."""
    u = np.sort(np.asarray(u))
    n = len(u)
    i = np.arange(n)
    return max(np.max(np.abs(u - i / n)), np.max(np.abs((i + 1) / n - u)))


# ---------------------------------------------------------------- units ----

def test_natural_units_against_independent_constants():
    # hbar*c = 197.3269804 MeV fm (CODATA 2018): 1 cm = 1/(1.973269804e-14 GeV)
    assert math.isclose(1.0 / cm, 1.973269804e-14, rel_tol=1e-15)
    # c = 1: 1 km/s = 1/299792.458
    assert math.isclose(km / sec, 1.0 / 299792.458, rel_tol=1e-15)
    # G = 1/M_Planck^2, M_Planck = 1.220890(14)e19 GeV (CODATA 2018)
    assert abs(G * 1.220890e19 ** 2 - 1.0) < 3e-5
    # 1 K = 8.617333262e-14 GeV (Boltzmann constant, exact since 2019)
    assert math.isclose(Kelvin, 8.617333262e-14, rel_tol=1e-15)
    # 1 amu = 931.49410242 MeV; 1 g = 5.60958860e23 GeV
    assert abs(gram / 5.60958860e23 - 1.0) < 1e-8
    # rSun = 6.957e8 m (IAU nominal); AU = 149597870700 m
    assert math.isclose(rSun * 1.973269804e-16, 6.957e8, rel_tol=1e-12)
    assert math.isclose(AU * 1.973269804e-16, 149597870700.0, rel_tol=1e-12)


# ---------------------------------------------------------------- Philox ----

def philox_py(ctr, key):
    """Philox4x32-10 written from the published algorithm (Salmon et al., SC'11)."""
    M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    c = list(ctr)
    k = list(key)
    for rnd in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k[0]) & 0xFFFFFFFF, p1 & 0xFFFFFFFF,
             ((p0 >> 32) ^ c[3] ^ k[1]) & 0xFFFFFFFF, p0 & 0xFFFFFFFF]
        k = [(k[0] + W0) & 0xFFFFFFFF, (k[1] + W1) & 0xFFFFFFFF]
    return c


def philox_c(ctr, key):
    a = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L().or_philox(a, k, o)
    return list(o)


@pytest.mark.parametrize("ctr,key,expected", [
    # Random123 known-answer vectors for philox4x32-10
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(ctr, key, expected):
    assert philox_c(ctr, key) == list(expected)
    assert philox_py(ctr, key) == list(expected)


def test_philox_matches_python_restatement():
    rng = np.random.default_rng(5)
    for _ in range(200):
        ctr = [int(x) for x in rng.integers(0, 2 ** 32, 4)]
        key = [int(x) for x in rng.integers(0, 2 ** 32, 2)]
        assert philox_c(ctr, key) == philox_py(ctr, key)


# ------------------------------------------------------ deterministic math --

def _probe(fn, x, y=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y if y is not None else x, dtype=np.float64)
    out = np.empty_like(x)
    L().or_probe_math(fn, dmsun.dptr(x), dmsun.dptr(y), dmsun.dptr(out), x.size)
    return out


def _ulp_err(got, exact):
    import mpmath
    err = []
    for g, e in zip(got, exact):
        e = mpmath.mpf(e)
        if e == 0:
            err.append(0.0 if g == 0 else float("inf"))
            continue
        u = mpmath.mpf(2) ** (mpmath.floor(mpmath.log(abs(e), 2)) - 52)
        err.append(float(abs(mpmath.mpf(g) - e) / u))
    return np.array(err)


@pytest.mark.parametrize("fn,name,lo,hi", [
    (2, "exp", -700.0, 700.0), (3, "log", 1e-300, 1e300), (4, "sin", -1e5, 1e5), (5, "cos", -1e5, 1e5),
    (6, "erf", -7.0, 7.0), (7, "acos", -1.0, 1.0),
])
def test_detmath_within_two_ulp(fn, name, lo, hi):
    """The elementary functions the kernels and the oracle share (built from
    IEEE +,-,*,/,sqrt only) against 50-digit mpmath."""
    import mpmath
    mpmath.mp.dps = 50
    rng = np.random.default_rng(fn)
    if name == "log":
        x = 10.0 ** rng.uniform(-300, 300, 3000)
    elif name in ("exp",):
        x = np.concatenate([rng.uniform(lo, hi, 2000), rng.uniform(-2, 2, 1000)])
    elif name in ("sin", "cos"):
        x = np.concatenate([rng.uniform(lo, hi, 1000), rng.uniform(-7, 7, 2000)])
    else:
        x = rng.uniform(lo, hi, 3000)
    f = {"exp": mpmath.exp, "log": mpmath.log, "sin": mpmath.sin, "cos": mpmath.cos, "erf": mpmath.erf,
         "acos": mpmath.acos}[name]
    got = _probe(fn, x)
    exact = [f(mpmath.mpf(float(v))) for v in x]
    assert np.max(_ulp_err(got, exact)) <= 2.0, name


def test_detmath_sqrt_and_division_correctly_rounded():
    rng = np.random.default_rng(3)
    x = 10.0 ** rng.uniform(-300, 300, 20000)
    y = 10.0 ** rng.uniform(-300, 300, 20000)
    assert np.array_equal(_probe(0, x), np.sqrt(x))
    with np.errstate(over="ignore", under="ignore"):
        assert np.array_equal(_probe(1, x, y), x / y)


def test_detmath_special_values():
    assert _probe(2, [0.0])[0] == 1.0 and _probe(3, [1.0])[0] == 0.0
    assert _probe(2, [-800.0])[0] == 0.0 and math.isinf(_probe(2, [800.0])[0])
    assert math.isnan(_probe(3, [-1.0])[0]) and _probe(3, [0.0])[0] == -math.inf
    assert _probe(6, [0.0])[0] == 0.0 and _probe(6, [10.0])[0] == 1.0 and _probe(6, [-10.0])[0] == -1.0
    assert _probe(7, [1.0])[0] == 0.0 and _probe(7, [-1.0])[0] == math.pi
    assert all(math.isnan(v) for v in _probe(2, [math.nan]))


# ------------------------------------------------------------ solar model --

def test_hydrogen_number_density_pins_proton_mass(oracle_solar):
    """tests/test_Solar_Model.cpp:30-41"""
    n_h1 = 0.36209 * (1.505e2 * gram / cm ** 3) / mProton
    for r in (0.0, 0.001 * rSun, 0.0015 * rSun):
        assert ulp_eq(oracle_solar.eval(100, r), n_h1)


def test_mass_temperature_known_answers(oracle_solar):
    """tests/test_Solar_Model.cpp:53-71"""
    S = oracle_solar
    assert S.eval(0, 0.0) == 0.0
    assert ulp_eq(S.eval(0, rSun), mSun)
    assert ulp_eq(S.eval(0, 0.0015 * rSun), 4e-7 * mSun)
    assert ulp_eq(S.eval(1, 0.0), 1.549e7 * Kelvin)
    assert abs(S.eval(1, rSun) - 5800.0 * Kelvin) < 1e-20
    assert ulp_eq(S.eval(1, 0.9695 * rSun), 1.544e5 * Kelvin)


def test_escape_speed_known_answers(oracle_solar):
    """tests/test_Solar_Model.cpp:73-81; the central value pins G, mSun, rSun
    and the quadrature of M(r)/r^2."""
    S = oracle_solar
    assert abs(S.eval(2, 0.0) - 1384.13 * km / sec) < 0.1 * km / sec
    assert ulp_eq(S.eval(2, rSun), math.sqrt(2.0 * G * mSun / rSun))
    assert ulp_eq(S.eval(2, 2 * rSun), math.sqrt(2.0 * G * mSun / 2.0 / rSun))
    assert ulp_eq(S.eval(2, 3 * rSun), math.sqrt(2.0 * G * mSun / (3.0 * rSun)))


def test_escape_speed_against_independent_quadrature(oracle_solar):
    """v_esc(r)^2 = 2 G M/R + 2 G int_r^R M(x)/x^2 dx with scipy's adaptive
    quadrature over the oracle's own M(r)."""
    S = oracle_solar
    for f in (0.0, 0.1, 0.35, 0.7, 0.95):
        r = f * rSun
        val, _ = integrate.quad(lambda x: S.eval(0, x) / x ** 2 if x > 0 else 0.0, r, rSun, limit=400,
                                epsrel=1e-11)
        expect = math.sqrt(2 * G * mSun / rSun + 2 * G * val)
        assert abs(S.eval(2, r) / expect - 1.0) < 2e-6, f


def test_number_densities(oracle_solar):
    """tests/test_Solar_Model.cpp:83-110"""
    S = oracle_solar
    iso = S.isotopes()
    r = 0.5 * rSun
    ne = 0.0
    for i, (Z, A, m, ab) in enumerate(iso):
        ne += Z * S.eval(100 + i, r)
        assert S.eval(100 + i, 1.5 * rSun) == 0.0
    assert ulp_eq(S.eval(4, r), ne)
    assert S.eval(4, 1.5 * rSun) == 0.0


def test_profiles_physical_structure(oracle_solar):
    """tests/test_Physics_Validation.cpp:124-157"""
    S = oracle_solar
    prev_m, prev_v = S.eval(0, 0.0), S.eval(2, 0.0)
    for i in range(1001):
        r = rSun * i / 1000
        m, T, rho, v = S.eval(0, r), S.eval(1, r), S.eval(3, r), S.eval(2, r)
        assert all(math.isfinite(q) for q in (m, T, rho, v))
        assert m >= prev_m and T > 0 and rho >= 0 and v <= prev_v
        prev_m, prev_v = m, v


def test_isotope_table(oracle_solar):
    iso = oracle_solar.isotopes()
    assert iso[0][:2] == (1, 1) and iso[0][2] == mProton
    for Z, A, m, ab in iso:
        assert 1 <= Z <= 28 and A >= Z
        if A > 1:
            assert m == A * O.const("mNucleon")
    # natural abundances of each element (Z >= 10) sum to one
    from collections import defaultdict
    tot = defaultdict(float)
    for Z, A, m, ab in iso:
        if Z >= 10:
            tot[Z] += ab
    assert all(abs(v - 1.0) < 2e-3 for v in tot.values())


# ------------------------------------------------------------ thermal rate --

def test_thermal_average_known_answers():
    """tests/test_Solar_Model.cpp:267-281, tests/test_Physics_Validation.cpp:232-240"""
    f = L().or_thermal_averaged_relative_speed
    assert abs(f(0.01 * Kelvin, mProton, 1e-3) - 1e-3) < 1e-10
    assert abs(f(1e7 * Kelvin, mProton, 1e-3) - 0.0017928) < 1e-7
    T = 1.2e7 * Kelvin
    mean = math.sqrt(8 * T / (math.pi * mProton))
    assert abs(f(T, mProton, 0.0) - mean) <= 1e-12 * mean
    assert math.isnan(f(0.0, mProton, 1e-3))
    assert math.isnan(f(1e7 * Kelvin, mProton, -1e-3))


def test_thermal_average_against_monte_carlo():
    """<|v - v_T|> over a Maxwell-Boltzmann target, brute force."""
    rng = np.random.default_rng(11)
    f = L().or_thermal_averaged_relative_speed
    for T_K, m, v in [(1.5e7, mProton, 3e-4), (5e6, 4 * mProton, 2e-3), (1e6, mElectron, 0.02)]:
        T = T_K * Kelvin
        s = math.sqrt(T / m)
        vt = rng.normal(0.0, s, (400000, 3))
        rel = np.linalg.norm(vt - np.array([0.0, 0.0, v]), axis=1)
        est, err = rel.mean(), rel.std() / math.sqrt(rel.size)
        assert abs(f(T, m, v) - est) < 5 * err


# ------------------------------------------------------------------ rates --

@pytest.fixture(scope="module")
def point_lowmass(oracle_solar):
    """obscura::DM_Particle_SI(0.01 GeV), low mass, sigma_p = 1 pb, 1000^2 grid
    (tests/test_Solar_Model.cpp:168-190)."""
    return O.Point(oracle_solar, dmsun.make_point(0.01, 1e-36, interpolation_points=1000), threads=8)


def test_rate_sums_and_outside_sun(oracle_solar, point_lowmass):
    """tests/test_Solar_Model.cpp:112-166"""
    P = point_lowmass
    n = len(oracle_solar.isotopes())
    r, v = 0.5 * rSun, 1e-3
    total = P.rate(2, r, v)
    for i in range(n):
        total += P.rate(3 + i, r, v)
    assert P.rate(1, r, v) >= 0.0
    assert ulp_eq(P.rate(1, r, v), total)
    assert P.rate(0, 1.5 * rSun, v) == 0.0 and P.rate(1, 1.5 * rSun, v) == 0.0


def test_rate_zero_cross_section(oracle_solar):
    P = O.Point(oracle_solar, dmsun.make_point(0.5, 0.0, interpolation_points=0))
    assert P.rate(1, 0.5 * rSun, 1e-3) == 0.0


def test_rate_grid_within_ten_percent(point_lowmass):
    """tests/test_Solar_Model.cpp:168-190: 500 points, r in [0, R], w in [0, 0.3]."""
    rng = np.random.default_rng(998)
    for _ in range(500):
        r = rng.uniform(0, rSun)
        w = rng.uniform(0, 0.3)
        exact = point_lowmass.rate(1, r, w)
        assert abs(point_lowmass.rate(0, r, w) - exact) <= 0.1 * exact


def test_rate_grid_exact_at_nodes(oracle_solar):
    """tests/test_Solar_Model.cpp:192-224: a 31 x 29 grid (radii x speeds) is
    exact at its nodes; (1, 29) turns the grid off and the direct sum is used."""
    Nr, Nv = 31, 29
    P = O.Point(oracle_solar, dmsun.make_point(0.01, 1e-36, interpolation_points=Nr, speed_points=Nv))
    g = P.grid()
    assert g.shape == (Nr, Nv)
    for i in (0, 7, 15, 30):
        for j in (0, 5, 14, 28):
            r = rSun * i / (Nr - 1)
            v = 0.75 * j / (Nv - 1)
            exact = P.rate(1, r, v)
            assert abs(P.rate(0, r, v) - exact) <= max(1e-30, 1e-10 * abs(exact))
            assert abs(g[i, j] - exact) <= max(1e-30, 1e-10 * abs(exact))
    # bilinear interpolation is exact on a plane: the grid's value between
    # nodes is the weighted mean of the four corners (radius-major, speed fastest)
    r = rSun * (7 + 0.25) / (Nr - 1)
    v = 0.75 * (5 + 0.5) / (Nv - 1)
    want = 0.75 * (0.5 * g[7, 5] + 0.5 * g[7, 6]) + 0.25 * (0.5 * g[8, 5] + 0.5 * g[8, 6])
    assert abs(P.rate(0, r, v) - want) <= 1e-12 * want
    # the same point with the axes' counts exchanged tabulates other nodes
    Pt = O.Point(oracle_solar, dmsun.make_point(0.01, 1e-36, interpolation_points=Nv, speed_points=Nr))
    assert Pt.grid().shape == (Nv, Nr) and not np.allclose(Pt.grid()[:, :Nv], g[:Nv, :], rtol=1e-6, atol=0.0)
    # either count below 2: no grid, the direct sum is used
    for nr, nv in ((1, Nv), (Nr, 1)):
        P1 = O.Point(oracle_solar, dmsun.make_point(0.01, 1e-36, interpolation_points=nr, speed_points=nv))
        assert P1.rate(0, 0.37 * rSun, 0.21) == P1.rate(1, 0.37 * rSun, 0.21)


def test_rate_nan_in_nan_out(oracle_solar):
    """tests/test_Solar_Model.cpp:244-255"""
    P = O.Point(oracle_solar, dmsun.make_point(0.01, 1e-36, interpolation_points=5))
    assert math.isnan(P.rate(0, math.nan, 1e-3))
    assert math.isnan(P.rate(0, 0.5 * rSun, math.nan))


# -------------------------------------------------------- halo and ICs ----

def test_entering_rate_known_answer(oracle_solar):
    """tests/test_Reflection_Spectrum.cpp:60-73 (obscura's default SHM: the
    Sun's velocity (11.1, 232.24, 7.25) km/s)."""
    h = dmsun.default_halo((11.1, 232.24, 7.25))
    r1 = L().or_dm_entering_rate(oracle_solar.h, C.byref(h), 1.0)
    r2 = L().or_dm_entering_rate(oracle_solar.h, C.byref(h), 2.0)
    assert ulp_eq(r1, 2.0 * r2)
    assert abs(r1 - 1.06689e30 / sec) < 1e27 / sec


def test_shm_speed_pdf_normalised_and_moments():
    h = dmsun.default_halo()
    f = lambda v: L().or_shm_pdf_speed(C.byref(h), v)
    vmax = (544.0 + 248.0) * km / sec
    norm, _ = integrate.quad(f, 0, vmax, limit=400, epsabs=0, epsrel=1e-10)
    mean, _ = integrate.quad(lambda v: v * f(v), 0, vmax, limit=400, epsabs=0, epsrel=1e-10)
    inv, _ = integrate.quad(lambda v: f(v) / v if v > 0 else 0.0, 0, vmax, limit=400, epsabs=0, epsrel=1e-10)
    assert abs(norm - 1.0) < 1e-6
    assert abs(L().or_shm_average_speed(C.byref(h)) / mean - 1.0) < 1e-6
    assert abs(L().or_shm_inverse_average_speed(C.byref(h)) / inv - 1.0) < 1e-6


def _ics(oracle_solar, halo, n, seed=3, r_final=0.0):
    ev, ok = O.initial_conditions(oracle_solar, halo, seed, 0, 0, n, r_final_rsun=r_final)
    a = np.ctypeslib.as_array(ev).view(np.float64).reshape(n, 7)
    return ev, ok, a[:, 1:4], a[:, 4:7]


def test_initial_conditions_bounds_and_area_law(oracle_solar):
    """tests/test_Simulation_Utilities.cpp:164-191 and
    tests/test_Physics_Validation.cpp:159-194"""
    n = 8000
    _, _, r, v = _ics(oracle_solar, dmsun.default_halo((11.1, 232.24, 7.25)), n, seed=20260714)
    R = np.linalg.norm(r, axis=1)
    V = np.linalg.norm(v, axis=1)
    J = np.linalg.norm(np.cross(r, v), axis=1)
    vesc_r = np.sqrt(2 * G * mSun / R)
    vesc_sun = oracle_solar.eval(2, rSun)
    Jmax = rSun * np.sqrt(V ** 2 + vesc_sun ** 2)
    assert np.all(R >= 1000 * AU * (1 - 1e-15))
    assert np.all(V >= vesc_r) and np.all(J <= Jmax * (1 + 1e-12)) and np.all(J >= 0)
    area = (J / Jmax) ** 2
    assert abs(area.mean() - 0.5) < 0.01
    assert abs((area ** 2).mean() - 1 / 3) < 0.01
    assert ks_uniform(np.minimum(area, 1.0)) < 1.75 / math.sqrt(n)


def test_initial_speed_distribution(oracle_solar):
    """Speeds at 1000 AU follow u f(u) (u^2 + v_esc(R)^2)/u^2 ... : the entering
    flux weight u * (1 + v_esc(R_sun)^2/u^2) * f_SHM(u), checked by KS against
    its quadrature CDF (u is the speed far from the Sun, u^2 = v^2 - v_esc(r)^2)."""
    n = 6000
    h = dmsun.default_halo()
    _, _, r, v = _ics(oracle_solar, h, n, seed=7)
    R = np.linalg.norm(r, axis=1)
    V = np.linalg.norm(v, axis=1)
    u = np.sqrt(np.maximum(V ** 2 - 2 * G * mSun / R, 0.0))
    vesc = oracle_solar.eval(2, rSun)
    f = lambda x: L().or_shm_pdf_speed(C.byref(h), x)
    w = lambda x: (x + vesc ** 2 / x) * f(x) if x > 0 else 0.0
    grid = np.linspace(0, (544 + 248) * km / sec, 4001)
    pdf = np.array([w(x) for x in grid])
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(grid))])
    cdf /= cdf[-1]
    d = stats.kstest(u, lambda x: np.interp(x, grid, cdf)).statistic
    assert d < 1.75 / math.sqrt(n)


def test_hyperbolic_kepler_shift_conserves_energy_and_momentum(oracle_solar):
    """tests/test_Simulation_Utilities.cpp:203-237"""
    ev, _, _, _ = _ics(oracle_solar, dmsun.default_halo((0.0, 0.0, 0.0)), 100, seed=123)

    def EJ(e):
        r = np.array(e.r)
        v = np.array(e.v)
        return 0.5 * v @ v - G * mSun / np.linalg.norm(r), np.linalg.norm(np.cross(r, v))

    for e in ev:
        E0, J0 = EJ(e)
        for target in (5.0, 1.0):
            assert L().or_hyperbolic_kepler_shift(C.byref(e), target * rSun) == 1
            r = np.array(e.r)
            v = np.array(e.v)
            assert abs(np.linalg.norm(r) - target * rSun) < 1e-10 * rSun
            assert r @ v < 0.0
            E, J = EJ(e)
            assert abs(E / E0 - 1) < 1e-10 and abs(J / J0 - 1) < 1e-10
            assert np.all(np.isfinite(r)) and np.all(np.isfinite(v))


def test_hyperbolic_kepler_shift_invalid_branches():
    """tests/test_Simulation_Utilities.cpp:239-255"""
    def ev(r, v):
        e = dmsun.Event()
        e.t = 0.0
        e.r[:] = r
        e.v[:] = v
        return e

    assert L().or_hyperbolic_kepler_shift(C.byref(ev([0.5 * rSun, 0, 0], [-1000 * km / sec, 0, 0])), 2 * rSun) == 0
    e = ev([5 * rSun, 0, 0], [-1000 * km / sec, 0, 0])
    assert L().or_hyperbolic_kepler_shift(C.byref(e), 6 * rSun) == 0
    assert list(e.r) == [5 * rSun, 0, 0] and list(e.v) == [-1000 * km / sec, 0, 0]
    assert L().or_hyperbolic_kepler_shift(C.byref(ev([2 * rSun, 0, 0], [-1 * km / sec, 0, 0])), rSun) == 0


# -------------------------------------------------------------- propagator --

def _prop(r, v):
    e = dmsun.Event()
    e.t = 0.0
    e.r[:] = r
    e.v[:] = v
    return L().or_prop_create(C.byref(e))


def _state(p):
    t, r, s, dt = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    e = dmsun.Event()
    L().or_prop_state(p, C.byref(t), C.byref(r), C.byref(s), C.byref(dt), C.byref(e))
    return t.value, r.value, s.value, dt.value, e


def _ellipse_error(step, duration):
    """tests/test_Physics_Validation.cpp:28-84"""
    a, ecc = 2.0 * rSun, 0.3
    mu = G * mSun
    n_mean = math.sqrt(mu / a ** 3)
    peri = a * (1 - ecc)
    vp = math.sqrt(mu * (1 + ecc) / (a * (1 - ecc)))
    p = _prop([peri, 0, 0], [0, vp, 0])
    steps = 0
    try:
        while _state(p)[0] < duration * (1 - 1e-14):
            L().or_prop_set_time_step(p, min(step, duration - _state(p)[0]))
            assert L().or_prop_step_mass(p, mSun) == 1
            steps += 1
            assert steps <= 4 * math.ceil(duration / step)
        t, _, _, _, e = _state(p)
    finally:
        L().or_prop_free(p)
    M = n_mean * duration
    E = M
    for _ in range(20):
        E -= (E - ecc * math.sin(E) - M) / (1 - ecc * math.cos(E))
    vf = math.sqrt(1 - ecc ** 2)
    rate = n_mean / (1 - ecc * math.cos(E))
    pos = np.array([a * (math.cos(E) - ecc), a * vf * math.sin(E), 0])
    vel = np.array([-a * math.sin(E) * rate, a * vf * math.cos(E) * rate, 0])
    assert abs(t - duration) <= 1e-12 * duration
    return np.linalg.norm(np.array(e.r) - pos), np.linalg.norm(np.array(e.v) - vel)


def test_rkf45_fourth_order_on_kepler_ellipse():
    """tests/test_Physics_Validation.cpp:106-122"""
    d = 4000.0 * sec
    c, m, f = (_ellipse_error(s * sec, d) for s in (400.0, 200.0, 100.0))
    assert c[0] / m[0] > 8 and m[0] / f[0] > 8 and c[1] / m[1] > 8 and m[1] / f[1] > 8
    assert f[0] / km < 0.01 and f[1] / (km / sec) < 1e-5


def test_rk_step_in_sun_moves_outward_and_slows(oracle_solar):
    """tests/test_Simulation_Trajectory.cpp:301-318"""
    r0 = [0.25 * rSun, 0.5 * rSun, 0.25 * rSun]
    v0 = [km / sec, 1000 * km / sec, km / sec]
    p = _prop(r0, v0)
    dt0 = _state(p)[3]
    L().or_prop_step_solar(p, oracle_solar.h)
    t, r, s, dt, _ = _state(p)
    L().or_prop_free(p)
    assert dt != dt0 and t > 0 and r > np.linalg.norm(r0) and s < np.linalg.norm(v0)


def test_rk_zero_error_growth_bounded():
    """tests/test_Simulation_Trajectory.cpp:320-341"""
    p = _prop([17 * km, 0, 0], [0, km / sec, 0])
    dt0 = _state(p)[3]
    for _ in range(3):
        L().or_prop_step_mass(p, 0.0)
    t, r, s, dt, _ = _state(p)
    L().or_prop_free(p)
    assert math.isfinite(t) and math.isfinite(dt) and t < 100 * dt0 and dt <= 64 * dt0 and r > 0 and s > 0


def test_rk_nan_mass_returns_false_with_finite_state():
    """tests/test_Simulation_Trajectory.cpp:350-372"""
    p = _prop([17 * km, 0, 0], [0, km / sec, 0])
    ok = 1
    for _ in range(5):
        ok = L().or_prop_step_mass(p, math.nan)
    t, r, s, dt, _ = _state(p)
    L().or_prop_free(p)
    assert ok == 0 and all(math.isfinite(q) for q in (t, dt, r, s)) and r >= 0


# ------------------------------------------------------- scattering pieces --

def test_dark_photon_angle_sampler_matches_cdf(oracle_solar):
    """tests/test_Physics_Validation.cpp:196-230: F_DM = 1/(q^2 + m_A'^2)
    ("General", m_A' = 10 keV), v = 1e-3, electron target.  The CDF is derived
    here: with t = (1 - cos)/2 and Q = q_max^2, dP/dt ~ 1/(m^2 + Q t)^2, so
    P(t' <= t) = t (m^2 + Q)/(m^2 + Q t)."""
    desc = dmsun.make_point(0.1, 0.0, sigma_e_cm2=1e-36, interaction=dmsun.DARK_PHOTON,
                            form_factor=dmsun.FF_GENERAL, mediator_mass_MeV=0.01, interpolation_points=0)
    P = O.Point(oracle_solar, desc)
    v, n = 1e-3, 10000
    out = np.empty(n)
    L().or_sample_angles(P.h, 314159, -1, v, n, dmsun.dptr(out))
    assert np.all(out >= -1) and np.all(out <= 1)
    m2 = (0.01 * MeV) ** 2
    mu = 0.1 * mElectron / (0.1 + mElectron)
    Q = 4 * (mu * v) ** 2
    t = (1 - out) / 2
    cdf_cos = 1.0 - t * (m2 + Q) / (m2 + Q * t)
    assert ks_uniform(cdf_cos) < 1.75 / math.sqrt(n)
    # a contact interaction gives an isotropic CM angle
    Pc = O.Point(oracle_solar, dmsun.make_point(0.5, 1e-36, interpolation_points=0))
    L().or_sample_angles(Pc.h, 7, 0, v, n, dmsun.dptr(out))
    assert ks_uniform((out + 1) / 2) < 1.75 / math.sqrt(n)


def test_sample_target_frequencies_follow_rates(oracle_solar):
    """Sample_Target (src/Simulation_Trajectory.cpp:1122-1158): target i is
    drawn with probability Gamma_i / Gamma (multinomial chi^2)."""
    P = O.Point(oracle_solar, dmsun.make_point(0.3, 1e-36, sigma_e_cm2=0.0, interpolation_points=0))
    n_iso = len(oracle_solar.isotopes())
    r, v = 0.3 * rSun, 1e-3
    rates = np.array([P.rate(2, r, v)] + [P.rate(3 + i, r, v) for i in range(n_iso)])
    p = rates / rates.sum()
    n = 200000
    out = np.empty(n, dtype=np.int32)
    L().or_sample_targets(P.h, 99, r, v, n, out.ctypes.data_as(C.POINTER(C.c_int32)))
    assert out.min() >= -1
    counts = np.bincount(out + 1, minlength=n_iso + 1)
    expect = n * p
    big = expect >= 20
    obs = np.append(counts[big], counts[~big].sum())
    exp = np.append(expect[big], expect[~big].sum())
    assert stats.chisquare(obs, exp).pvalue > 1e-4


def test_target_velocity_follows_relative_speed_weighted_maxwellian():
    """Sample_Target_Velocity (Romano & Walsh): v_T ~ |v - v_T| f_MB(v_T);
    its speed and direction moments against 2-D quadrature."""
    T, m, vdm = 1.2e7 * Kelvin, mProton, 6e-4
    kappa = math.sqrt(m / (2 * T))
    n = 100000
    vel = np.array([0.0, 0.0, vdm])
    out = np.empty(3 * n)
    failed = L().or_sample_target_velocities(17, T, m, dmsun.dptr(vel), n, dmsun.dptr(out))
    assert failed == 0
    vt = out.reshape(n, 3)
    speed = np.linalg.norm(vt, axis=1)
    mu = vt[:, 2] / speed

    def w(u, c):
        return math.sqrt(u * u + vdm * vdm - 2 * u * vdm * c) * u * u * math.exp(-(kappa * u) ** 2)

    umax = 8.0 / kappa
    Z = integrate.dblquad(lambda c, u: w(u, c), 0, umax, -1, 1, epsrel=1e-9)[0]
    Eu = integrate.dblquad(lambda c, u: u * w(u, c), 0, umax, -1, 1, epsrel=1e-9)[0] / Z
    Emu = integrate.dblquad(lambda c, u: c * w(u, c), 0, umax, -1, 1, epsrel=1e-9)[0] / Z
    assert abs(speed.mean() - Eu) < 5 * speed.std() / math.sqrt(n)
    assert abs(mu.mean() - Emu) < 5 * mu.std() / math.sqrt(n)
    assert Emu < 0  # targets moving against the DM are favoured


def test_new_dm_velocity_conserves_momentum_and_energy():
    """Elastic kinematics of New_DM_Velocity (src/Simulation_Trajectory.cpp:1228-1247):
    with the recoil fixed by momentum conservation, kinetic energy is conserved,
    |v' - v_cm| = m_T |v - v_T| / (m_T + m_DM), and the CM angle to the axis is
    the sampled one."""
    rng = np.random.default_rng(2)
    for k in range(200):
        mDM = 10 ** rng.uniform(-1, 1)
        mT = [mProton, 4 * mProton, mElectron, 56 * mProton][k % 4]
        v = rng.normal(0, 1e-3, 3)
        vT = rng.normal(0, 3e-4, 3)
        cos_a = rng.uniform(-1, 1)
        out = v.copy()
        assert L().or_new_dm_velocity(k, cos_a, mDM, mT, dmsun.dptr(out), dmsun.dptr(vT)) == 1
        vT_new = vT + (mDM / mT) * (v - out)
        e0 = 0.5 * mDM * v @ v + 0.5 * mT * vT @ vT
        e1 = 0.5 * mDM * out @ out + 0.5 * mT * vT_new @ vT_new
        assert abs(e1 / e0 - 1) < 1e-12
        vcm = (mDM * v + mT * vT) / (mDM + mT)
        d = out - vcm
        # out - v_cm cancels when m_T << m_DM: compare on the scale of |v|
        scale = np.linalg.norm(v) + np.linalg.norm(vT)
        assert abs(np.linalg.norm(d) - mT * np.linalg.norm(v - vT) / (mT + mDM)) < 1e-13 * scale
        assert abs(d @ v / (np.linalg.norm(d) * np.linalg.norm(v)) - cos_a) < 1e-13 * scale / np.linalg.norm(d)


# ------------------------------------------------------------ trajectories --

def _simulate_one(oracle_solar, desc, job, r, v):
    P = O.Point(oracle_solar, desc)
    e = (dmsun.Event * 1)()
    e[0].t = 0.0
    e[0].r[:] = r
    e[0].v[:] = v
    recs, _, _ = P.simulate_events(job, e, threads=1)
    return recs[0]


def test_max_free_steps_termination(oracle_solar):
    """tests/test_Simulation_Trajectory.cpp:216-233"""
    desc = dmsun.make_point(0.5, 0.1e-36, interpolation_points=0)
    job = dmsun.make_job(mode=dmsun.MODE_CAPTURE, max_free_steps=0, max_scatterings=10)
    r = 0.5 * rSun
    rec = _simulate_one(oracle_solar, desc, job, [r, 0, 0], [0, 1.1 * oracle_solar.eval(2, r), 0])
    assert dmsun.REASONS[rec.reason] == "MaxFreeSteps"
    assert not rec.flags & (dmsun.F_TRUNCATED | dmsun.F_SURVIVAL_VALID | dmsun.F_FREE | dmsun.F_REFLECTED)


def test_uncaptured_bound_orbit_is_numerical_failure(oracle_solar):
    """tests/test_Simulation_Trajectory.cpp:235-259"""
    desc = dmsun.make_point(0.5, 1e-40, interpolation_points=0)
    job = dmsun.make_job(mode=dmsun.MODE_CAPTURE, max_free_steps=100, max_scatterings=10)
    r = 1.5 * rSun
    rec = _simulate_one(oracle_solar, desc, job, [r, 0, 0], [0.5 * oracle_solar.eval(2, r), 0, 0])
    assert dmsun.REASONS[rec.reason] == "NumericalFailure"
    assert not rec.flags & (dmsun.F_CAPTURED | dmsun.F_SURVIVAL_VALID)
    assert rec.n_scatterings == 0 and rec.t_termination_sec == 0.0


def test_vanishing_cross_section_gives_only_free_particles(oracle_solar):
    """tests/test_Data_Generation.cpp:384-405: sigma = 1e-100 pb -> free ratio 1."""
    desc = dmsun.make_point(1.0, 1e-136, interpolation_points=0)
    job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=200, capture_target=0, n_bins=50,
                         max_scatterings=10**14)
    res = O.Point(oracle_solar, desc).run_job(job, threads=4)
    c = res["counters"]
    assert c.trajectories == 200 and c.total_scatterings == 0 and c.free_particles == 200
    assert c.captured == 0 and c.completed_outward_escapes == 200


def test_simulate_smoke_reflected_or_scattered(oracle_solar):
    """tests/test_Simulation_Trajectory.cpp:261-281"""
    desc = dmsun.make_point(0.5, 0.1e-36, interpolation_points=0)
    job = dmsun.make_job(mode=dmsun.MODE_CAPTURE, max_scatterings=1000)
    ev, ok = O.initial_conditions(oracle_solar, dmsun.default_halo(), job.seed, 0, 0, 5, r_final_rsun=1.5)
    assert ok.all()
    recs, _, _ = O.Point(oracle_solar, desc).simulate_events(job, ev)
    for rec in recs:
        if rec.flags & (dmsun.F_FREE | dmsun.F_REFLECTED):
            assert np.linalg.norm(np.array(rec.final.r)) >= 2.0 * rSun * (1 - 1e-12)
        else:
            assert rec.n_scatterings > 0
