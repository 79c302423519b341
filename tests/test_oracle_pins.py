"""Known-answer, closed-form and brute-force pins of the oracle functions that
the reference's own tests leave open: the SI and dark-photon cross sections,
Bound_Kepler_Return_At_Same_Radius, Solve_Linear_Rate_Optical_Depth_Fraction,
Solar_Interior_Fraction_Interval (surface crossings) and the time-weighted
bincount.  Every check comes with a test that a plausible mistake in the
function it pins (a swapped operand, a dropped factor, a wrong root or bin
width) fails it, so that a misreading of the reference shared by the oracle
and the kernel cannot pass unnoticed.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import dmsun
import oracle_lib as O

L = O.lib
cm, km, sec = O.const("cm"), O.const("km"), O.const("sec")
MeV, GeV = O.const("MeV"), O.const("GeV")
mProton, mElectron, aEM = O.const("mProton"), O.const("mElectron"), O.const("aEM")
G, mSun, rSun = O.const("G_Newton"), O.const("mSun"), O.const("rSun")
MU_KM = G * mSun / km ** 3 * sec ** 2  # G M_sun in km^3 / s^2


def reduced(a, b):
    return a * b / (a + b)


def rel_close(a, b, tol):
    return abs(a - b) <= tol * max(abs(a), abs(b))


# ----------------------------------------------------------- cross sections --
# obscura is not in the tree: the SI nuclear cross section of the low-mass
# contact interaction is the standard sigma_N = sigma_p (mu_N / mu_p)^2
# (Z + f_n/f_p (A - Z))^2; the dark photon's are in-tree (src/Dark_Photon.cpp).

def sigma_si_expected(sigma_p, m_dm, Z, A, m_N, fnfp, mu_p=None):
    mu_p = reduced(m_dm, mProton) if mu_p is None else mu_p
    return sigma_p * (reduced(m_dm, m_N) / mu_p) ** 2 * (Z + fnfp * (A - Z)) ** 2


def check_si(point, solar, m_GeV, sigma_p, fnfp, expected=sigma_si_expected):
    for i, (Z, A, mN, _) in enumerate(solar.isotopes()):
        got = L().or_point_sigma_target(point.h, i, 1e-3)
        if not rel_close(got, expected(sigma_p, m_GeV * GeV, Z, A, mN, fnfp), 1e-13):
            return False
    return True


@pytest.mark.parametrize("m_GeV,fnfp", [(0.3, 1.0), (1.0, 0.0), (5.0, 0.5), (0.1, 1.0)])
def test_si_nuclear_cross_sections(oracle_solar, m_GeV, fnfp):
    """sigma_N = sigma_p (mu_N/mu_p)^2 (Z + f_n/f_p (A-Z))^2 for every target
    isotope, velocity independent (contact, low mass); sigma_e is the input."""
    sp = 1e-36
    desc = dmsun.make_point(m_GeV, sp, sigma_e_cm2=2e-40, interpolation_points=0, fn_over_fp=fnfp)
    P = O.Point(oracle_solar, desc)
    assert check_si(P, oracle_solar, m_GeV, sp, fnfp)
    assert L().or_point_sigma_target(P.h, 3, 0.0) == L().or_point_sigma_target(P.h, 3, 0.3)
    assert L().or_point_sigma_target(P.h, -1, 1e-3) == 2e-40
    s_p, s_e = C.c_double(), C.c_double()
    L().or_point_cross_sections(P.h, C.byref(s_p), C.byref(s_e))
    assert rel_close(s_p.value, sp, 1e-15) and rel_close(s_e.value, 2e-40, 1e-15)


def test_si_cross_section_check_catches_plausible_mistakes(oracle_solar):
    """The check above fails for A and Z exchanged, for the coupling not
    squared, for mu_N / mu_e in place of mu_N / mu_p and for the DM mass in
    place of the nucleus in mu_N."""
    P = O.Point(oracle_solar, dmsun.make_point(0.3, 1e-36, interpolation_points=0, fn_over_fp=0.0))
    mutants = [
        lambda sp, m, Z, A, mN, f: sigma_si_expected(sp, m, A, Z, mN, f),
        lambda sp, m, Z, A, mN, f: sp * (reduced(m, mN) / reduced(m, mProton)) ** 2 * (Z + f * (A - Z)),
        lambda sp, m, Z, A, mN, f: sigma_si_expected(sp, m, Z, A, mN, f, mu_p=reduced(m, mElectron)),
        lambda sp, m, Z, A, mN, f: sp * (m / reduced(m, mProton)) ** 2 * (Z + f * (A - Z)) ** 2,
    ]
    assert check_si(P, oracle_solar, 0.3, 1e-36, 0.0)
    for mut in mutants:
        assert not check_si(P, oracle_solar, 0.3, 1e-36, 0.0, expected=mut)


def dp_point(solar, m_GeV, sigma_e, ff=dmsun.FF_CONTACT, mediator_MeV=0.0):
    return O.Point(solar, dmsun.make_point(m_GeV, 0.0, sigma_e_cm2=sigma_e, interaction=dmsun.DARK_PHOTON,
                                           form_factor=ff, mediator_mass_MeV=mediator_MeV, interpolation_points=0))


def dp_sigmas(P):
    s_p, s_e = C.c_double(), C.c_double()
    L().or_point_cross_sections(P.h, C.byref(s_p), C.byref(s_e))
    return s_p.value, s_e.value


def check_dp_ratio(s_p, s_e, m, ratio=lambda m: (reduced(m, mElectron) / reduced(m, mProton)) ** 2):
    return rel_close(s_e / s_p, ratio(m), 1e-14)


@pytest.mark.parametrize("m_GeV", [0.1, 0.01, 1.0])
def test_dark_photon_sigma_ratio_round_trip_and_epsilon_scaling(oracle_solar, m_GeV):
    """tests/test_Dark_Photon.cpp:42-77 (TestSigmaRatio, TestEpsilon) and
    :94-106 (TestSetSigma): sigma_e / sigma_p = (mu_e / mu_p)^2 at fixed
    epsilon; Set_Sigma_Electron(sigma) then Sigma_Electron() returns sigma;
    doubling epsilon (four times sigma_e) quadruples sigma_p."""
    pb = 1e-36
    s_p, s_e = dp_sigmas(dp_point(oracle_solar, m_GeV, 10 * pb))
    assert rel_close(s_e, 10 * pb, 4e-16)
    assert check_dp_ratio(s_p, s_e, m_GeV * GeV)
    s_p4, s_e4 = dp_sigmas(dp_point(oracle_solar, m_GeV, 40 * pb))
    assert rel_close(s_p4, 4.0 * s_p, 4e-16) and rel_close(s_e4, 40 * pb, 4e-16)


def test_dark_photon_ratio_check_catches_plausible_mistakes(oracle_solar):
    """mu_e and mu_p exchanged, the ratio not squared, or the bare masses in
    place of the reduced ones fail the ratio check."""
    m = 0.1 * GeV
    s_p, s_e = dp_sigmas(dp_point(oracle_solar, 0.1, 1e-37))
    assert check_dp_ratio(s_p, s_e, m)
    for mut in (lambda m: (reduced(m, mProton) / reduced(m, mElectron)) ** 2,
                lambda m: reduced(m, mElectron) / reduced(m, mProton),
                lambda m: (mElectron / mProton) ** 2):
        assert not check_dp_ratio(s_p, s_e, m, ratio=mut)


def dp_nucleus_expected(s_p, m, Z, A, mN, v, md=None, q_ref=aEM * mElectron, weight="Z"):
    w = Z if weight == "Z" else A
    s = s_p * (reduced(m, mN) / reduced(m, mProton)) ** 2 * w * w
    if md is not None:
        q2max = 4.0 * (reduced(m, mN) * v) ** 2
        s *= (q_ref ** 2 + md ** 2) ** 2 / md ** 2 / (md ** 2 + q2max)
    return s


def check_dp_nuclei(P, solar, m, s_p, md=None, weight="Z"):
    for v in (0.0, 1e-3, 4e-3):
        for i, (Z, A, mN, _) in enumerate(solar.isotopes()):
            if not rel_close(L().or_point_sigma_target(P.h, i, v), dp_nucleus_expected(s_p, m, Z, A, mN, v, md, weight=weight),
                             1e-13):
                return False
    return True


@pytest.mark.parametrize("ff,md_MeV", [(dmsun.FF_CONTACT, 0.0), (dmsun.FF_GENERAL, 0.01), (dmsun.FF_GENERAL, 10.0)])
def test_dark_photon_nuclear_and_electron_cross_sections(oracle_solar, ff, md_MeV):
    """src/Dark_Photon.cpp:171-213: nuclei scatter with sigma_p (mu_N/mu_p)^2
    Z^2 (charge, not mass number); the General form factor multiplies by
    (q_ref^2 + m_A'^2)^2 / (m_A'^2 (m_A'^2 + q_max^2)) with q_max = 2 mu v, for
    nuclei and electrons alike (q_ref = alpha m_e)."""
    m = 0.1
    P = dp_point(oracle_solar, m, 1e-37, ff, md_MeV)
    s_p, s_e = dp_sigmas(P)
    md = md_MeV * MeV if ff == dmsun.FF_GENERAL else None
    assert check_dp_nuclei(P, oracle_solar, m * GeV, s_p, md)
    assert not check_dp_nuclei(P, oracle_solar, m * GeV, s_p, md, weight="A")
    for v in (0.0, 1e-3, 4e-3):
        want = s_e
        if md is not None:
            q2 = 4.0 * (reduced(m * GeV, mElectron) * v) ** 2
            want *= ((aEM * mElectron) ** 2 + md ** 2) ** 2 / md ** 2 / (md ** 2 + q2)
        assert rel_close(L().or_point_sigma_target(P.h, -1, v), want, 1e-13)


def test_rates_assemble_density_cross_section_and_thermal_speed(oracle_solar):
    """src/Solar_Model.cpp:263-303: Gamma_i(r, v) = n_i(r) sigma_i(v)
    <|v - v_T|>(T(r), m_i, v) for each target, electrons with m_e and n_e."""
    P = O.Point(oracle_solar, dmsun.make_point(0.1, 0.0, sigma_e_cm2=1e-37, interaction=dmsun.DARK_PHOTON,
                                               form_factor=dmsun.FF_GENERAL, mediator_mass_MeV=0.01,
                                               interpolation_points=0))
    iso = oracle_solar.isotopes()
    for r in (0.0, 0.31 * rSun, 0.87 * rSun):
        T = oracle_solar.eval(1, r)
        for v in (2e-4, 2e-3):
            for i in (0, 1, 7, len(iso) - 1):
                want = oracle_solar.eval(100 + i, r) * L().or_point_sigma_target(P.h, i, v) * cm * cm \
                    * L().or_thermal_averaged_relative_speed(T, iso[i][2], v)
                assert rel_close(P.rate(3 + i, r, v), want, 1e-14)
                wrong = oracle_solar.eval(100 + i, r) * L().or_point_sigma_target(P.h, i, v) * cm * cm \
                    * L().or_thermal_averaged_relative_speed(T, 0.1 * GeV, v)
                assert not rel_close(P.rate(3 + i, r, v), wrong, 1e-6)
            want_e = oracle_solar.eval(4, r) * L().or_point_sigma_target(P.h, -1, v) * cm * cm \
                * L().or_thermal_averaged_relative_speed(T, mElectron, v)
            assert rel_close(P.rate(2, r, v), want_e, 1e-14)


# ----------------------------------------------------- bound Kepler return --

def kepler_return_numeric(r0_km, v0_kms):
    """Point-mass two-body orbit integrated from the outbound state until it
    is back at the same radius (DOP853, rtol 1e-13): time, position, velocity."""
    R0 = np.linalg.norm(r0_km)

    def f(t, y):
        x = y[:3]
        return np.concatenate([y[3:], -MU_KM / np.linalg.norm(x) ** 3 * x])

    def back(t, y):
        return np.linalg.norm(y[:3]) - R0

    back.terminal = True
    back.direction = -1
    sol = solve_ivp(f, (0.0, 1e9), np.concatenate([r0_km, v0_kms]), method="DOP853", rtol=1e-13, atol=1e-10,
                    events=back, first_step=1.0)
    t = sol.t_events[0][0]
    y = sol.y_events[0][0]
    return t, y[:3], y[3:]


def outbound_bound_events():
    """Bound outbound states at 1.2 .. 3 R_sun with eccentricities from near
    circular to near parabolic."""
    out = []
    rng = np.random.default_rng(17)
    for k in range(12):
        R = rng.uniform(1.2, 3.0) * rSun / km
        vesc = math.sqrt(2 * MU_KM / R)
        speed = vesc * rng.uniform(0.72, 0.97)
        ang = rng.uniform(0.05, 1.5)  # angle of v from the radial direction
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        t_ = np.cross(d, rng.normal(size=3))
        t_ /= np.linalg.norm(t_)
        out.append((R * d, speed * (math.cos(ang) * d + math.sin(ang) * t_)))
    return out


def to_event(r_km, v_kms, t_sec=0.0):
    e = dmsun.Event()
    e.t = t_sec * sec
    e.r[:] = [x * km for x in r_km]
    e.v[:] = [x * km / sec for x in v_kms]
    return e


def check_kepler_return(t_oracle, t_num, tol=1e-9):
    return abs(t_oracle - t_num) <= tol * t_num


def test_bound_kepler_return_against_integrated_orbit():
    """Bound_Kepler_Return_At_Same_Radius (src/Simulation_Trajectory.cpp:152-201):
    the return time (2 pi - 2 E + 2 e sin E) / n equals the time of flight of
    the integrated orbit back to the same radius; the returned event keeps the
    position and the speed, flips the radial velocity and keeps the tangential
    one (|J| conserved)."""
    for r0, v0 in outbound_bound_events():
        o = to_event(r0, v0, 100.0)
        i = dmsun.Event()
        assert L().or_bound_kepler_return(C.byref(o), C.byref(i)) == 1
        t_num, r_num, v_num = kepler_return_numeric(r0, v0)
        t_or = (i.t - o.t) / sec
        assert check_kepler_return(t_or, t_num)
        ri = np.array(i.r[:]) / km
        vi = np.array(i.v[:]) / (km / sec)
        assert np.allclose(ri, r0, rtol=0, atol=1e-9 * np.linalg.norm(r0))
        rhat = r0 / np.linalg.norm(r0)
        assert abs(np.linalg.norm(vi) - np.linalg.norm(v0)) < 1e-12 * np.linalg.norm(v0)
        assert abs(vi @ rhat + v0 @ rhat) < 1e-12 * np.linalg.norm(v0)
        # the integrated orbit arrives with the same speed and |v_r|, and |J|
        assert abs(np.linalg.norm(v_num) - np.linalg.norm(vi)) < 1e-7 * np.linalg.norm(v0)
        assert abs(v_num @ (r_num / np.linalg.norm(r_num)) - vi @ rhat) < 1e-6 * np.linalg.norm(v0)
        assert abs(np.linalg.norm(np.cross(ri, vi)) - np.linalg.norm(np.cross(r_num, v_num))) \
            < 1e-7 * np.linalg.norm(np.cross(r0, v0))


def test_bound_kepler_return_rejects_inbound_unbound_and_interior():
    """The function returns false (no jump) for inbound, unbound, radial-free
    or interior states (:157-186)."""
    R = 2.0 * rSun / km
    vesc = math.sqrt(2 * MU_KM / R)
    cases = [([R, 0, 0], [-0.5 * vesc, 0.3 * vesc, 0]),   # inbound
             ([R, 0, 0], [1.1 * vesc, 0.1 * vesc, 0]),    # unbound
             ([0.8 * rSun / km, 0, 0], [0.3 * vesc, 0.3 * vesc, 0]),  # inside the Sun
             ([R, 0, 0], [0.0, 0.0, 0.0])]                # at rest
    for r, v in cases:
        o = to_event(np.array(r), np.array(v))
        i = dmsun.Event()
        assert L().or_bound_kepler_return(C.byref(o), C.byref(i)) == 0


def test_kepler_return_check_catches_plausible_mistakes():
    """Return times with the sign of e sin E flipped, with the full period,
    with E in place of the mean anomaly or with half the arc fail the check."""
    r0, v0 = outbound_bound_events()[3]
    t_num, _, _ = kepler_return_numeric(r0, v0)
    R = np.linalg.norm(r0)
    s = np.linalg.norm(v0)
    vr = v0 @ r0 / R
    eps = 0.5 * s * s - MU_KM / R
    a = -MU_KM / (2 * eps)
    J = np.linalg.norm(np.cross(r0, v0))
    e = math.sqrt(1 + 2 * eps * J * J / MU_KM ** 2)
    E = math.acos((1 - R / a) / e)
    n = math.sqrt(MU_KM / a ** 3)
    right = (2 * math.pi - 2 * E + 2 * e * math.sin(E)) / n
    assert vr > 0 and check_kepler_return(right, t_num, 1e-8)
    for wrong in ((2 * math.pi - 2 * E - 2 * e * math.sin(E)) / n, 2 * math.pi / n, (2 * math.pi - 2 * E) / n,
                  (math.pi - E + e * math.sin(E)) / n):
        assert not check_kepler_return(wrong, t_num, 1e-3)


# ------------------------------------------- linear-rate optical-depth root --

def tau_linear(rs, re, dt, f):
    return dt * (rs * f + 0.5 * (re - rs) * f * f)


def bisect_fraction(rs, re, dt, target):
    lo, hi = 0.0, 1.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if tau_linear(rs, re, dt, mid) < target:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def linear_rate_cases(seed=5, n=4000):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        kind = rng.integers(0, 5)
        rs = 10 ** rng.uniform(-30, 5)
        re = {0: 10 ** rng.uniform(-30, 5), 1: rs, 2: 0.0, 3: rs * (1 + 1e-15), 4: rs * rng.uniform(0, 3)}[int(kind)]
        if kind == 2 and rng.random() < 0.5:
            rs, re = re, rs
        dt = 10 ** rng.uniform(-8, 6)
        total = 0.5 * (rs + re) * dt
        target = total * rng.uniform(1e-9, 1.0 - 1e-9)
        out.append((rs, re, dt, target))
    return out


def check_linear_rate(solve, cases):
    """f in [0, 1] and within the rounding of the reference's formula of the
    root: the quadratic (-b + sqrt(b^2 + 4 a t)) / (2 a), a = (re - rs) / 2,
    b = rs, loses about eps b / |a| to cancellation when the end rates nearly
    agree (the reference computes it that way, :434-446); elsewhere the root
    solves the trapezoid depth to 1e-10 and matches bisection to 1e-9."""
    eps = np.finfo(float).eps
    for rs, re, dt, target in cases:
        try:
            f = solve(rs, re, dt, target)
        except (ValueError, ZeroDivisionError):
            return False
        if not (0.0 <= f <= 1.0):
            return False
        a = 0.5 * (re - rs)
        slack = 8.0 * eps * rs / abs(a) if a != 0.0 else 0.0
        if abs(f - bisect_fraction(rs, re, dt, target)) > 1e-9 + slack:
            return False
        if slack < 1e-12 and abs(tau_linear(rs, re, dt, f) - target) > 1e-10 * target + 1e-300:
            return False
    return True


def test_linear_rate_root_solves_trapezoid_depth():
    """Solve_Linear_Rate_Optical_Depth_Fraction (src/Simulation_Trajectory.cpp:416-464):
    for a rate linear from rs to re over dt the returned fraction f in [0, 1]
    has dt (rs f + (re - rs) f^2 / 2) equal to the target depth, and agrees with
    a 200-step bisection -- rising, falling, constant (a = 0 branch) and
    vanishing rates; nearly equal end rates within the cancellation of the
    reference's quadratic formula."""
    assert check_linear_rate(L().or_solve_linear_rate, linear_rate_cases())


def test_linear_rate_root_edge_cases():
    f = L().or_solve_linear_rate
    assert f(1.0, 2.0, 1.0, 0.0) == 0.0 and f(1.0, 2.0, 1.0, -1.0) == 0.0  # nothing to cover
    assert f(1.0, 2.0, 1.0, 1.5) == 1.0 and f(1.0, 2.0, 1.0, 7.0) == 1.0  # target at or beyond the step
    assert f(1.0, 2.0, 0.0, 0.5) == 1.0 and f(-1.0, 2.0, 1.0, 0.1) == 1.0  # no duration, negative rate
    assert f(math.nan, 2.0, 1.0, 0.1) == 1.0 and f(1.0, 2.0, 1.0, math.inf) == 1.0
    assert f(0.0, 0.0, 1.0, 0.1) == 1.0  # no depth at all
    assert f(3.0, 3.0, 2.0, 3.0) == 0.5  # constant rate: target / (rs dt)


def test_linear_rate_check_catches_plausible_mistakes():
    """A solver without the 1/2 of the trapezoid, one taking the other root
    first, or a linear interpolation of the depth fails the check."""
    cases = linear_rate_cases(seed=6, n=400)

    def no_half(rs, re, dt, t):
        a, b = re - rs, rs
        if abs(a) < 1e-300:
            return min(1.0, t / dt / b)
        return min(1.0, max(0.0, (-b + math.sqrt(b * b + 4 * a * t / dt)) / (2 * a)))

    def other_root(rs, re, dt, t):
        a, b = 0.5 * (re - rs), rs
        if abs(a) < 1e-300:
            return min(1.0, t / dt / b)
        disc = b * b + 4 * a * t / dt
        return min(1.0, max(0.0, (-b - math.sqrt(max(0.0, disc))) / (2 * a)))

    def linear_depth(rs, re, dt, t):
        return min(1.0, t / (0.5 * (rs + re) * dt))

    assert check_linear_rate(L().or_solve_linear_rate, cases)
    for bad in (no_half, other_root, linear_depth):
        assert not check_linear_rate(bad, cases)


# --------------------------------------------------- surface crossings ------

def interior_brute(b, a, n=20001):
    """Sub-interval of the chord b + f (a - b), f in [0, 1], inside R_sun:
    dense sampling, then 100 bisections on each inside/outside change."""
    R = rSun

    def inside(f):
        return np.linalg.norm(b + f * (a - b)) < R

    fs = np.linspace(0.0, 1.0, n)
    ins = np.linalg.norm(b[None, :] + fs[:, None] * (a - b)[None, :], axis=1) < R
    if not ins.any():
        return None

    def refine(lo, hi):
        lo_in = inside(lo)
        for _ in range(100):
            mid = 0.5 * (lo + hi)
            if inside(mid) == lo_in:
                lo = mid
            else:
                hi = mid
        return 0.5 * (lo + hi)

    k = np.nonzero(ins)[0]
    s = 0.0 if k[0] == 0 else refine(fs[k[0] - 1], fs[k[0]])
    e = 1.0 if k[-1] == n - 1 else refine(fs[k[-1]], fs[k[-1] + 1])
    return s, e


def chord_cases(seed=9, n=300):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        kind = rng.integers(0, 4)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if kind == 0:    # both ends inside
            b = d * rng.uniform(0, 0.99) * rSun
            a = b + rng.normal(size=3) * 0.1 * rSun
            a *= min(1.0, 0.99 * rSun / np.linalg.norm(a))
        elif kind == 1:  # entering
            b = d * rng.uniform(1.01, 1.5) * rSun
            a = -d * rng.uniform(0.0, 0.9) * rSun + rng.normal(size=3) * 0.05 * rSun
        elif kind == 2:  # leaving
            a = d * rng.uniform(1.01, 1.5) * rSun
            b = -d * rng.uniform(0.0, 0.9) * rSun + rng.normal(size=3) * 0.05 * rSun
        else:            # both ends outside: through the Sun or past it
            b = d * rng.uniform(1.05, 2.0) * rSun
            off = np.cross(d, rng.normal(size=3))
            off /= np.linalg.norm(off)
            if rng.random() < 0.5:
                a = -d * rng.uniform(1.05, 2.0) * rSun + off * rng.uniform(0.0, 2.0) * rSun
            else:
                a = b - d * rng.uniform(0.0, 1.0) * rSun + off * rng.uniform(0.5, 3.0) * rSun
        out.append((b, a))
    return out


def oracle_interval(b, a):
    eb, ea = to_event(b / km, [0, 0, 0]), to_event(a / km, [0, 0, 0])
    s, e = C.c_double(), C.c_double()
    ok = L().or_interior_interval(C.byref(eb), C.byref(ea), float(np.linalg.norm(b)), float(np.linalg.norm(a)),
                                  C.byref(s), C.byref(e))
    return (s.value, e.value) if ok else None


def check_interval(got, want, tol=1e-8):
    if want is None or got is None:
        return got is None and want is None
    return abs(got[0] - want[0]) <= tol and abs(got[1] - want[1]) <= tol


def test_interior_interval_against_dense_sampling():
    """Solar_Interior_Fraction_Interval with Surface_Crossing_Fractions
    (src/Simulation_Trajectory.cpp:231-325) against the brute-force interior
    of the chord: both ends inside, entering, leaving, through and missing."""
    n_through = n_miss = 0
    for b, a in chord_cases():
        want = interior_brute(b, a)
        got = oracle_interval(b, a)
        assert check_interval(got, want), (b, a, got, want)
        if np.linalg.norm(b) > rSun and np.linalg.norm(a) > rSun:
            n_through += want is not None
            n_miss += want is None
    assert n_through > 10 and n_miss > 10


def test_interior_interval_check_catches_plausible_mistakes():
    """Taking the far crossing when entering, the near one when leaving, or
    the whole step when the chord passes through fails the check."""
    cases = chord_cases(seed=10, n=120)

    def swapped(b, a):
        got = oracle_interval(b, a)
        if got is None:
            return None
        rb, ra = np.linalg.norm(b), np.linalg.norm(a)
        d = a - b
        qa, qb, qc = d @ d, 2 * b @ d, b @ b - rSun * rSun
        sd = math.sqrt(max(0.0, qb * qb - 4 * qa * qc))
        first, second = (-qb - sd) / (2 * qa), (-qb + sd) / (2 * qa)
        if rb >= rSun > ra:
            return (min(1.0, max(0.0, second)), 1.0)
        if rb < rSun <= ra:
            return (0.0, min(1.0, max(0.0, first)))
        return got

    def whole(b, a):
        got = oracle_interval(b, a)
        return (0.0, 1.0) if got is not None else None

    assert all(check_interval(oracle_interval(b, a), interior_brute(b, a)) for b, a in cases)
    for bad in (swapped, whole):
        assert not all(check_interval(bad(b, a), interior_brute(b, a)) for b, a in cases)


# --------------------------------------------------------------- bincount --

def sigma0_pass(oracle_solar, periapsis_rsun, nbins, speed_kms=700.0):
    """One trajectory with a vanishing cross section, injected at 2 R_sun with
    the given speed and a Keplerian periapsis: the oracle's histogram row and
    record, and the orbit integrated independently (oracle mass profile M(r)
    inside the Sun, point mass outside)."""
    R = 2.0 * rSun
    v = speed_kms * km / sec
    rp = periapsis_rsun * rSun
    mu = G * mSun
    vp = math.sqrt(v * v - 2 * mu / R + 2 * mu / rp)
    vt = rp * vp / R
    vr = -math.sqrt(v * v - vt * vt)
    P = O.Point(oracle_solar, dmsun.make_point(0.3, 1e-100, interpolation_points=0))
    job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=1, capture_target=0, n_bins=nbins,
                         max_scatterings=10 ** 14)
    e = (dmsun.Event * 1)()
    e[0].t = 0.0
    e[0].r[:] = [R, 0, 0]
    e[0].v[:] = [vr, vt, 0]
    recs, rows, counts = P.simulate_events(job, e, want_rows=True, threads=1)
    Rkm = R / km

    def mass(r_km):
        return oracle_solar.eval(0, r_km * km) / mSun if r_km * km < rSun else 1.0

    def f(t, y):
        x = y[:3]
        r = np.linalg.norm(x)
        return np.concatenate([y[3:], -MU_KM * mass(r) / r ** 3 * x])

    def out(t, y):
        return np.linalg.norm(y[:3]) - Rkm

    out.terminal = True
    out.direction = 1
    y0 = np.array([Rkm, 0.0, 0.0, vr / (km / sec), vt / (km / sec), 0.0])
    sol = solve_ivp(f, (0.0, 1e6), y0, method="DOP853", rtol=1e-11, atol=1e-8, events=out, dense_output=True)
    T = sol.t_events[0][0]
    ts = np.linspace(0.0, T, 400001)
    ys = sol.sol(ts)
    r_t = np.linalg.norm(ys[:3], axis=0)
    v2_t = np.sum(ys[3:] ** 2, axis=0)
    return recs[0], rows[0, 0], rows[0, 1], counts[0], dict(T=T, ts=ts, r=r_t, v2=v2_t)


def max_step_sec():
    """Largest accepted free-flight step inside 2 R_sun: the exterior cap
    0.1 sqrt(r^3 / G M_sun) at 2 R_sun (src/Simulation_Trajectory.cpp:65-81)."""
    R = 2.0 * rSun / km
    return 0.1 * math.sqrt(R ** 3 / MU_KM)


def check_bincount(dt_row, v2dt_row, nbins, orbit, step):
    """Time-weighted bincount of one pass against the integrated orbit: total
    time within 2 R_sun and mean v^2 within one step's worth, cumulative time
    below every bin edge within two steps (a step straddles each radius at
    most once inbound and once outbound), nothing below the periapsis."""
    w = 2.0 * rSun / km / nbins
    T, ts, r, v2 = orbit["T"], orbit["ts"], orbit["r"], orbit["v2"]
    if abs(dt_row.sum() - T) > step:
        return False
    mean_v2 = np.trapezoid(v2, ts) / T
    if abs(v2dt_row.sum() / dt_row.sum() - mean_v2) > 0.1 * mean_v2:
        return False
    dt_sample = ts[1] - ts[0]
    edges = np.arange(1, nbins + 1) * w
    c_hist = np.cumsum(dt_row)
    c_true = np.array([np.sum(r < x) * dt_sample for x in edges])
    if np.max(np.abs(c_hist - c_true)) > 2 * step:
        return False
    lowest = np.nonzero(dt_row)[0]
    return lowest.size > 0 and (lowest[0] + 1) * w > r.min()


@pytest.mark.parametrize("periapsis,nbins", [(0.3, 2000), (0.6, 1000), (1.3, 200), (0.05, 50)])
def test_bincount_of_a_free_pass_matches_the_orbit(oracle_solar, periapsis, nbins):
    """Accumulate_Bincount_Step (src/Simulation_Trajectory.cpp:640-669): for a
    pass with a vanishing cross section the time per radial shell (left-point
    rule, bins of 2 R_sun / n_bins) and the v^2-weighted time agree with the
    independently integrated orbit to within its step size; every committed
    step that starts inside 2 R_sun is counted once."""
    rec, dt_row, v2_row, cnt, orbit = sigma0_pass(oracle_solar, periapsis, nbins)
    assert dmsun.REASONS[rec.reason] == "OutwardEscape" and rec.n_scatterings == 0
    assert check_bincount(dt_row, v2_row, nbins, orbit, max_step_sec())
    assert np.all((cnt > 0) == (dt_row > 0))
    assert cnt.sum() <= rec.n_steps_interior + rec.n_steps_exterior
    # the pass spends all but its first step (which starts at 2 R_sun exactly,
    # outside the binned range) and the part of its last beyond 2 R_sun
    assert dt_row.sum() <= rec.final.t / sec


def test_bincount_check_catches_plausible_mistakes(oracle_solar):
    """Histograms with bins of R_sun / n_bins, with v in place of v^2, with dt
    in milliseconds, or with each step's time put in the bin of its end point
    shifted outward by a bin, fail the check."""
    nbins = 1000
    rec, dt_row, v2_row, cnt, orbit = sigma0_pass(oracle_solar, 0.3, nbins)
    step = max_step_sec()
    assert check_bincount(dt_row, v2_row, nbins, orbit, step)
    half_width = np.zeros(nbins)
    half_width[: nbins // 2] = dt_row[0::2] + dt_row[1::2]  # bins twice as wide as the row claims
    assert not check_bincount(half_width, v2_row, nbins, orbit, step)
    v_not_v2 = np.where(dt_row > 0, np.sqrt(np.maximum(v2_row / np.maximum(dt_row, 1e-300), 0)) * dt_row, 0)
    assert not check_bincount(dt_row, v_not_v2, nbins, orbit, step)
    assert not check_bincount(1e3 * dt_row, 1e3 * v2_row, nbins, orbit, step)
    shifted = np.roll(dt_row, 40)
    assert not check_bincount(shifted, np.roll(v2_row, 40), nbins, orbit, step)


def test_generate_data_round_longer_than_a_chunk(oracle_solar):
    """A sync round longer than the oracle driver's 4096-trajectory chunks
    (sigma_p = 1e-38 cm^2: the reference's interval is then far above 4096,
    src/Data_Generation.cpp:59-74) is simulated in several chunks, in id order:
    every trajectory once, the counters the records' sums."""
    P = O.Point(oracle_solar, dmsun.make_point(0.1, 1e-38, interpolation_points=100))
    n = 10000
    job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=n, capture_target=0, max_scatterings=2000,
                         n_bins=100, record=1, halo=dmsun.default_halo((0.0, 248.0, 0.0)))
    res = P.run_job(job, threads=4, record_capacity=n)
    c, recs = res["counters"], res["records"]
    assert c.sync_interval > 4096 and c.trajectories == n and len(recs) == n
    assert [r.trajectory_id for r in recs] == list(range(n))
    assert c.captured == sum(1 for r in recs if r.flags & dmsun.F_CAPTURED)
    assert c.total_scatterings == sum(r.n_scatterings for r in recs)
