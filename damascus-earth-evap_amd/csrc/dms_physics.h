/*
 * Solar model, cross sections, scattering rates and orbit helpers of the
 * trajectory path as __host__ __device__ functions (FP64).
 *
 * Layouts are chosen for gfx950: the 63 nuclear number-density splines and the
 * electron-density spline are interleaved per profile interval
 * (dens[j * 64 + k], 32 B per entry), so the 64 lanes of a wave that
 * cooperate on one scattering read one contiguous 2 KiB line.
 *
 * Operation order follows the reference sources cited per function so that
 * results are bit-identical to the CPU restatement under -ffp-contract=off.
 */
#ifndef DMS_PHYSICS_H
#define DMS_PHYSICS_H

#include <hip/hip_runtime.h>

#include "detmath.h"
#include "dms_units.h"

#define DMS_HD __host__ __device__ inline

struct __attribute__((aligned(32))) DmsCubic
{
	double a, b, c, d;
};

struct DmsSolarView
{
	int n;        /* padded profile nodes */
	int n_iso;    /* nuclear targets (<= 63) */
	double x1;    /* second node (first interior grid point) */
	double inv_h; /* 1 / uniform interior spacing, for the O(1) interval guess */
	const double* x;
	const DmsCubic* mass;
	const DmsCubic* temp;
	const DmsCubic* vesc2;
	const DmsCubic* dens; /* [(n-1) * 64] */
	const double* iso_pref; /* abundance prefactor per isotope */
	const double* iso_mass;
	const int* iso_Z;
};

struct DmsPointView
{
	double mass;
	double mu_e;
	int interaction; /* dms_interaction */
	int ff;          /* dms_form_factor */
	double m_dark_photon, q_reference;
	double sigma_e;
	const double* sigma_N; /* [64] velocity-independent part */
	int use_grid, Nr, Nv;
	double inv_dr, inv_dv, vmax;
	const double* grid;
	const float* grid32; /* the same table in FP32 and 1/s (mixed-precision kernel) */
};

/* Interval j with x_j <= x < x_{j+1}, clamped to [0, n-2]: the same index a
 * bisection returns, found from the uniform interior spacing plus exact
 * neighbour checks. */
DMS_HD int dms_interval(const DmsSolarView& S, double x)
{
	const int n = S.n;
	const double t = (x - S.x1) * S.inv_h;
	int j;
	if(!(t >= 0.0))
		j = 0;
	else if(t >= (double) n)
		j = n - 2;
	else
		j = (int) t + 1;
	if(j > n - 2)
		j = n - 2;
	while(j < n - 2 && S.x[j + 1] <= x)
		j++;
	while(j > 0 && S.x[j] > x)
		j--;
	return j;
}

DMS_HD double dms_cubic(const DmsCubic& c, double t) { return ((c.a * t + c.b) * t + c.c) * t + c.d; }

/* Cubic piece j of table T evaluated at x, for x in [x_0, x_{n-1}] (finite).
 * The interval guess from the uniform spacing is almost always exact, so the
 * node pair and the coefficients of the guessed interval are read together
 * (one memory round trip on the dependent chain); a wrong guess falls back to
 * dms_interval.  The result is that of dms_cubic(T[j], x - x_j) with j from
 * dms_interval in every case. */
DMS_HD double dms_spline_at(const DmsSolarView& S, const DmsCubic* T, double x)
{
	const int n = S.n;
	const double t = (x - S.x1) * S.inv_h;
	int j;
	if(!(t >= 0.0))
		j = 0;
	else if(t >= (double) n)
		j = n - 2;
	else
		j = (int) t + 1;
	if(j > n - 2)
		j = n - 2;
	const double xj = S.x[j];
	const double xn = S.x[j + 1];
	const DmsCubic c = T[j];
	const bool ok = (j == n - 2 || !(xn <= x)) && (j == 0 || !(xj > x));
	if(__builtin_expect(ok, 1))
		return dms_cubic(c, x - xj);
	const int k = dms_interval(S, x);
	return dms_cubic(T[k], x - S.x[k]);
}

/* Solar_Model::Mass, src/Solar_Model.cpp:168-179 */
DMS_HD double dms_mass(const DmsSolarView& S, double r)
{
	if(__builtin_expect(!dm_isfinite(r), 0))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(r > DMS_rSun)
		return DMS_mSun;
	return 1.0 * dms_spline_at(S, S.mass, r);
}

/* Solar_Model::Temperature, src/Solar_Model.cpp:194-202 */
DMS_HD double dms_temperature(const DmsSolarView& S, double r)
{
	if(!dm_isfinite(r))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	const int j = dms_interval(S, r);
	return 1.0 * dms_cubic(S.temp[j], r - S.x[j]);
}

/* Solar_Model::Local_Escape_Speed, src/Solar_Model.cpp:204-218 */
DMS_HD double dms_local_escape_speed(const DmsSolarView& S, double r)
{
	if(!dm_isfinite(r))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(r > DMS_rSun)
		return dm_sqrt(2 * DMS_G_Newton * DMS_mSun / r);
	return dm_sqrt(1.0 * dms_spline_at(S, S.vesc2, r));
}

/* Number density of target k (k < n_iso: isotope, k == 63: electrons) at a
 * radius already known to lie in [0, R_sun]; src/Solar_Model.cpp:28-39,250-261 */
DMS_HD double dms_density_at(const DmsSolarView& S, int j, double t, int k)
{
	const double pref = (k == DMS_MAX_TARGETS - 1) ? 1.0 : S.iso_pref[k];
	return pref * dms_cubic(S.dens[(size_t) j * DMS_MAX_TARGETS + k], t);
}

/* Thermal_Averaged_Relative_Speed, src/Solar_Model.cpp:493-521 */
DMS_HD double dms_thermal_averaged_relative_speed(double temperature, double mass_target, double v_DM)
{
	if(!dm_isfinite(temperature) || temperature <= 0.0 || !dm_isfinite(mass_target) || mass_target <= 0.0
	   || !dm_isfinite(v_DM) || v_DM < 0.0)
		return dm_nan();
	const double kappa = dm_sqrt(mass_target / (2.0 * temperature));
	if(!dm_isfinite(kappa) || kappa <= 0.0)
		return dm_nan();
	if(v_DM == 0.0)
		return 2.0 / DM_SQRT_PI / kappa;
	const double x = kappa * v_DM;
	if(!dm_isfinite(x))
		return v_DM;
	if(x < 1.0e-4)
	{
		const double x2 = x * x;
		return 2.0 / DM_SQRT_PI / kappa * (1.0 + x2 / 3.0 - x2 * x2 / 30.0 + x2 * x2 * x2 / 210.0);
	}
	if(x > 26.0)
		return v_DM * (1.0 + 0.5 / (x * x));
	const double x2 = x * x;
	return v_DM * ((1.0 + 2.0 * x2) * dm_erf(x) / (2.0 * x2) + dm_exp(-x2) / (DM_SQRT_PI * x));
}

DMS_HD double dms_reduced_mass(double a, double b) { return a * b / (a + b); }

/* Sigma_Total_Nucleus / Sigma_Total_Electron (src/Dark_Photon.cpp:171-213;
 * SI low-mass contact is velocity independent). */
DMS_HD double dms_sigma_target(const DmsPointView& P, const DmsSolarView& S, int k, double v)
{
	const bool electron = (k == DMS_MAX_TARGETS - 1);
	double sigmatot = electron ? P.sigma_e : P.sigma_N[k];
	if(P.interaction == 2 && P.ff == 1)
	{
		const double mu = electron ? P.mu_e : dms_reduced_mass(P.mass, S.iso_mass[k]);
		const double q2max = 4.0 * ((mu * v) * (mu * v));
		const double md = P.m_dark_photon;
		const double Q = P.q_reference * P.q_reference + md * md;
		sigmatot = sigmatot * ((Q * Q) / md / md / (md * md + q2max));
	}
	return sigmatot;
}

/* DM_Scattering_Rate_Nucleus / _Electron, src/Solar_Model.cpp:263-303,
 * for target k at (r, v); k == 63 is the electron channel. */
DMS_HD double dms_rate_target(const DmsPointView& P, const DmsSolarView& S, double r, double v, int k)
{
	if(!dm_isfinite(r) || !dm_isfinite(v))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(v < 0.0)
		v = 0.0;
	if(r > DMS_rSun)
		return 0.0;
	const bool electron = (k == DMS_MAX_TARGETS - 1);
	const double m_target = electron ? DMS_mElectron : S.iso_mass[k];
	/* dms_interval's interval of r: the guess from the uniform spacing, its
	 * nodes and the temperature and density pieces read together (one memory
	 * round trip), dms_interval itself when the guess is off */
	const int nn = S.n;
	const double t = (r - S.x1) * S.inv_h;
	int j = !(t >= 0.0) ? 0 : ((t >= (double) nn) ? nn - 2 : (int) t + 1);
	j = (j > nn - 2) ? nn - 2 : j;
	double xj = S.x[j];
	const double xn = S.x[j + 1];
	DmsCubic ct = S.temp[j];
	DmsCubic cd = S.dens[(size_t) j * DMS_MAX_TARGETS + k];
	const bool ok = ((j == nn - 2) | !(xn <= r)) & ((j == 0) | !(xj > r));
	if(__builtin_expect(!ok, 0))
	{
		j = dms_interval(S, r);
		xj = S.x[j];
		ct = S.temp[j];
		cd = S.dens[(size_t) j * DMS_MAX_TARGETS + k];
	}
	const double dr = r - xj;
	/* Solar_Model::Temperature and the number density at r */
	const double v_rel = dms_thermal_averaged_relative_speed(1.0 * dms_cubic(ct, dr), m_target, v);
	const double pref = electron ? 1.0 : S.iso_pref[k];
	const double n = pref * dms_cubic(cd, dr);
	return n * dms_sigma_target(P, S, k, v) * v_rel;
}

/* Total_DM_Scattering_Rate_Computed, src/Solar_Model.cpp:324-342 */
DMS_HD double dms_rate_computed(const DmsPointView& P, const DmsSolarView& S, double r, double v)
{
	if(!dm_isfinite(r) || !dm_isfinite(v))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(v < 0.0)
		v = 0.0;
	if(r > DMS_rSun)
		return 0.0;
	double total = dms_rate_target(P, S, r, v, DMS_MAX_TARGETS - 1);
	for(int i = 0; i < S.n_iso; i++)
		total += dms_rate_target(P, S, r, v, i);
	return total;
}

/* Total_DM_Scattering_Rate_Interpolated, src/Solar_Model.cpp:344-393 */
DMS_HD double dms_rate_interpolated(const DmsPointView& P, const DmsSolarView& S, double r, double v)
{
	if(!dm_isfinite(r) || !dm_isfinite(v))
		return dm_nan();
	if(r > DMS_rSun || r < 0.0)
		return 0.0;
	if(v < 0.0)
		return 0.0;
	if(v > P.vmax)
		return dms_rate_computed(P, S, r, v);
	const double rc = r * P.inv_dr;
	const double vc = v * P.inv_dv;
	if(!dm_isfinite(rc) || !dm_isfinite(vc) || rc < 0.0 || vc < 0.0)
		return dms_rate_computed(P, S, r, v);
	unsigned ri = (unsigned) rc;
	if(ri > (unsigned) (P.Nr - 2))
		ri = (unsigned) (P.Nr - 2);
	unsigned vi = (unsigned) vc;
	if(vi > (unsigned) (P.Nv - 2))
		vi = (unsigned) (P.Nv - 2);
	const double rf = rc - ri;
	const double vf = vc - vi;
	const size_t lo = (size_t) ri * (size_t) P.Nv + vi;
	const size_t up = lo + (size_t) P.Nv;
	const double* g = P.grid;
	const double g0 = g[lo], g1 = g[lo + 1], g2 = g[up], g3 = g[up + 1];
	const double lower = g0 + vf * (g1 - g0);
	const double upper = g2 + vf * (g3 - g2);
	const double val = lower + rf * (upper - lower);
	return dm_isfinite(val) ? dm_max(0.0, val) : dm_nan();
}

/* Total_DM_Scattering_Rate, src/Solar_Model.cpp:305-322 */
DMS_HD double dms_rate_total(const DmsPointView& P, const DmsSolarView& S, double r, double v)
{
	if(!dm_isfinite(r) || !dm_isfinite(v))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(v < 0.0)
		v = 0.0;
	if(P.use_grid && v <= P.vmax)
		return dms_rate_interpolated(P, S, r, v);
	return dms_rate_computed(P, S, r, v);
}

/* ---- 3-vectors (libphysica::Vector semantics) ------------------------------ */
DMS_HD double dms_dot(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
DMS_HD double dms_norm(const double* a) { return dm_sqrt(dms_dot(a, a)); }
DMS_HD void dms_cross(const double* a, const double* b, double* o)
{
	const double x = a[1] * b[2] - a[2] * b[1];
	const double y = a[2] * b[0] - a[0] * b[2];
	const double z = a[0] * b[1] - a[1] * b[0];
	o[0] = x;
	o[1] = y;
	o[2] = z;
}
DMS_HD void dms_normalize(const double* a, double* o)
{
	const double n = dms_norm(a);
	o[0] = a[0] / n;
	o[1] = a[1] / n;
	o[2] = a[2] / n;
}

/* Orthonormal frame about a unit axis e3: e1 = (ref x e3)/|.|, e2 = (e3 x e1)/|.|
 * with ref = z unless e3 is within ~26 degrees of it
 * (src/Simulation_Trajectory.cpp:118-123, src/Simulation_Utilities.cpp:233-238). */
DMS_HD void dms_frame(const double* e3, double* e1, double* e2)
{
	const bool zref = dm_fabs(e3[2]) < 0.9;
	const double ref[3] = {zref ? 0.0 : 1.0, 0.0, zref ? 1.0 : 0.0};
	double t[3];
	dms_cross(ref, e3, t);
	dms_normalize(t, e1);
	dms_cross(e3, e1, t);
	dms_normalize(t, e2);
}

/* Hyperbolic_Kepler_Shift, src/Simulation_Utilities.cpp:252-321 (event time unchanged) */
DMS_HD bool dms_kepler_shift(double* r, double* v, double R_final)
{
	const double mu = DMS_G_Newton * DMS_mSun;
	const double Ri = dms_norm(r);
	const double speed = dms_norm(v);
	const double tol = 1.0e-10;
	if(R_final < DMS_rSun || Ri < DMS_rSun)
		return false;
	if(!(dm_isfinite(R_final) && R_final > 0.0) || !(dm_isfinite(Ri) && Ri > 0.0) || !(dm_isfinite(speed) && speed > 0.0))
		return false;
	if(dm_fabs(R_final - Ri) <= tol * dm_max(Ri, R_final))
		return true;
	const double vr = dms_dot(r, v) / Ri;
	if(!dm_isfinite(vr) || vr == 0.0)
		return false;
	const double dir = (vr < 0.0) ? -1.0 : 1.0;
	if(dir < 0.0 && R_final > Ri)
		return false;
	if(dir > 0.0 && R_final < Ri)
		return false;
	const double energy = 0.5 * speed * speed - mu / Ri;
	if(!(dm_isfinite(energy) && energy > 0.0))
		return false;
	const double rhat[3] = {r[0] / Ri, r[1] / Ri, r[2] / Ri};
	double hv[3];
	dms_cross(r, v, hv);
	const double h = dms_norm(hv);
	const double h_scale = Ri * speed;
	if(!dm_isfinite(h) || h <= tol * h_scale)
	{
		const double fss = 2.0 * (energy + mu / R_final);
		if(!(dm_isfinite(fss) && fss > 0.0))
			return false;
		const double s = dir * dm_sqrt(fss);
		for(int c = 0; c < 3; c++)
		{
			r[c] = R_final * rhat[c];
			v[c] = s * rhat[c];
		}
		return true;
	}
	const double p = h * h / mu;
	double vxh[3];
	dms_cross(v, hv, vxh);
	const double ev[3] = {vxh[0] / mu - rhat[0], vxh[1] / mu - rhat[1], vxh[2] / mu - rhat[2]};
	const double ecc = dms_norm(ev);
	if(!(dm_isfinite(p) && p > 0.0) || !dm_isfinite(ecc) || ecc <= 1.0)
		return false;
	const double cos_raw = (p / R_final - 1.0) / ecc;
	if(!dm_isfinite(cos_raw) || cos_raw > 1.0 + tol || cos_raw < -1.0 - tol)
		return false;
	double ax[3] = {ev[0] / ecc, ev[1] / ecc, ev[2] / ecc};
	const double az[3] = {hv[0] / h, hv[1] / h, hv[2] / h};
	double ay[3];
	dms_cross(az, ax, ay);
	const double ayn = dms_norm(ay);
	if(!(dm_isfinite(ayn) && ayn > 0.0))
		return false;
	for(int c = 0; c < 3; c++)
		ay[c] = ay[c] / ayn;
	double t[3];
	dms_cross(ay, az, t);
	dms_normalize(t, ax);
	const double theta = dir * dm_acos(dm_clamp_cos(cos_raw));
	const double ct = dm_cos(theta), st = dm_sin(theta);
	const double sv = dm_sqrt(mu / p);
	for(int c = 0; c < 3; c++)
	{
		r[c] = R_final * (ct * ax[c] + st * ay[c]);
		v[c] = sv * (-st * ax[c] + (ecc + ct) * ay[c]);
	}
	return true;
}

#endif
