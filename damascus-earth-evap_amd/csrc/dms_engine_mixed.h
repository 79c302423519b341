/*
 * Mixed-precision hot path of the trajectory kernel (DMS_MIXED).
 *
 * FP64 is kept for the RKF45 state, its step and error control, the 3-D event
 * reconstruction, the capture energy test and every time sum (bincount,
 * capture/unbinding times).  FP32 is used for the rate-table interpolation
 * (an FP32 copy of the table in units of 1/s), the optical-depth pieces with a
 * Kahan-compensated running optical depth, the per-target rates of
 * Sample_Target, the thermal target velocity and the scattering kinematics.
 * FP32 work runs in rescaled units (radii in R_sun, rates in 1/s, times in s),
 * so nothing under- or overflows.  The RKF45 right-hand side uses one
 * reciprocal of r per stage instead of the reference's division chain.
 *
 * These are not bit-compatible with the FP64 restatement: the contract is
 * statistical (capture fraction, histograms, evaporation times against the
 * FP64 path on the same trajectory ids).
 */
#ifndef DMS_ENGINE_MIXED_H
#define DMS_ENGINE_MIXED_H

/* Included by dms_engine.h after the lane state and the shared helpers. */
#include <math.h>

/* ---- FP32 elementary functions (hardware approximations on the device) ---- */
#if defined(__HIP_DEVICE_COMPILE__)
DMS_HD float mx_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
DMS_HD float mx_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
DMS_HD float mx_exp(float x) { return __expf(x); }
DMS_HD float mx_log(float x) { return __logf(x); }
DMS_HD float mx_cos(float x) { return __cosf(x); }
DMS_HD float mx_sin(float x) { return __sinf(x); }
DMS_HD float mx_erf(float x) { return erff(x); }
#else
DMS_HD float mx_sqrt(float x) { return sqrtf(x); }
DMS_HD float mx_rcp(float x) { return 1.0f / x; }
DMS_HD float mx_exp(float x) { return expf(x); }
DMS_HD float mx_log(float x) { return logf(x); }
DMS_HD float mx_cos(float x) { return cosf(x); }
DMS_HD float mx_sin(float x) { return sinf(x); }
DMS_HD float mx_erf(float x) { return erff(x); }
#endif
DMS_HD bool mx_isfinite(float x) { return x - x == 0.0f; }

#define DMS_MX_INV_RSUN (1.0 / DMS_rSun)
#define DMS_MX_INV_SEC (1.0 / DMS_sec)

/* Mass(r): the uniform-spacing interval guess without the exact neighbour
 * check (the Steffen pieces are C1 across nodes, so a guess off by one node
 * evaluates the neighbouring piece a rounding error outside its interval). */
DMS_HD double dms_mass_mx(const DmsSolarView& S, double r)
{
	r = (r < 0.0) ? 0.0 : r;
	const double t = (r - S.x1) * S.inv_h;
	int j = (t >= 0.0) ? ((t < (double) S.n) ? (int) t + 1 : S.n - 2) : 0;
	j = (j > S.n - 2) ? S.n - 2 : j;
	const double m = dms_cubic(S.mass[j], r - S.x[j]);
	return (r > DMS_rSun) ? DMS_mSun : m;
}

/* One RKF45 attempt (Runge_Kutta_45_Step, src/Simulation_Trajectory.cpp:1487-1602)
 * with the right-hand side J^2/r^3 - G M(r)/r^2, J/r^2 from one refined
 * reciprocal per stage.  PAIR: the stages as pairs (dms_rk45_stages_paired;
 * on the device lanes 0 and 1 of the wave run the attempt together, second =
 * lane 1), abandoned with p unchanged (returns -2) when *watch moves off
 * watch_for. */
template <bool PAIR = false>
DMS_HD int dms_rk45_attempt_mx(DmsProp& p, const DmsSolarView& S, int inner_iter, bool second = false,
                               const unsigned int* watch = nullptr, unsigned int watch_for = 0u)
{
	const double tol0 = 1.0 * DMS_km, tol1 = 1.0e-3 * DMS_km / DMS_sec, tol2 = 1.0e-7;
	const double h = p.ts;
	const double J2 = p.J * p.J;
	const bool angular = !(p.radial || p.J == 0.0);
	double kr[6], kv[6], kp[6];
	auto rhs = [&](double ri, double& dv, double& dphi) {
		/* one refined reciprocal (within an ulp) of the unclamped radius, which
		 * dphi/dt uses; dv/dt clamps the radius to 1 km, whose reciprocal is a
		 * constant */
		const double qu = dm_recip_shared(ri).y;
		const double q = (ri < 1.0 * DMS_km) ? 1.0 / DMS_km : qu;
		const double m = dms_mass_mx(S, ri);
		dv = (J2 * q - DMS_G_Newton * m) * (q * q);
		dphi = (angular && ri > 0.0) ? p.J * (qu * qu) : 0.0;
	};
	if constexpr(PAIR)
	{
		if(!dms_rk45_stages_paired(p, h, rhs, second, kr, kv, kp, watch, watch_for))
			return -2;
	}
	else
	{
		(void) second;
		(void) watch;
		(void) watch_for;
		double dv, dp;
		rhs(p.radius, dv, dp);
		kr[0] = h * p.vr;
		kv[0] = h * dv;
		kp[0] = h * dp;
		double ri = p.radius + kr[0] / 4.0;
		kr[1] = h * (p.vr + kv[0] / 4.0);
		rhs(ri, dv, dp);
		kv[1] = h * dv;
		ri = p.radius + 3.0 / 32.0 * kr[0] + 9.0 / 32.0 * kr[1];
		kr[2] = h * (p.vr + 3.0 / 32.0 * kv[0] + 9.0 / 32.0 * kv[1]);
		rhs(ri, dv, dp);
		kv[2] = h * dv;
		kp[2] = h * dp;
		ri = p.radius + 1932.0 / 2197.0 * kr[0] - 7200.0 / 2197.0 * kr[1] + 7296.0 / 2197.0 * kr[2];
		kr[3] = h * (p.vr + 1932.0 / 2197.0 * kv[0] - 7200.0 / 2197.0 * kv[1] + 7296.0 / 2197.0 * kv[2]);
		rhs(ri, dv, dp);
		kv[3] = h * dv;
		kp[3] = h * dp;
		ri = p.radius + 439.0 / 216.0 * kr[0] - 8.0 * kr[1] + 3680.0 / 513.0 * kr[2] - 845.0 / 4104.0 * kr[3];
		kr[4] = h * (p.vr + 439.0 / 216.0 * kv[0] - 8.0 * kv[1] + 3680.0 / 513.0 * kv[2] - 845.0 / 4104.0 * kv[3]);
		rhs(ri, dv, dp);
		kv[4] = h * dv;
		kp[4] = h * dp;
		ri = p.radius - 8.0 / 27.0 * kr[0] + 2.0 * kr[1] - 3544.0 / 2565.0 * kr[2] + 1859.0 / 4104.0 * kr[3] - 11.0 / 40.0 * kr[4];
		kr[5] = h * (p.vr - 8.0 / 27.0 * kv[0] + 2.0 * kv[1] - 3544.0 / 2565.0 * kv[2] + 1859.0 / 4104.0 * kv[3] - 11.0 / 40.0 * kv[4]);
		rhs(ri, dv, dp);
		kv[5] = h * dv;
		kp[5] = h * dp;
	}
	const double r4 = p.radius + 25.0 / 216.0 * kr[0] + 1408.0 / 2565.0 * kr[2] + 2197.0 / 4104.0 * kr[3] - 1.0 / 5.0 * kr[4];
	const double v4 = p.vr + 25.0 / 216.0 * kv[0] + 1408.0 / 2565.0 * kv[2] + 2197.0 / 4104.0 * kv[3] - 1.0 / 5.0 * kv[4];
	const double p4 = p.phi + 25.0 / 216.0 * kp[0] + 1408.0 / 2565.0 * kp[2] + 2197.0 / 4104.0 * kp[3] - 1.0 / 5.0 * kp[4];
	/* y5 - y4 directly from the stage differences (Fehlberg's error weights) */
	const double e0 = dm_fabs(1.0 / 360.0 * kr[0] - 128.0 / 4275.0 * kr[2] - 2197.0 / 75240.0 * kr[3] + 1.0 / 50.0 * kr[4] + 2.0 / 55.0 * kr[5]);
	const double e1 = dm_fabs(1.0 / 360.0 * kv[0] - 128.0 / 4275.0 * kv[2] - 2197.0 / 75240.0 * kv[3] + 1.0 / 50.0 * kv[4] + 2.0 / 55.0 * kv[5]);
	const double e2 = dm_fabs(1.0 / 360.0 * kp[0] - 128.0 / 4275.0 * kp[2] - 2197.0 / 75240.0 * kp[3] + 1.0 / 50.0 * kp[4] + 2.0 / 55.0 * kp[5]);
	double h_new;
	const bool finite = dm_isfinite(e0) && dm_isfinite(e1) && dm_isfinite(e2);
	{
		/* min over components of 0.84 (tol/err)^(1/4) = 0.84 (min tol/err)^(1/4) */
		double ratio = 1.0e300; /* tol/err by refined reciprocals: an ulp does not matter here */
		ratio = (e0 > 0.0) ? dm_min(ratio, tol0 * dm_recip_shared(e0).y) : ratio;
		ratio = (e1 > 0.0) ? dm_min(ratio, tol1 * dm_recip_shared(e1).y) : ratio;
		ratio = (e2 > 0.0) ? dm_min(ratio, tol2 * dm_recip_shared(e2).y) : ratio;
		double factor = dm_min(DMS_RK45_MAX_STEP_FACTOR, 0.84 * dm_pow_quarter(ratio));
		factor = dm_max(DMS_RK45_MIN_STEP_FACTOR, factor);
		h_new = finite ? dms_rk45_sanitized(factor * h) : DMS_RK45_MIN_STEP_FACTOR * h;
	}
	const bool within = finite && !(e0 > tol0) && !(e1 > tol1) && !(e2 > tol2);
	if(within)
	{
		p.time = p.time + h;
		p.radius = (r4 < 0.0) ? 0.0 : r4;
		p.vr = v4;
		p.phi = p4;
		p.ts = h_new;
		return 1;
	}
	const double abs_min_step = 1.0e-8 * DMS_sec;
	if(h_new < abs_min_step || inner_iter >= DMS_RK45_MAX_INNER_RETRIES)
	{
		const bool state_ok = dm_isfinite(r4) && dm_isfinite(v4) && dm_isfinite(p4);
		p.time = p.time + abs_min_step;
		if(state_ok)
		{
			p.radius = (r4 < 0.0) ? 0.0 : r4;
			p.vr = v4;
			p.phi = p4;
		}
		p.ts = abs_min_step;
		return 0;
	}
	p.ts = h_new;
	return -1;
}

/* ---- rates in FP32 (1/s), radius in R_sun --------------------------------- */

/* Thermal_Averaged_Relative_Speed in FP32, x = kappa v */
DMS_HD float dms_thermal_average_mx(float kappa, float v)
{
	const float x = kappa * v;
	const float two_over_sqrt_pi = 1.1283791670955126f;
	if(!(x > 1.0e-3f))
		return two_over_sqrt_pi / kappa * (1.0f + x * x * (1.0f / 3.0f));
	if(x > 26.0f)
		return v * (1.0f + 0.5f * mx_rcp(x * x));
	const float x2 = x * x;
	return v * ((1.0f + 2.0f * x2) * mx_erf(x) * mx_rcp(2.0f * x2) + mx_exp(-x2) * mx_rcp(1.7724538509055159f * x));
}

/* DM_Scattering_Rate_Nucleus / _Electron in 1/s for target k (k == 63:
 * electrons); the density-times-cross-section prefactor is formed in FP64 and
 * rescaled before the FP32 thermal average. */
DMS_HD float dms_rate_target_mx(const DmsPointView& P, const DmsSolarView& S, double r, double v, int k)
{
	if(!(r <= DMS_rSun))
		return 0.0f;
	r = (r < 0.0) ? 0.0 : r;
	const bool electron = (k == DMS_MAX_TARGETS - 1);
	const double m_target = electron ? DMS_mElectron : S.iso_mass[k];
	const double t = (r - S.x1) * S.inv_h;
	int j = (t >= 0.0) ? ((t < (double) S.n) ? (int) t + 1 : S.n - 2) : 0;
	j = (j > S.n - 2) ? S.n - 2 : j;
	const double dr = r - S.x[j];
	const double temperature = dms_cubic(S.temp[j], dr);
	const double n = dms_density_at(S, j, dr, k);
	const float pref = (float) (n * dms_sigma_target(P, S, k, v) * DMS_sec);
	const float kappa = (float) dm_sqrt(m_target / (2.0 * temperature));
	const float vr = dms_thermal_average_mx(kappa, (float) v);
	const float rate = pref * vr;
	return (temperature > 0.0 && mx_isfinite(rate)) ? rate : __builtin_nanf("");
}

DMS_HD float dms_rate_computed_mx(const DmsPointView& P, const DmsSolarView& S, double r, double v)
{
	float total = dms_rate_target_mx(P, S, r, v, DMS_MAX_TARGETS - 1);
	for(int i = 0; i < S.n_iso; i++)
		total += dms_rate_target_mx(P, S, r, v, i);
	return total;
}

/* Total_DM_Scattering_Rate in 1/s at radius rr (R_sun units) and speed v, from
 * the FP32 table (bilinear) or, without a table or above v_max, directly. */
DMS_HD float dms_rate_total_mx(const DmsPointView& P, const DmsSolarView& S, float rr, float v)
{
	if(!mx_isfinite(rr) || !mx_isfinite(v))
		return __builtin_nanf("");
	rr = (rr < 0.0f) ? 0.0f : rr;
	v = (v < 0.0f) ? 0.0f : v;
	if(rr > 1.0f)
		return 0.0f;
	if(!(P.use_grid && v <= (float) P.vmax))
		return dms_rate_computed_mx(P, S, (double) rr * DMS_rSun, (double) v);
	const float rc = rr * (float) (P.Nr - 1);
	const float vc = v * (float) P.inv_dv;
	const unsigned ri = (unsigned) fminf(rc, (float) (P.Nr - 2));
	const unsigned vi = (unsigned) fminf(vc, (float) (P.Nv - 2));
	const float rf = rc - (float) ri, vf = vc - (float) vi;
	const float* q = P.grid32 + ri * (unsigned) P.Nv + vi;
	const float lower = q[0] + vf * (q[1] - q[0]);
	const float upper = q[P.Nv] + vf * (q[P.Nv + 1] - q[P.Nv]);
	const float val = lower + rf * (upper - lower);
	return mx_isfinite(val) ? fmaxf(0.0f, val) : __builtin_nanf("");
}

/* One element of dms_rates_at_fractions_mx up to its table read (FP32, radius
 * in R_sun): radius and speed at fraction f of the step (b + f d) or the
 * end-point query (xr, xv), the lookup cell, and which evaluation it needs. */
struct DmsRateSlotMx
{
	float res, rad, spd, rf, vf;
	unsigned lo;
	bool grid, comp;
};

DMS_HD DmsRateSlotMx dms_rate_slot_mx(const DmsPointView& P, bool use_grid, bool active, bool at_fraction, const float* b,
                                      const float* d, float f, float xr, float xv)
{
	DmsRateSlotMx q;
	float r, v;
	bool outside = false;
	if(!at_fraction)
	{
		r = xr;
		v = xv;
	}
	else
	{
		f = fminf(1.0f, fmaxf(0.0f, f));
		const float p0 = b[0] + f * d[0], p1 = b[1] + f * d[1], p2 = b[2] + f * d[2];
		const float q0 = b[3] + f * d[3], q1 = b[4] + f * d[4], q2 = b[5] + f * d[5];
		r = mx_sqrt(p0 * p0 + p1 * p1 + p2 * p2);
		v = mx_sqrt(q0 * q0 + q1 * q1 + q2 * q2);
		outside = r >= 1.0f;
	}
	/* conditions combined with & (no short-circuit): selects, not branches */
	const bool finite = mx_isfinite(r) & mx_isfinite(v);
	q.rad = r;
	q.spd = v;
	const bool table = use_grid & (v <= (float) P.vmax);
	const bool live = active & !outside & finite & !(r > 1.0f);
	q.grid = live & table;
	q.comp = live & !table;
	const float rc = r * (float) (P.Nr - 1);
	const float vc = v * (float) P.inv_dv;
	const unsigned ri = q.grid ? (unsigned) fminf(rc, (float) (P.Nr - 2)) : 0u;
	const unsigned vi = q.grid ? (unsigned) fminf(vc, (float) (P.Nv - 2)) : 0u;
	q.rf = rc - (float) ri;
	q.vf = vc - (float) vi;
	q.lo = ri * (unsigned) P.Nv + vi;
	q.res = (!finite && !outside) ? __builtin_nanf("") : 0.0f;
	return q;
}

DMS_HD float dms_rate_slot_bilinear_mx(const DmsRateSlotMx& q, const float* g)
{
	const float lower = g[0] + q.vf * (g[1] - g[0]);
	const float upper = g[2] + q.vf * (g[3] - g[2]);
	const float val = lower + q.rf * (upper - lower);
	return q.grid ? (mx_isfinite(val) ? fmaxf(0.0f, val) : __builtin_nanf("")) : q.res;
}

/* Scattering_Rate_At_Fraction for n fractions and the next step's rate at the
 * end point (xr, xv), FP32.  Elements 1..n-1 and the end-point query take
 * one batch of table reads; element 0, the step's start point, is evaluated
 * only when its cached rate is not available (use_cached). */
/* Query k of dms_rates_at_fractions_mx alone (k < NQ - 1: fraction k, k ==
 * NQ - 1: the extra query), the same evaluation: the lanes of a helper wave
 * evaluate one query each. */
DMS_HD float dms_rate_fraction_slot_mx(const DmsPointView& P, const DmsSolarView& S, int k, const float* b, const float* a,
                                       const float* fr, int n, bool use_cached, float cached_rate, bool extra, float xr,
                                       float xv)
{
	constexpr int NQ = DMS_MAX_OPTICAL_DEPTH_PIECES + 2;
	if((k == 0) & use_cached)
		return cached_rate;
	const bool use_grid = P.use_grid != 0;
	const bool is_extra = (k == NQ - 1);
	const float d[6] = {a[0] - b[0], a[1] - b[1], a[2] - b[2], a[3] - b[3], a[4] - b[4], a[5] - b[5]};
	DmsRateSlotMx q = dms_rate_slot_mx(P, use_grid, is_extra ? extra : (k < n), !is_extra, b, d, fr[is_extra ? 0 : k], xr, xv);
	if(use_grid)
	{
		const unsigned nv = (unsigned) P.Nv;
		const float* t = P.grid32 + q.lo;
		const float g[4] = {t[0], t[1], t[nv], t[nv + 1]};
		q.res = dms_rate_slot_bilinear_mx(q, g);
	}
	if(__builtin_expect(q.comp, 0))
		q.res = dms_rate_computed_mx(P, S, (double) q.rad * DMS_rSun, (double) q.spd);
	return q.res;
}

DMS_HD void dms_rates_at_fractions_mx(const DmsPointView& P, const DmsSolarView& S, const float* b, const float* a,
                                      const float* fr, int n, bool use_cached, float cached_rate, float* out, bool extra,
                                      float xr, float xv, float* xout)
{
	constexpr int NQ = DMS_MAX_OPTICAL_DEPTH_PIECES + 2;
	const bool use_grid = P.use_grid != 0;
	const unsigned nv = (unsigned) P.Nv;
	const float d[6] = {a[0] - b[0], a[1] - b[1], a[2] - b[2], a[3] - b[3], a[4] - b[4], a[5] - b[5]};
	DmsRateSlotMx q[NQ];
#pragma unroll
	for(int i = 1; i < NQ; i++)
	{
		const bool is_extra = (i == NQ - 1);
		q[i] = dms_rate_slot_mx(P, use_grid, is_extra ? extra : (i < n), !is_extra, b, d, fr[is_extra ? 0 : i], xr, xv);
	}
	if(use_grid)
	{
		float g[NQ][4];
#pragma unroll
		for(int i = 1; i < NQ; i++)
		{
			const float* t = P.grid32 + q[i].lo;
			g[i][0] = t[0];
			g[i][1] = t[1];
			g[i][2] = t[nv];
			g[i][3] = t[nv + 1];
		}
#pragma unroll
		for(int i = 1; i < NQ; i++)
			q[i].res = dms_rate_slot_bilinear_mx(q[i], g[i]);
	}
	if(use_cached)
	{
		q[0].res = cached_rate;
		q[0].comp = false;
	}
	else
	{
		q[0] = dms_rate_slot_mx(P, use_grid, 0 < n, true, b, d, fr[0], xr, xv);
		if(use_grid)
		{
			const float* t = P.grid32 + q[0].lo;
			const float g0[4] = {t[0], t[1], t[nv], t[nv + 1]};
			q[0].res = dms_rate_slot_bilinear_mx(q[0], g0);
		}
	}
	bool any = false;
#pragma unroll
	for(int i = 0; i < NQ; i++)
		any = any || q[i].comp;
	if(__builtin_expect(any, 0)) /* no table, or v above it: rare */
	{
#pragma unroll
		for(int i = 0; i < NQ; i++)
			if(q[i].comp)
				q[i].res = dms_rate_computed_mx(P, S, (double) q[i].rad * DMS_rSun, (double) q[i].spd);
	}
	/* out has NQ - 1 elements; those from n on are written but not read */
#pragma unroll
	for(int i = 0; i < NQ - 1; i++)
		out[i] = q[i].res;
	if(extra)
		*xout = q[NQ - 1].res;
}

/* Solve_Linear_Rate_Optical_Depth_Fraction (:416-464) in FP32: the fraction at
 * which a rate linear in time from rs to re accumulates target_tau over dt. */
DMS_HD float dms_solve_linear_rate_mx(float rs, float re, float dt, float target_tau)
{
	if(!(target_tau > 0.0f))
		return 0.0f;
	const float total = 0.5f * (rs + re) * dt;
	if(!mx_isfinite(total) || !(total > 0.0f) || target_tau >= total)
		return 1.0f;
	const float a = 0.5f * (re - rs) * dt, bq = rs * dt;
	float x;
	if(fabsf(a) <= 1.0e-6f * fabsf(bq))
		x = target_tau / bq;
	else
	{
		const float disc = fmaxf(0.0f, bq * bq + 4.0f * a * target_tau);
		/* stable root of a x^2 + bq x - tau = 0 */
		x = (2.0f * target_tau) / (bq + mx_sqrt(disc));
	}
	return (mx_isfinite(x) && x >= 0.0f) ? fminf(1.0f, x) : 1.0f;
}

/* Kahan-compensated subtraction of dtau from the remaining optical depth. */
DMS_HD void dms_kahan_sub_mx(float& sum, float& comp, float dtau)
{
	const float y = -dtau - comp;
	const float t = sum + y;
	comp = (t - sum) - y;
	sum = t;
}

/* ---- per-step bookkeeping (FP64, fewer divisions and square roots) -------- */

/* Event_In_3D (:1715-1742) with one reciprocal for J/r, plus the speed. */
DMS_HD double dms_prop_event_mx(const DmsProp& p, double& t, double* r, double* v)
{
	const double safe_radius = (dm_isfinite(p.radius) && p.radius > 0.0) ? p.radius : 0.0;
	t = p.time;
	if(p.radial || p.J == 0.0 || safe_radius == 0.0)
	{
		for(int c = 0; c < 3; c++)
		{
			r[c] = safe_radius * p.ax[c];
			v[c] = p.vr * p.ax[c];
		}
		return dm_fabs(p.vr);
	}
	double sin_phi, cos_phi;
	dm_sincos(p.phi, &sin_phi, &cos_phi);
	const double tangential = p.J * dm_recip_shared(safe_radius).y; /* refined reciprocal, within an ulp */
	const double vx = p.vr * cos_phi - tangential * sin_phi;
	const double vy = p.vr * sin_phi + tangential * cos_phi;
	for(int c = 0; c < 3; c++)
	{
		r[c] = safe_radius * (cos_phi * p.ax[c] + sin_phi * p.ay[c]);
		v[c] = vx * p.ax[c] + vy * p.ay[c];
	}
	return dm_sqrt(p.vr * p.vr + tangential * tangential);
}

/* E = m (v^2 - v_esc^2) / 2 in eV from v^2, without the square root of v_esc^2
 * (outside the Sun 2 G M / r by a refined reciprocal). */
DMS_HD double dms_capture_energy_mx(const DmsSolarView& S, double dm_mass, double radius, double speed2)
{
	double vesc2;
	if(radius > DMS_rSun)
		vesc2 = (2.0 * DMS_G_Newton * DMS_mSun) * dm_recip_shared(radius).y;
	else
	{
		const double r = (radius < 0.0) ? 0.0 : radius;
		const double t = (r - S.x1) * S.inv_h;
		int j = (t >= 0.0) ? ((t < (double) S.n) ? (int) t + 1 : S.n - 2) : 0;
		j = (j > S.n - 2) ? S.n - 2 : j;
		vesc2 = dms_cubic(S.vesc2[j], r - S.x[j]);
	}
	return 0.5 * dm_mass * (speed2 - vesc2) * (1.0 / DMS_eV);
}

/* The commit lambda (:842-874) for a free-flight step: bincount of the previous
 * point and Update_Capture_State(..., allow_new = false), given the committed
 * event's radius and squared speed. */
DMS_HD void dms_commit_mx(DmsLane& L, const DmsRows& R, const DmsJobView& J, const DmsSolarView& S, double dm_mass, double lt,
                          double radius, double speed2)
{
	const double t_now = lt * (1.0 / DMS_sec);
	if(!J.terminate_on_capture)
	{
		if(L.prev_time_sec >= 0.0)
			dms_accumulate_bincount(L, R, J, L.prev_v2, t_now - L.prev_time_sec);
		L.prev_time_sec = t_now;
		L.prev_r_km = radius * (1.0 / DMS_km);
		L.prev_v2 = speed2 * ((DMS_sec / DMS_km) * (DMS_sec / DMS_km));
		dms_bin_anchor(L, R, J);
	}
	L.last_E = dm_nan();
	if(J.terminate_on_capture && !L.captured)
		return;
	const double E = dms_capture_energy_mx(S, dm_mass, radius, speed2);
	L.last_E = E;
	L.prev_cap_E = E;
	if((L.captured != 0) & dm_isfinite(L.ff_ref_E))
	{
		const double drift = dm_fabs(E - L.ff_ref_E);
		L.drift_eV = (dm_isfinite(drift) & (drift > L.drift_eV)) ? drift : L.drift_eV;
		const double rel = drift * dm_recip_shared(dm_max(dm_fabs(L.ff_ref_E), DMS_FREE_ENERGY_DRIFT_REL_E_SCALE_EV)).y;
		L.drift_rel = (dm_isfinite(rel) & (rel > L.drift_rel)) ? rel : L.drift_rel;
	}
	if((L.bound != 0) & (L.captured != 0)) /* a store, not a select: the field lives in scratch */
		L.t_last_bound = (lt + L.time_origin) * (1.0 / DMS_sec);
}

/* ---- scattering in FP32 -------------------------------------------------- */

/* Unit_Vector_At_Angle_From_Axis (:109-128) in FP32. */
DMS_HD bool dms_unit_vector_at_angle_mx(float cos_theta, float phi, const float* axis, float* out)
{
	const float an = mx_sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
	if(!(an > 0.0f) || !mx_isfinite(an) || !mx_isfinite(cos_theta) || !mx_isfinite(phi))
		return false;
	const float ia = mx_rcp(an);
	const float e3[3] = {axis[0] * ia, axis[1] * ia, axis[2] * ia};
	const bool zref = fabsf(e3[2]) < 0.9f;
	/* e1 = normalize(ref x e3), e2 = e3 x e1 (unit already) */
	float t0 = zref ? -e3[1] : 0.0f;
	float t1 = zref ? e3[0] : -e3[2];
	float t2 = zref ? 0.0f : e3[1];
	const float it = mx_rcp(mx_sqrt(t0 * t0 + t1 * t1 + t2 * t2));
	const float e1[3] = {t0 * it, t1 * it, t2 * it};
	const float e2[3] = {e3[1] * e1[2] - e3[2] * e1[1], e3[2] * e1[0] - e3[0] * e1[2], e3[0] * e1[1] - e3[1] * e1[0]};
	cos_theta = fminf(1.0f, fmaxf(-1.0f, cos_theta));
	const float st = mx_sqrt(fmaxf(0.0f, 1.0f - cos_theta * cos_theta));
	const float cp = mx_cos(phi), sp = mx_sin(phi);
	for(int c = 0; c < 3; c++)
		out[c] = cos_theta * e3[c] + st * cp * e1[c] + st * sp * e2[c];
	return true;
}

/* Sample_Target_Velocity (Romano & Walsh) in FP32; velocities in units of c. */
DMS_HD bool dms_sample_target_velocity_mx(DmsLane& L, double temperature, double target_mass, const float* vel, float* out)
{
	if(!(temperature > 0.0) || !(target_mass > 0.0))
		return false;
	const float kappa = (float) dm_sqrt(target_mass / (2.0 * temperature));
	const float vDM = mx_sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
	if(!mx_isfinite(kappa) || !(vDM > 0.0f) || !mx_isfinite(vDM))
		return false;
	const float y = kappa * vDM;
	const float branch = 2.0f * mx_rcp(1.7724538509055159f * y + 2.0f);
	float x = y, mu = 1.0f;
	for(unsigned long attempts = 0;; attempts++)
	{
		L.n_tv++;
		const float raw = fmaxf(0.0f, x * x + y * y - 2.0f * x * y * mu);
		const float ratio = mx_sqrt(raw) * mx_rcp(x + y);
		if(!mx_isfinite(ratio))
			return false;
		if(!(fminf(1.0f, ratio) < (float) dm_uniform(&L.rng)))
			break;
		if(attempts >= DMS_TARGET_VELOCITY_MAX_REJECTION_ATTEMPTS)
			return false;
		mu = -1.0f + 2.0f * (float) dm_uniform(&L.rng);
		const float xi_1 = (float) dm_uniform(&L.rng);
		float z;
		if(xi_1 < branch)
		{
			const float xi_2 = fmaxf((float) dm_uniform(&L.rng), 1.0e-38f);
			const float xi_3 = fmaxf((float) dm_uniform(&L.rng), 1.0e-38f);
			z = -mx_log(xi_2) - mx_log(xi_3);
		}
		else
		{
			const float xi_2 = fmaxf((float) dm_uniform(&L.rng), 1.0e-38f);
			const float xi_3 = (float) dm_uniform(&L.rng);
			const float xi_4 = fmaxf((float) dm_uniform(&L.rng), 1.0e-38f);
			const float c = mx_cos(1.5707963267948966f * xi_3);
			z = -mx_log(xi_2) - c * c * mx_log(xi_4);
		}
		x = mx_sqrt(z);
	}
	const float vT = x * mx_rcp(kappa);
	const float phi = 6.283185307179586f * (float) dm_uniform(&L.rng);
	float u[3];
	if(!dms_unit_vector_at_angle_mx(mu, phi, vel, u))
		return false;
	for(int c = 0; c < 3; c++)
		out[c] = vT * u[c];
	return true;
}

/* Scatter after the target choice, FP32 kinematics (New_DM_Velocity,
 * :1228-1247); the new velocity is written back in FP64. */
DMS_HD bool dms_scatter_rest_mx(DmsLane& L, const DmsPointView& P, const DmsSolarView& S, int target)
{
	const double r = dms_norm(L.ev_r);
	const float v64[3] = {(float) L.ev_v[0], (float) L.ev_v[1], (float) L.ev_v[2]};
	const float v = mx_sqrt(v64[0] * v64[0] + v64[1] * v64[1] + v64[2] * v64[2]);
	const double mT = (target == -1) ? DMS_mElectron : S.iso_mass[target];
	float vT[3];
	if(!dms_sample_target_velocity_mx(L, dms_temperature(S, r), mT, v64, vT))
		return false;
	const float xi = (float) dm_uniform(&L.rng);
	float cos_a;
	if(P.interaction == 2 && P.ff == 1)
	{
		const double mu = (target < 0) ? P.mu_e : P.mass * S.iso_mass[target] / (P.mass + S.iso_mass[target]);
		const float m2 = (float) (P.m_dark_photon / mu) * (float) (P.m_dark_photon / mu); /* in units of mu^2 */
		const float q2max = 4.0f * v * v;
		cos_a = (m2 * (2.0f * xi - 1.0f) + q2max * xi) * mx_rcp(m2 + q2max * xi);
	}
	else
		cos_a = 2.0f * xi - 1.0f;
	cos_a = fminf(1.0f, fmaxf(-1.0f, cos_a));
	const float phi = 6.283185307179586f * (float) dm_uniform(&L.rng);
	float n[3];
	if(!dms_unit_vector_at_angle_mx(cos_a, phi, v64, n))
		return false;
	const float d0 = vT[0] - v64[0], d1 = vT[1] - v64[1], d2 = vT[2] - v64[2];
	const float rel = mx_sqrt(d0 * d0 + d1 * d1 + d2 * d2);
	const float wT = (float) (mT / (mT + P.mass)), wD = (float) (P.mass / (mT + P.mass));
	const float f = wT * rel;
	float out[3];
	for(int c = 0; c < 3; c++)
		out[c] = f * n[c] + (wD * v64[c] + wT * vT[c]);
	if(!mx_isfinite(out[0]) || !mx_isfinite(out[1]) || !mx_isfinite(out[2]))
		return false;
	for(int c = 0; c < 3; c++)
		L.ev_v[c] = (double) out[c];
	return true;
}

#endif
