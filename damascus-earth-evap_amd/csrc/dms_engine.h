/*
 * Per-lane trajectory state machine of the FP64 kernel.
 *
 * The reference runs Trajectory_Simulator::Simulate (nested loops:
 * Simulate -> Propagate_Freely -> RKF45 retries) on one CPU thread.  Here the
 * loops are flattened into phases so that a wavefront executes one RKF45
 * attempt per iteration of the persistent loop in every busy lane, and the
 * rarer phases (scatter, segment start, trajectory end) diverge briefly:
 *
 *   PREP  top of the Propagate_Freely step loop (step caps, rate_before)
 *   RK    one RKF45 attempt (inner retry loop unrolled across iterations)
 *   POST  optical depth, scatter location, boundary / Kepler return, commit
 *   SCAT  Scatter (target choice is done wave-cooperatively by the kernel)
 *   SEG   Propagate_Freely entry (new free-flight segment)
 *   FIN   post-loop classification of Simulate
 *
 * All arithmetic follows src/Simulation_Trajectory.cpp in the reference with
 * the same operation order, so under -ffp-contract=off a lane reproduces the
 * CPU restatement bit for bit.
 */
#ifndef DMS_ENGINE_H
#define DMS_ENGINE_H

#include "dms_physics.h"
#include "philox.h"
#include "../../include/dms.h"

enum DmsPhase
{
	PH_DONE = 0,
	PH_NEED = 1,
	PH_START = 2,
	PH_SEG = 3,
	PH_PREP = 4,
	PH_RK = 5,
	PH_POST = 6,
	PH_SCAT = 7,
	PH_FIN = 8,
	PH_WAIT = 9 /* serving launch: holds a slot index not yet published */
};

/* Per-trajectory summary the host needs for run accounting. */
struct DmsSummary
{
	int32_t reason;
	uint32_t flags;
	uint64_t n_scat;
	double t_capture_sec;
	double t_unbind_sec;
	uint64_t n_rk;
	uint32_t n_int, n_ext, n_rej, n_tv;
};

struct DmsJobView
{
	uint64_t max_time_steps, max_scatterings, rk_budget;
	double max_distance;
	int terminate_on_capture;
	int nbins;
	double bin_width_km;
	uint32_t seed, point_id;
	uint64_t first_id;
	int mixed; /* DMS_MIXED: run dms_traj_kernel_mixed */
};

struct DmsProp
{
	double time, radius, phi, vr, J;
	double ax[3], ay[3];
	double ts;
	int radial;
};

struct DmsLane
{
	int phase;
	uint32_t slot;  /* trajectory slot in the launch */
	uint32_t row;   /* histogram row (slot, or the lane in per-lane mode) */
	uint32_t job;   /* job of the slot */
	uint32_t n_touched;
	uint64_t id;    /* trajectory id (Philox counter) */
	dm_rng rng;
	double ev_t, ev_r[3], ev_v[3];
	uint64_t n_scat;
	int reason;
	/* capture / survival bookkeeping (TrajectoryBincount scalars) */
	int captured, bound, event_observed, boundary_escape_observed, survival_valid, nie;
	double t_capture, t_last_bound, t_unbind, t_boundary, r_first_neg, E_first_neg, dE_first_neg;
	double drift_eV, drift_rel;
	double prev_time_sec, prev_r_km, prev_v2, prev_cap_E, ff_ref_E;
	/* bincount: running accumulator of the bin currently being filled */
	int cur_bin;
	double acc_dt, acc_v2;
	uint32_t acc_cnt;
	/* the bin of prev_r_km (-1: none) and its stored values, read when the
	 * anchor moved (dms_bin_anchor) */
	int pf_bin;
	uint32_t pf_cnt;
	double pf_dt, pf_v2;
	/* work counters */
	uint64_t rk_all;
	uint32_t n_int, n_ext, n_rej, n_tv;
	/* Propagate_Freely segment */
	double time_origin, mlx;
	uint64_t time_steps;
	uint32_t od_retries;
	DmsProp prop;
	/* step-level state carried from PREP to POST */
	double pb_time, pb_radius, pb_phi, pb_vr, pb_ts;
	double eb_t, eb_r[3], eb_v[3];
	double r_before, v_before, rate_before;
	int eb_valid; /* eb/r_before/v_before already hold Event_In_3D of the current propagator */
	int rb_cached; /* rate_before already holds Total_DM_Scattering_Rate(r_before, v_before) */
	/* mixed precision: rate_before in 1/s, remaining optical depth (Kahan pair) */
	float rate_before_f, mlx_f, mlx_c;
	int inner_iter, rk_ok;
	double last_E; /* capture energy of the last committed event (NaN when not computed) */
	/* scatter hand-off */
	double sc_r, sc_v, sc_xi;
	int sc_target;
};

/* One bin of a per-lane scratch row: the trajectory's time, v^2-weighted time
 * and visits in that bin, in one 32-byte record (one cache line per bin change
 * instead of one per array). */
struct DmsBinAcc
{
	double dt, v2dt;
	uint32_t cnt, pad0;
	uint64_t pad1;
};

/* Per-trajectory radial histogram storage.  Per-trajectory mode: one row per
 * trajectory slot (dt[nbins], v2dt[nbins] in FP64, counts[nbins]), summed
 * after the launch.  Per-lane mode: one scratch row of DmsBinAcc records per
 * resident lane plus the list of bins the current trajectory touched; at the
 * end of a trajectory its bins are added (FP64 atomics) to the sums of its
 * class and the scratch is cleared for the next trajectory of the lane. */
struct DmsRows
{
	double* rows;       /* per-trajectory mode: [row][2][nbins] or NULL */
	uint32_t* counts;   /* per-trajectory mode: [row][nbins] or NULL */
	DmsBinAcc* bins;    /* per-lane mode: [row][nbins] */
	uint16_t* touched;  /* per-lane mode: [row][nbins] touched-bin list */
	int per_lane;
	double* hist;       /* per-lane mode: [job][8][nbins] class sums */
	unsigned long long* hcounts; /* per-lane mode: [job][2][nbins] visit counts */
};

DMS_HD double dms_rk45_sanitized(double step)
{
	return (!dm_isfinite(step) | (step <= 0.0)) ? 0.1 * DMS_sec : dm_min(step, 1.0e6 * DMS_sec);
}

/* Free_Particle_Propagator constructor, src/Simulation_Trajectory.cpp:1403-1464 */
DMS_HD void dms_prop_init(DmsProp& p, double t, const double* r, const double* v)
{
	p.time = t;
	p.radius = dms_norm(r);
	p.phi = 0.0;
	p.radial = 0;
	p.ts = 0.1 * DMS_sec;
	p.J = 0.0;
	const double speed = dms_norm(v);
	const double radius_eps = 1.0e-12 * DMS_km;
	const double angular_scale = dm_max(p.radius * speed, 1.0 * DMS_km * DMS_km / DMS_sec);
	const double angular_eps = 1.0e-12 * angular_scale;
	double L[3];
	dms_cross(r, v, L);
	const double Lnorm = dms_norm(L);
	if(dm_isfinite(p.radius) && p.radius > radius_eps)
	{
		for(int c = 0; c < 3; c++)
			p.ax[c] = r[c] / p.radius;
		p.vr = dms_dot(r, v) / p.radius;
	}
	else if(dm_isfinite(speed) && speed > 0.0)
	{
		for(int c = 0; c < 3; c++)
			p.ax[c] = v[c] / speed;
		p.vr = speed;
		p.radius = 0.0;
		p.radial = 1;
	}
	else
	{
		p.ax[0] = 1.0;
		p.ax[1] = 0.0;
		p.ax[2] = 0.0;
		p.vr = 0.0;
		p.radius = 0.0;
		p.radial = 1;
	}
	if(!p.radial && dm_isfinite(Lnorm) && Lnorm > angular_eps)
	{
		double az[3];
		for(int c = 0; c < 3; c++)
			az[c] = L[c] / Lnorm;
		dms_cross(az, p.ax, p.ay);
		const double ayn = dms_norm(p.ay);
		if(dm_isfinite(ayn) && ayn > 0.0)
		{
			for(int c = 0; c < 3; c++)
				p.ay[c] = p.ay[c] / ayn;
			double t3[3];
			dms_cross(p.ay, az, t3);
			dms_normalize(t3, p.ax);
			p.J = Lnorm;
		}
		else
			p.radial = 1;
	}
	else
		p.radial = 1;
	if(p.radial)
	{
		/* Build_Stable_Orthonormal_Basis, :130-142 */
		const double n = dms_norm(p.ax);
		const double e1[3] = {p.ax[0] / n, p.ax[1] / n, p.ax[2] / n};
		p.ax[0] = e1[0];
		p.ax[1] = e1[1];
		p.ax[2] = e1[2];
		dms_frame(e1, p.ay, L); /* L reused for the (unused) third axis */
		p.J = 0.0;
	}
}

DMS_HD double dms_prop_radius(const DmsProp& p) { return (p.radius < 0.0) ? 0.0 : p.radius; }

DMS_HD double dms_prop_speed(const DmsProp& p)
{
	const double r = (p.radius < 0.0) ? 0.0 : p.radius;
	if(__builtin_expect(r == 0 || p.J == 0, 0))
		return dm_fabs(p.vr);
	return dm_sqrt(p.vr * p.vr + p.J * p.J / r / r);
}

/* Event_In_3D, :1715-1742 */
DMS_HD void dms_prop_event(const DmsProp& p, double& t, double* r, double* v)
{
	const double safe_radius = (dm_isfinite(p.radius) && p.radius > 0.0) ? p.radius : 0.0;
	t = p.time;
	if(__builtin_expect(p.radial || p.J == 0.0 || safe_radius == 0.0, 0))
	{
		for(int c = 0; c < 3; c++)
		{
			r[c] = safe_radius * p.ax[c];
			v[c] = p.vr * p.ax[c];
		}
		return;
	}
	double sin_phi, cos_phi;
	dm_sincos(p.phi, &sin_phi, &cos_phi);
	const double tangential = p.J / safe_radius;
	const double vx = p.vr * cos_phi - tangential * sin_phi;
	const double vy = p.vr * sin_phi + tangential * cos_phi;
	for(int c = 0; c < 3; c++)
	{
		r[c] = safe_radius * (cos_phi * p.ax[c] + sin_phi * p.ay[c]);
		v[c] = vx * p.ax[c] + vy * p.ay[c];
	}
}

/* x in [2^lo, 2^hi) for exponents of normal numbers lo < hi: one unsigned
 * compare of the high word, false for negative, NaN and infinite x. */
DMS_HD bool dms_in_binades(double x, int lo, int hi)
{
	const uint32_t w = (uint32_t) (dm_bits(x) >> 32);
	const uint32_t a = (uint32_t) (lo + 1023) << 20, b = (uint32_t) (hi + 1023) << 20;
	return w - a < b - a;
}

/* dms_prop_speed and dms_prop_event of one propagator state.  On the device,
 * for an angular state with radius in [2^63, 2^132) (inside [1 km, 1e40]) and
 * |J| in [2^-99, 2^165), the three divisions by the radius share one refined
 * reciprocal (dm_div_shared: the dividends J^2, J^2 / r and J stay in its
 * exact domain); otherwise, and on the host, the two functions. */
DMS_HD double dms_prop_event_speed(const DmsProp& p, double& t, double* r, double* v)
{
#if defined(__HIP_DEVICE_COMPILE__)
	if(__builtin_expect(!p.radial & dms_in_binades(p.radius, 63, 132) & dms_in_binades(dm_fabs(p.J), -99, 165), 1))
	{
		const DmRecip R = dm_recip_shared(p.radius);
		const double speed = dm_sqrt(p.vr * p.vr + dm_div_shared(dm_div_shared(p.J * p.J, R), R));
		t = p.time;
		double sin_phi, cos_phi;
		dm_sincos(p.phi, &sin_phi, &cos_phi);
		const double tangential = dm_div_shared(p.J, R);
		const double vx = p.vr * cos_phi - tangential * sin_phi;
		const double vy = p.vr * sin_phi + tangential * cos_phi;
		for(int c = 0; c < 3; c++)
		{
			r[c] = p.radius * (cos_phi * p.ax[c] + sin_phi * p.ay[c]);
			v[c] = vx * p.ax[c] + vy * p.ay[c];
		}
		return speed;
	}
#endif
	const double speed = dms_prop_speed(p);
	dms_prop_event(p, t, r, v);
	return speed;
}

DMS_HD double dms_dv_dt(const DmsProp& p, double r, double mass)
{
	if(r < 1.0 * DMS_km)
		r = 1.0 * DMS_km;
	return p.J * p.J / r / r / r - DMS_G_Newton * mass / r / r;
}

DMS_HD double dms_dphi_dt(const DmsProp& p, double r)
{
	if(p.radial || p.J == 0.0 || r <= 0.0)
		return 0.0;
	return p.J / r / r;
}

/* Mass(r) (dms_mass), dms_dv_dt and (want_phi) dms_dphi_dt of one RKF45
 * stage.  On the device the common case is Mass() on the interval guess of
 * dms_spline_at (inside the Sun; M_sun outside) and the seven divisions by r
 * sharing one reciprocal (dm_div_shared), exact when every dividend and r lie
 * in its domain: r in [1 km, 1e40], |J| zero or in [1e-30, 1e50] (j_ok,
 * checked once per attempt) and G M zero or in [1e-60, 1e100] keep every
 * dividend of the chains in [1e-180, 1e100] (the tests use power-of-two
 * bounds inside these).  A missed interval guess or operands outside that
 * domain recompute the stage the reference's way in one rarely taken branch;
 * the host (CPU replay of the kernel code) always does. */
DMS_HD void dms_stage(const DmsProp& p, const DmsSolarView& S, double r, bool j_ok, bool want_phi, double& dv, double& dphi)
{
#if defined(__HIP_DEVICE_COMPILE__)
	const double rm = (r < 0.0) ? 0.0 : r;
	const bool finite = dm_isfinite(r);
	double mass = finite ? DMS_mSun : dm_nan();
	bool guess_ok = true;
	if(finite & (rm <= DMS_rSun)) /* outside the Sun Mass() is M_sun: no table read */
	{
		const int n = S.n;
		const double t = (rm - S.x1) * S.inv_h;
		/* dms_spline_at's guess (0 for t < 0, n - 2 for t >= n - 2) with the
		 * conversion kept in range: t clamped to [-1, n] */
		const double tc = (t >= 0.0) ? dm_min(t, (double) n) : -1.0;
		const int jt = (int) tc + 1;
		const int j = (jt > n - 2) ? n - 2 : jt;
		const double xj = S.x[j];
		const double xn = S.x[j + 1];
		const DmsCubic c = S.mass[j];
		mass = 1.0 * dms_cubic(c, rm - xj);
		guess_ok = (j == n - 2 || !(xn <= rm)) & (j == 0 || !(xj > rm));
	}
	const double gm = DMS_G_Newton * mass;
	/* [2^63, 2^132) lies inside [1 km, 1e40] (1 km = 2^62.14) and
	 * [2^-199, 2^332) inside [1e-60, 1e100] */
	const bool fast = j_ok & dms_in_binades(r, 63, 132) & ((gm == 0.0) | dms_in_binades(gm, -199, 332));
	if(__builtin_expect(guess_ok & fast, 1))
	{
		const DmRecip R = dm_recip_shared(r);
		const double jj = p.J * p.J;
		dv = dm_div_shared(dm_div_shared(dm_div_shared(jj, R), R), R) - dm_div_shared(dm_div_shared(gm, R), R);
		if(want_phi)
			dphi = (p.radial || p.J == 0.0) ? 0.0 : dm_div_shared(dm_div_shared(p.J, R), R);
		return;
	}
#endif
	const double m = dms_mass(S, r);
	dv = dms_dv_dt(p, r, m);
	if(want_phi)
		dphi = dms_dphi_dt(p, r);
}

#if defined(__HIP_DEVICE_COMPILE__)
/* lane l's value of x (l uniform over the wave, lane l active) */
__device__ inline double dms_lane_value(double x, int l)
{
	const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
	const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
	return __hiloint2double(hi, lo);
}
#endif

/* Two RKF45 stages whose radii are both known: F(ra) into (dva, dpa) and F(rb)
 * into (dvb, dpb), F one right-hand side evaluation (dv/dt, dphi/dt at a
 * radius).  On the device lanes 0 and 1 of the calling wave (both active, with
 * the same inputs; second = lane 1) evaluate one each in a single instruction
 * stream, and v_readlane hands both results to both lanes; on the host the two
 * evaluations run one after the other.  Each evaluation is the one a serial
 * attempt makes, so the results are the same bits. */
template <class F>
DMS_HD void dms_stage_pair(F&& stage, double ra, double rb, bool second, double& dva, double& dpa, double& dvb, double& dpb)
{
#if defined(__HIP_DEVICE_COMPILE__)
	double dv, dp;
	stage(second ? rb : ra, dv, dp);
	dva = dms_lane_value(dv, 0);
	dpa = dms_lane_value(dp, 0);
	dvb = dms_lane_value(dv, 1);
	dpb = dms_lane_value(dp, 1);
#else
	(void) second;
	stage(ra, dva, dpa);
	stage(rb, dvb, dpb);
#endif
}

/* Whether the word a helper watches (an LDS word of its workgroup; none on
 * the host) no longer holds v. */
DMS_HD bool dms_watch_moved(const unsigned int* w, unsigned int v)
{
#if defined(__HIP_DEVICE_COMPILE__)
	return (w != nullptr) && __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != v;
#else
	(void) w;
	(void) v;
	return false;
#endif
}

/* The six stage increments of an RKF45 attempt with step h (the tableau of
 * :1503-1540).  Stage i's radius needs the velocity increments of stages
 * before i - 1 only, so stages 0 and 1, 2 and 3, 4 and 5 are evaluated as
 * pairs (dms_stage_pair): three dependent right-hand side evaluations instead
 * of six.  Every increment is the same expression of the same values as in
 * the serial order.  Between the pairs the word *watch is checked: when it no
 * longer holds watch_for, the increments are abandoned (returns false). */
template <class F>
DMS_HD bool dms_rk45_stages_paired(const DmsProp& p, double h, F&& stage, bool second, double* kr, double* kv, double* kp,
                                   const unsigned int* watch, unsigned int watch_for)
{
	double dv0, dp0, dv1, dp1;
	kr[0] = h * p.vr;
	dms_stage_pair(stage, p.radius, p.radius + kr[0] / 4.0, second, dv0, dp0, dv1, dp1);
	kv[0] = h * dv0;
	kp[0] = h * dp0;
	kr[1] = h * (p.vr + kv[0] / 4.0);
	kv[1] = h * dv1;
	kp[1] = 0.0;
	if(dms_watch_moved(watch, watch_for))
		return false;

	const double r2 = p.radius + 3.0 / 32.0 * kr[0] + 9.0 / 32.0 * kr[1];
	kr[2] = h * (p.vr + 3.0 / 32.0 * kv[0] + 9.0 / 32.0 * kv[1]);
	const double r3 = p.radius + 1932.0 / 2197.0 * kr[0] - 7200.0 / 2197.0 * kr[1] + 7296.0 / 2197.0 * kr[2];
	double dv2, dp2, dv3, dp3;
	dms_stage_pair(stage, r2, r3, second, dv2, dp2, dv3, dp3);
	kv[2] = h * dv2;
	kp[2] = h * dp2;
	kv[3] = h * dv3;
	kp[3] = h * dp3;
	if(dms_watch_moved(watch, watch_for))
		return false;

	kr[3] = h * (p.vr + 1932.0 / 2197.0 * kv[0] - 7200.0 / 2197.0 * kv[1] + 7296.0 / 2197.0 * kv[2]);
	const double r4 = p.radius + 439.0 / 216.0 * kr[0] - 8.0 * kr[1] + 3680.0 / 513.0 * kr[2] - 845.0 / 4104.0 * kr[3];
	kr[4] = h * (p.vr + 439.0 / 216.0 * kv[0] - 8.0 * kv[1] + 3680.0 / 513.0 * kv[2] - 845.0 / 4104.0 * kv[3]);
	const double r5 = p.radius - 8.0 / 27.0 * kr[0] + 2.0 * kr[1] - 3544.0 / 2565.0 * kr[2] + 1859.0 / 4104.0 * kr[3] - 11.0 / 40.0 * kr[4];
	double dv4, dp4, dv5, dp5;
	dms_stage_pair(stage, r4, r5, second, dv4, dp4, dv5, dp5);
	kv[4] = h * dv4;
	kp[4] = h * dp4;
	kr[5] = h * (p.vr - 8.0 / 27.0 * kv[0] + 2.0 * kv[1] - 3544.0 / 2565.0 * kv[2] + 1859.0 / 4104.0 * kv[3] - 11.0 / 40.0 * kv[4]);
	kv[5] = h * dv5;
	kp[5] = h * dp5;
	return true;
}

/* One attempt of Runge_Kutta_45_Step (:1487-1602).  Returns 1 accepted,
 * 0 forced-accepted (retry cap or minimum step: the step "fails"), -1 retry.
 * PAIR: the stages as pairs (dms_rk45_stages_paired; on the device lanes 0
 * and 1 of the wave run the attempt together, second = lane 1), abandoned
 * with p unchanged (returns -2) when *watch moves off watch_for. */
template <bool PAIR = false>
DMS_HD int dms_rk45_attempt(DmsProp& p, const DmsSolarView& S, int inner_iter, bool second = false,
                            const unsigned int* watch = nullptr, unsigned int watch_for = 0u)
{
	const double tol0 = 1.0 * DMS_km, tol1 = 1.0e-3 * DMS_km / DMS_sec, tol2 = 1.0e-7;
	const double h = p.ts;
	double kr[6], kv[6], kp[6], ri;
	const double aJ = dm_fabs(p.J);
	const bool j_ok = (p.J == 0.0) | dms_in_binades(aJ, -99, 166); /* inside [1e-30, 1e50] */
	if constexpr(PAIR)
	{
		(void) ri;
		if(!dms_rk45_stages_paired(
		       p, h, [&](double r, double& dv, double& dph) { dms_stage(p, S, r, j_ok, true, dv, dph); }, second, kr, kv, kp,
		       watch, watch_for))
			return -2;
	}
	else
	{
		(void) second;
		(void) watch;
		(void) watch_for;
		double dv, dph = 0.0;
		kr[0] = h * p.vr;
		dms_stage(p, S, p.radius, j_ok, true, dv, dph);
		kv[0] = h * dv;
		kp[0] = h * dph;

		ri = p.radius + kr[0] / 4.0;
		kr[1] = h * (p.vr + kv[0] / 4.0);
		dms_stage(p, S, ri, j_ok, false, dv, dph);
		kv[1] = h * dv;

		ri = p.radius + 3.0 / 32.0 * kr[0] + 9.0 / 32.0 * kr[1];
		kr[2] = h * (p.vr + 3.0 / 32.0 * kv[0] + 9.0 / 32.0 * kv[1]);
		dms_stage(p, S, ri, j_ok, true, dv, dph);
		kv[2] = h * dv;
		kp[2] = h * dph;

		ri = p.radius + 1932.0 / 2197.0 * kr[0] - 7200.0 / 2197.0 * kr[1] + 7296.0 / 2197.0 * kr[2];
		kr[3] = h * (p.vr + 1932.0 / 2197.0 * kv[0] - 7200.0 / 2197.0 * kv[1] + 7296.0 / 2197.0 * kv[2]);
		dms_stage(p, S, ri, j_ok, true, dv, dph);
		kv[3] = h * dv;
		kp[3] = h * dph;

		ri = p.radius + 439.0 / 216.0 * kr[0] - 8.0 * kr[1] + 3680.0 / 513.0 * kr[2] - 845.0 / 4104.0 * kr[3];
		kr[4] = h * (p.vr + 439.0 / 216.0 * kv[0] - 8.0 * kv[1] + 3680.0 / 513.0 * kv[2] - 845.0 / 4104.0 * kv[3]);
		dms_stage(p, S, ri, j_ok, true, dv, dph);
		kv[4] = h * dv;
		kp[4] = h * dph;

		ri = p.radius - 8.0 / 27.0 * kr[0] + 2.0 * kr[1] - 3544.0 / 2565.0 * kr[2] + 1859.0 / 4104.0 * kr[3] - 11.0 / 40.0 * kr[4];
		kr[5] = h * (p.vr - 8.0 / 27.0 * kv[0] + 2.0 * kv[1] - 3544.0 / 2565.0 * kv[2] + 1859.0 / 4104.0 * kv[3] - 11.0 / 40.0 * kv[4]);
		dms_stage(p, S, ri, j_ok, true, dv, dph);
		kv[5] = h * dv;
		kp[5] = h * dph;
	}

	const double r4 = p.radius + 25.0 / 216.0 * kr[0] + 1408.0 / 2565.0 * kr[2] + 2197.0 / 4104.0 * kr[3] - 1.0 / 5.0 * kr[4];
	const double v4 = p.vr + 25.0 / 216.0 * kv[0] + 1408.0 / 2565.0 * kv[2] + 2197.0 / 4104.0 * kv[3] - 1.0 / 5.0 * kv[4];
	const double p4 = p.phi + 25.0 / 216.0 * kp[0] + 1408.0 / 2565.0 * kp[2] + 2197.0 / 4104.0 * kp[3] - 1.0 / 5.0 * kp[4];
	const double r5 = p.radius + 16.0 / 135.0 * kr[0] + 6656.0 / 12825.0 * kr[2] + 28561.0 / 56430.0 * kr[3] - 9.0 / 50.0 * kr[4] + 2.0 / 55.0 * kr[5];
	const double v5 = p.vr + 16.0 / 135.0 * kv[0] + 6656.0 / 12825.0 * kv[2] + 28561.0 / 56430.0 * kv[3] - 9.0 / 50.0 * kv[4] + 2.0 / 55.0 * kv[5];
	const double p5 = p.phi + 16.0 / 135.0 * kp[0] + 6656.0 / 12825.0 * kp[2] + 28561.0 / 56430.0 * kp[3] - 9.0 / 50.0 * kp[4] + 2.0 / 55.0 * kp[5];
	const double e0 = dm_fabs(r5 - r4), e1 = dm_fabs(v5 - v4), e2 = dm_fabs(p5 - p4);
	/* RK45_Next_Step_Size, :477-502, and the acceptance, written with selects
	 * (both arms are plain values; only the forced acceptance branches) */
	const bool finite_err = dm_isfinite(e0) & dm_isfinite(e1) & dm_isfinite(e2);
	const double ee[3] = {e0, e1, e2};
	const double tt[3] = {tol0, tol1, tol2};
	/* The reference takes the minimum of 0.84 (tol_i / e_i)^(1/4) over the
	 * components with e_i > 0, skipping non-finite candidates.  Rounded
	 * square roots and the product by 0.84 are monotone, so that minimum
	 * is the candidate of the smallest quotient (in (0, inf], finite
	 * exactly when its candidate is): one candidate instead of three. */
	double q = 0.0;
	bool any = false;
	for(int i = 0; i < 3; i++)
	{
		const bool use = !(ee[i] <= 0.0);
		const double qi = tt[i] / ee[i];
		q = use ? (any ? dm_min(q, qi) : qi) : q;
		any = any | use;
	}
	const double candidate = 0.84 * dm_pow_quarter(q);
	double factor = (any & dm_isfinite(candidate)) ? dm_min(DMS_RK45_MAX_STEP_FACTOR, candidate) : DMS_RK45_MAX_STEP_FACTOR;
	factor = dm_max(DMS_RK45_MIN_STEP_FACTOR, dm_min(factor, DMS_RK45_MAX_STEP_FACTOR));
	const double h_new = finite_err ? dms_rk45_sanitized(factor * h) : DMS_RK45_MIN_STEP_FACTOR * h;
	/* RK45_Errors_Within_Tolerance, :466-475 */
	const bool within = finite_err & !(e0 > tol0) & !(e1 > tol1) & !(e2 > tol2);
	const double abs_min_step = 1.0e-8 * DMS_sec;
	if(__builtin_expect(!within & ((h_new < abs_min_step) | (inner_iter >= DMS_RK45_MAX_INNER_RETRIES)), 0))
	{
		const bool state_ok = dm_isfinite(r4) && dm_isfinite(v4) && dm_isfinite(p4);
		p.time = p.time + abs_min_step;
		if(state_ok)
		{
			p.radius = r4;
			if(p.radius < 0.0)
				p.radius = 0.0;
			p.vr = v4;
			p.phi = p4;
		}
		p.ts = abs_min_step;
		return 0;
	}
	p.time = within ? p.time + h : p.time;
	p.radius = within ? ((r4 < 0.0) ? 0.0 : r4) : p.radius;
	p.vr = within ? v4 : p.vr;
	p.phi = within ? p4 : p.phi;
	p.ts = h_new;
	return within ? 1 : -1;
}

/* ---- helpers of Propagate_Freely, :53-464 ------------------------------------ */
DMS_HD double dms_free_propagation_cap(double radius, double speed, double maximum_distance)
{
	double cap = 1.0e6 * DMS_sec;
	const double safe_speed = dm_max(dm_fabs(speed), 1.0e-12 * DMS_km / DMS_sec);
	const double crossing_scale = dm_max(0.25 * maximum_distance, 10.0 * DMS_km);
	cap = dm_min(cap, crossing_scale / safe_speed);
	const double safe_radius = dm_max(radius, 1.0 * DMS_km);
	const double dynamical_time = dm_sqrt(safe_radius * safe_radius * safe_radius / (DMS_G_Newton * DMS_mSun));
	if(dm_isfinite(dynamical_time) && dynamical_time > 0.0)
		cap = dm_min(cap, 0.1 * dynamical_time);
	return dm_max(cap, 1.0e-8 * DMS_sec);
}

DMS_HD bool dms_outward_escaping(const DmsSolarView& S, const double* r, const double* v, double boundary)
{
	const double radius = dms_norm(r);
	if(radius < boundary)
		return false;
	const double vr = (radius > 0.0) ? dms_dot(r, v) / radius : 0.0;
	return vr > 0.0 && dms_norm(v) > dms_local_escape_speed(S, radius);
}

/* Unit_Vector_At_Angle_From_Axis, :109-128 (false where the reference throws) */
DMS_HD bool dms_unit_vector_at_angle(double cos_theta, double phi, const double* axis, double* out)
{
	if(!dm_isfinite(cos_theta) || cos_theta < -1.0 - 1.0e-12 || cos_theta > 1.0 + 1.0e-12 || !dm_isfinite(phi))
		return false;
	const double axis_norm = dms_norm(axis);
	if(!dm_isfinite(axis_norm) || axis_norm <= 0.0)
		return false;
	const double e3[3] = {axis[0] / axis_norm, axis[1] / axis_norm, axis[2] / axis_norm};
	double e1[3], e2[3];
	dms_frame(e3, e1, e2);
	cos_theta = dm_clamp_cos(cos_theta);
	const double sin_theta = dm_sqrt(dm_max(0.0, 1.0 - cos_theta * cos_theta));
	const double cp = dm_cos(phi), sp = dm_sin(phi);
	for(int c = 0; c < 3; c++)
		out[c] = cos_theta * e3[c] + sin_theta * cp * e1[c] + sin_theta * sp * e2[c];
	return true;
}

/* Bound_Kepler_Return_At_Same_Radius, :152-201 (times are segment-local) */
DMS_HD bool dms_bound_kepler_return(double ot, const double* orr, const double* ov, double& it, double* ir, double* iv)
{
	const double radius = dms_norm(orr);
	const double speed = dms_norm(ov);
	const double vr = (radius <= 0.0) ? 0.0 : dms_dot(orr, ov) / radius;
	if(!dm_isfinite(radius) || !dm_isfinite(speed) || !dm_isfinite(vr) || radius <= DMS_rSun || speed <= 0.0 || vr <= 0.0)
		return false;
	const double mu = DMS_G_Newton * DMS_mSun;
	const double specific_energy = 0.5 * speed * speed - mu / radius;
	if(!dm_isfinite(specific_energy) || specific_energy >= 0.0)
		return false;
	double L[3];
	dms_cross(orr, ov, L);
	const double J = dms_norm(L);
	if(!dm_isfinite(J))
		return false;
	const double a = -mu / (2.0 * specific_energy);
	if(!dm_isfinite(a) || a <= 0.0)
		return false;
	double e2 = 1.0 + 2.0 * specific_energy * J * J / (mu * mu);
	if(!dm_isfinite(e2) || e2 < 0.0)
		return false;
	e2 = dm_max(0.0, e2);
	const double ecc = dm_sqrt(e2);
	if(ecc <= 1.0e-14 || ecc >= 1.0)
		return false;
	const double cosE = dm_clamp_cos((1.0 - radius / a) / ecc);
	const double E = dm_acos(cosE);
	const double sinE = dm_sqrt(dm_max(0.0, 1.0 - cosE * cosE));
	const double n = dm_sqrt(mu / (a * a * a));
	if(!dm_isfinite(n) || n <= 0.0)
		return false;
	const double return_time = (2.0 * DM_PI - 2.0 * E + 2.0 * ecc * sinE) / n;
	if(!dm_isfinite(return_time) || return_time <= 0.0)
		return false;
	it = ot + return_time;
	for(int c = 0; c < 3; c++)
	{
		ir[c] = orr[c];
		const double ru = orr[c] / radius;
		iv[c] = ov[c] - 2.0 * vr * ru;
	}
	return it > ot && dm_isfinite(dms_norm(iv));
}

DMS_HD void dms_interpolate_event(double bt, const double* br, const double* bv, double at, const double* ar,
                                  const double* av, double f, double& t, double* r, double* v)
{
	f = dm_clamp_unit(f);
	for(int c = 0; c < 3; c++)
	{
		r[c] = br[c] + f * (ar[c] - br[c]);
		v[c] = bv[c] + f * (av[c] - bv[c]);
	}
	t = bt + f * (at - bt);
}

DMS_HD double dms_radius_at_fraction(const double* br, const double* ar, double f)
{
	f = dm_clamp_unit(f);
	double rs = 0.0;
	for(int c = 0; c < 3; c++)
	{
		const double q = br[c] + f * (ar[c] - br[c]);
		rs += q * q;
	}
	return dm_sqrt(dm_max(0.0, rs));
}

/* Solar_Interior_Fraction_Interval with Surface_Crossing_Fractions and
 * Find_Surface_Crossing_Fraction, :231-325 */
DMS_HD bool dms_interior_interval(const double* br, const double* ar, double r_before, double r_after, double& start, double& end)
{
	const bool before_inside = r_before < DMS_rSun;
	const bool after_inside = r_after < DMS_rSun;
	if(before_inside & after_inside)
	{
		/* the chord's crossings are not used in this case */
		start = 0.0;
		end = 1.0;
		return true;
	}
	double first = 0.0, second = 0.0;
	bool has = false;
	{
		double qa = 0.0, qb = 0.0, qc = -DMS_rSun * DMS_rSun;
		for(int c = 0; c < 3; c++)
		{
			const double d = ar[c] - br[c];
			qa += d * d;
			qb += 2.0 * br[c] * d;
			qc += br[c] * br[c];
		}
		if(qa > 0.0)
		{
			const double disc = qb * qb - 4.0 * qa * qc;
			if(dm_isfinite(disc) & !(disc < 0.0))
			{
				const double sd = dm_sqrt(dm_max(0.0, disc));
				first = (-qb - sd) / (2.0 * qa);
				second = (-qb + sd) / (2.0 * qa);
				if(first > second)
				{
					const double t = first;
					first = second;
					second = t;
				}
				has = true;
			}
		}
	}
	const double tol = 1.0e-12;
	const bool f_in = (first >= -tol) & (first <= 1.0 + tol);
	const bool s_in = (second >= -tol) & (second <= 1.0 + tol);
	bool need_bisection = false;
	if(!before_inside && after_inside)
	{
		if(has && f_in)
			start = dm_clamp_unit(first);
		else if(has && s_in)
			start = dm_clamp_unit(second);
		else
			need_bisection = true;
		end = 1.0;
	}
	else if(before_inside && !after_inside)
	{
		start = 0.0;
		if(has && s_in)
			end = dm_clamp_unit(second);
		else if(has && f_in)
			end = dm_clamp_unit(first);
		else
			need_bisection = true;
	}
	else
	{
		if(has && f_in && s_in)
		{
			start = dm_clamp_unit(first);
			end = dm_clamp_unit(second);
			return end > start;
		}
		return false;
	}
	if(need_bisection)
	{
		double lower = 0.0, upper = 1.0;
		double r_lower = dms_radius_at_fraction(br, ar, lower) - DMS_rSun;
		for(int it = 0; it < 60; it++)
		{
			const double mid = 0.5 * (lower + upper);
			const double r_mid = dms_radius_at_fraction(br, ar, mid) - DMS_rSun;
			if((r_lower <= 0.0 && r_mid <= 0.0) || (r_lower >= 0.0 && r_mid >= 0.0))
			{
				lower = mid;
				r_lower = r_mid;
			}
			else
				upper = mid;
		}
		const double f = dm_clamp_unit(0.5 * (lower + upper));
		if(!before_inside)
			start = f;
		else
			end = f;
	}
	return end > start;
}

/* Scattering_Rate_At_Fraction, :327-345 */
DMS_HD double dms_rate_at_fraction(const DmsPointView& P, const DmsSolarView& S, const double* br, const double* bv,
                                   const double* ar, const double* av, double f)
{
	f = dm_clamp_unit(f);
	double rs = 0.0, vs = 0.0;
	for(int c = 0; c < 3; c++)
	{
		const double pos = br[c] + f * (ar[c] - br[c]);
		const double vel = bv[c] + f * (av[c] - bv[c]);
		rs += pos * pos;
		vs += vel * vel;
	}
	const double radius = dm_sqrt(dm_max(0.0, rs));
	if(radius >= DMS_rSun)
		return 0.0;
	return dms_rate_total(P, S, radius, dm_sqrt(dm_max(0.0, vs)));
}

/* One element of dms_rates_at_fractions up to its table read: radius and
 * speed (at fraction f of the step, or (xr, xv) for the end-point query), the
 * lookup cell, and which evaluation it needs (grid: the table; comp: the
 * direct 64-target sum); res holds the value of an element that needs
 * neither.  The arithmetic is that of dms_rate_at_fraction / dms_rate_total. */
struct DmsRateSlot
{
	double res, rad, spd, rf, vf;
	unsigned lo;
	bool grid, comp;
};

DMS_HD DmsRateSlot dms_rate_slot(const DmsPointView& P, bool use_grid, bool active, bool at_fraction, const double* br,
                                 const double* bv, const double* ar, const double* av, double f, double xr, double xv)
{
	DmsRateSlot q;
	double r, v;
	bool outside = false; /* Scattering_Rate_At_Fraction: radius >= R_sun gives 0 */
	if(!at_fraction)
	{
		r = xr;
		v = xv;
	}
	else
	{
		f = dm_clamp_unit(f);
		double rs = 0.0, vs = 0.0;
		for(int c = 0; c < 3; c++)
		{
			const double pos = br[c] + f * (ar[c] - br[c]);
			const double vel = bv[c] + f * (av[c] - bv[c]);
			rs += pos * pos;
			vs += vel * vel;
		}
		r = dm_sqrt(dm_max(0.0, rs));
		v = dm_sqrt(dm_max(0.0, vs));
		outside = r >= DMS_rSun;
	}
	/* Total_DM_Scattering_Rate */
	/* conditions combined with & and | (no short-circuit): selects, not branches */
	const bool finite = dm_isfinite(r) & dm_isfinite(v);
	r = (r < 0.0) ? 0.0 : r;
	v = (v < 0.0) ? 0.0 : v;
	q.rad = r;
	q.spd = v;
	const bool table = use_grid & (v <= P.vmax);
	/* Total_DM_Scattering_Rate_Interpolated */
	const bool zero_interp = table & ((r > DMS_rSun) | (r < 0.0) | (v < 0.0));
	const double rc = r * P.inv_dr;
	const double vc = v * P.inv_dv;
	const bool coords_ok = dm_isfinite(rc) & dm_isfinite(vc) & !(rc < 0.0) & !(vc < 0.0);
	const bool live = active & !outside & finite;
	q.grid = live & table & !zero_interp & coords_ok;
	q.comp = live & !zero_interp & !(table & coords_ok);
	const double rcs = coords_ok ? dm_min(rc, (double) (P.Nr - 2)) : 0.0;
	const double vcs = coords_ok ? dm_min(vc, (double) (P.Nv - 2)) : 0.0;
	const unsigned ri = (unsigned) rcs;
	const unsigned vi = (unsigned) vcs;
	q.rf = rc - ri;
	q.vf = vc - vi;
	q.lo = q.grid ? ri * (unsigned) P.Nv + vi : 0u;
	q.res = (!finite && !outside) ? dm_nan() : 0.0;
	return q;
}

/* The element's value from the four table nodes of its cell (lookup slots)
 * or the value it already holds (the others). */
DMS_HD double dms_rate_slot_bilinear(const DmsRateSlot& q, const double* g)
{
	const double lower = g[0] + q.vf * (g[1] - g[0]);
	const double upper = g[2] + q.vf * (g[3] - g[2]);
	const double val = lower + q.rf * (upper - lower);
	const double gv = dm_isfinite(val) ? dm_max(0.0, val) : dm_nan();
	return q.grid ? gv : q.res;
}

/* Scattering_Rate_At_Fraction (:327-345) for n fractions of one step, plus
 * optionally Total_DM_Scattering_Rate (src/Solar_Model.cpp:305-322) at one more
 * (xr, xv) -- the rate the next step's preparation needs at the end point.
 * Per element the arithmetic is that of dms_rate_at_fraction / dms_rate_total.
 * fr and out hold DMS_MAX_OPTICAL_DEPTH_PIECES + 1 elements (fr all defined). */
DMS_HD void dms_rates_at_fractions(const DmsPointView& P, const DmsSolarView& S, const double* br, const double* bv,
                                   const double* ar, const double* av, const double* fr, int n, bool use_cached,
                                   double cached_rate, double* out, bool extra, double xr, double xv, double* xout)
{
	/* Elements 1..n-1 and the end-point query are written without branches
	 * (selects only), their table reads issued together, so that their
	 * dependent chains interleave in one instruction stream.  Element 0, the
	 * step's start point, usually has the rate the step preparation used
	 * (use_cached) and is evaluated only when it has not.  Elements that need
	 * the direct 64-target sum (no table, v > v_max, non-finite table
	 * coordinates) take a separate branch afterwards. */
	constexpr int NQ = DMS_MAX_OPTICAL_DEPTH_PIECES + 2; /* fractions, then the extra query */
	const bool use_grid = P.use_grid != 0;
	const unsigned nv = (unsigned) P.Nv;
	DmsRateSlot q[NQ];
#pragma unroll
	for(int i = 1; i < NQ; i++)
	{
		const bool is_extra = (i == NQ - 1);
		q[i] = dms_rate_slot(P, use_grid, is_extra ? extra : (i < n), !is_extra, br, bv, ar, av, fr[is_extra ? 0 : i], xr, xv);
	}
	if(use_grid)
	{
		double g[NQ][4];
#pragma unroll
		for(int i = 1; i < NQ; i++)
		{
			const double* t = P.grid + q[i].lo;
			g[i][0] = t[0];
			g[i][1] = t[1];
			g[i][2] = t[nv];
			g[i][3] = t[nv + 1];
		}
#pragma unroll
		for(int i = 1; i < NQ; i++)
			q[i].res = dms_rate_slot_bilinear(q[i], g[i]);
	}
	if(use_cached)
	{
		q[0].res = cached_rate;
		q[0].comp = false;
	}
	else
	{
		q[0] = dms_rate_slot(P, use_grid, 0 < n, true, br, bv, ar, av, fr[0], xr, xv);
		if(use_grid)
		{
			const double* t = P.grid + q[0].lo;
			const double g0[4] = {t[0], t[1], t[nv], t[nv + 1]};
			q[0].res = dms_rate_slot_bilinear(q[0], g0);
		}
	}
	bool any_comp = false;
#pragma unroll
	for(int i = 0; i < NQ; i++)
		any_comp = any_comp || q[i].comp;
	if(__builtin_expect(any_comp, 0)) /* no table, or v above it: rare */
	{
#pragma unroll
		for(int i = 0; i < NQ; i++)
			if(q[i].comp)
				q[i].res = dms_rate_computed(P, S, q[i].rad, q[i].spd);
	}
	/* out has NQ - 1 elements; those from n on are written but not read */
#pragma unroll
	for(int i = 0; i < NQ - 1; i++)
		out[i] = q[i].res;
	if(extra)
		*xout = q[NQ - 1].res;
}

/* Element k of dms_rates_at_fractions (k < DMS_MAX_OPTICAL_DEPTH_PIECES + 1:
 * fraction fr[k]; k = DMS_MAX_OPTICAL_DEPTH_PIECES + 1: the end-point query)
 * evaluated on its own: the value that function leaves in out[k] (in *xout for
 * the end-point query), by the same slot, table read and direct-sum steps. */
DMS_HD double dms_rate_fraction_slot(const DmsPointView& P, const DmsSolarView& S, int k, const double* br, const double* bv,
                                     const double* ar, const double* av, const double* fr, int n, bool use_cached,
                                     double cached_rate, bool extra, double xr, double xv)
{
	constexpr int NQ = DMS_MAX_OPTICAL_DEPTH_PIECES + 2;
	if((k == 0) & use_cached)
		return cached_rate;
	const bool use_grid = P.use_grid != 0;
	const bool is_extra = (k == NQ - 1);
	DmsRateSlot q = dms_rate_slot(P, use_grid, is_extra ? extra : (k < n), !is_extra, br, bv, ar, av, fr[is_extra ? 0 : k], xr, xv);
	if(use_grid)
	{
		const unsigned nv = (unsigned) P.Nv;
		const double* t = P.grid + q.lo;
		const double g[4] = {t[0], t[1], t[nv], t[nv + 1]};
		q.res = dms_rate_slot_bilinear(q, g);
	}
	if(__builtin_expect(q.comp, 0))
		q.res = dms_rate_computed(P, S, q.rad, q.spd);
	return q.res;
}

/* Solve_Linear_Rate_Optical_Depth_Fraction, :416-464 */
DMS_HD double dms_solve_linear_rate(double rs, double re, double dt, double target_tau)
{
	if(target_tau <= 0.0)
		return 0.0;
	if(!dm_isfinite(target_tau) || dt <= 0.0)
		return 1.0;
	if(!dm_isfinite(rs) || !dm_isfinite(re) || rs < 0.0 || re < 0.0)
		return 1.0;
	const double total = 0.5 * (rs + re) * dt;
	if(!dm_isfinite(total) || total <= 0.0)
		return 1.0;
	if(target_tau >= total)
		return 1.0;
	const double target = target_tau / dt;
	const double qa = 0.5 * (re - rs);
	const double qb = rs;
	if(dm_fabs(qa) < 1.0e-300)
		return (qb > 0.0) ? dm_clamp_unit(target / qb) : 1.0;
	const double disc = qb * qb + 4.0 * qa * target;
	if(dm_isfinite(disc) && disc >= 0.0)
	{
		const double sd = dm_sqrt(disc);
		const double r0 = (-qb + sd) / (2.0 * qa);
		if(dm_isfinite(r0) && r0 >= 0.0 && r0 <= 1.0)
			return dm_clamp_unit(r0);
		const double r1 = (-qb - sd) / (2.0 * qa);
		if(dm_isfinite(r1) && r1 >= 0.0 && r1 <= 1.0)
			return dm_clamp_unit(r1);
	}
	double low = 0.0, high = 1.0;
	for(int it = 0; it < 30; it++)
	{
		const double mid = 0.5 * (low + high);
		const double tau_mid = dt * (rs * mid + 0.5 * (re - rs) * mid * mid);
		if(tau_mid < target_tau)
			low = mid;
		else
			high = mid;
	}
	return dm_clamp_unit(0.5 * (low + high));
}

/* ---- trajectory bookkeeping --------------------------------------------------- */
DMS_HD double dms_capture_energy_eV(const DmsSolarView& S, double dm_mass, double radius, double speed)
{
	const double vesc = dms_local_escape_speed(S, radius);
	const double E = 0.5 * dm_mass * (speed * speed - vesc * vesc);
	return dm_div_const(E, DMS_eV, 1.0 / DMS_eV);
}

/* true where the reference's abort_if_uncaptured_bound lambda fires (:767-778) */
DMS_HD bool dms_uncaptured_bound(const DmsLane& L, const DmsSolarView& S, double dm_mass, const double* r, const double* v)
{
	if(L.captured)
		return false;
	const double E = dms_capture_energy_eV(S, dm_mass, dms_norm(r), dms_norm(v));
	return !(dm_isfinite(E) && E >= 0.0);
}

/* Update_Capture_State, :689-760 */
DMS_HD int dms_update_capture_state(DmsLane& L, const DmsSolarView& S, double dm_mass, int terminate_on_capture,
                                    double radius, double speed, double time, int allow_new)
{
	L.last_E = dm_nan();
	if((terminate_on_capture != 0) & !allow_new & !L.captured)
		return 0;
	const double E = dms_capture_energy_eV(S, dm_mass, radius, speed);
	L.last_E = E;
	double dE = dm_nan();
	if(dm_isfinite(L.prev_cap_E))
		dE = E - L.prev_cap_E;
	L.prev_cap_E = E;
	const int negative = E < 0.0;
	const double t_now = dm_div_const(time, DMS_sec, 1.0 / DMS_sec);
	if(!allow_new)
	{
		if((L.captured != 0) & dm_isfinite(L.ff_ref_E))
		{
			const double drift = dm_fabs(E - L.ff_ref_E);
			L.drift_eV = (dm_isfinite(drift) & (drift > L.drift_eV)) ? drift : L.drift_eV;
			const double scale = dm_max(dm_fabs(L.ff_ref_E), DMS_FREE_ENERGY_DRIFT_REL_E_SCALE_EV);
			const double rel = drift / scale;
			L.drift_rel = (dm_isfinite(rel) & (rel > L.drift_rel)) ? rel : L.drift_rel;
		}
		if((L.bound != 0) & (L.captured != 0)) /* a store, not a select: the field lives in scratch */
			L.t_last_bound = t_now;
		return 0;
	}
	const int was_bound = L.bound;
	L.bound = negative;
	L.ff_ref_E = E;
	if(negative)
	{
		if(!L.captured)
		{
			L.captured = 1;
			L.t_capture = t_now;
			L.t_last_bound = t_now;
			L.r_first_neg = radius / DMS_km;
			L.E_first_neg = E;
			L.dE_first_neg = dE;
		}
		else if(!was_bound)
			L.t_unbind = dm_nan();
		L.t_last_bound = t_now;
		return 1;
	}
	if(was_bound && L.captured)
		L.t_unbind = t_now;
	return 0;
}

/* Accumulate_Bincount_Step (:640-669) into the trajectory's row, keeping the
 * bin being filled in registers; the per-bin addition sequence is unchanged. */
DMS_HD void dms_bin_flush(DmsLane& L, const DmsRows& R, int nbins)
{
	if(L.cur_bin < 0)
		return;
	if(R.per_lane)
	{
		DmsBinAcc* b = R.bins + (size_t) L.row * (size_t) nbins + L.cur_bin;
		b->dt = L.acc_dt;
		b->v2dt = L.acc_v2;
		b->cnt = L.acc_cnt;
		L.cur_bin = -1;
		return;
	}
	if(R.rows)
	{
		double* row = R.rows + (size_t) L.row * 2 * (size_t) nbins;
		row[L.cur_bin] = L.acc_dt;
		row[nbins + L.cur_bin] = L.acc_v2;
	}
	if(R.counts)
		R.counts[(size_t) L.row * (size_t) nbins + L.cur_bin] = L.acc_cnt;
	L.cur_bin = -1;
}

/* Accumulate_Bincount_Step's bin of a radius in km: -1 outside [0, 2 R_sun)
 * or past the last bin. */
DMS_HD int dms_bin_index(const DmsJobView& J, double r_km)
{
	const double bin_max = 2.0 * DMS_R_SUN_KM;
	int bin = -1;
	if(dm_isfinite(r_km) & (r_km >= 0.0) & (r_km < bin_max))
	{
		bin = (int) (r_km / J.bin_width_km);
		if(bin < 0)
			bin = 0;
		if(bin >= J.nbins)
			bin = -1;
	}
	return bin;
}

/* The bincount anchor (prev_r_km) moved: the next accumulation falls into its
 * bin, whose stored values are read now, a step before they are used, so that
 * the load is not waited for in the step that switches bins.  They are read
 * whatever the bin (bin 0 for none): when it is the bin in registers they are
 * not used.  No store to that bin's record can come between this read and the
 * switch: a record is written only when the lane leaves its bin. */
DMS_HD void dms_bin_anchor(DmsLane& L, const DmsRows& R, const DmsJobView& J)
{
	const int bin = dms_bin_index(J, L.prev_r_km);
	L.pf_bin = bin;
	const size_t b = (size_t) (bin < 0 ? 0 : bin);
	if(R.per_lane)
	{
		const DmsBinAcc* rec = R.bins + (size_t) L.row * (size_t) J.nbins + b;
		L.pf_dt = rec->dt;
		L.pf_v2 = rec->v2dt;
		L.pf_cnt = rec->cnt;
		return;
	}
	if(R.rows)
	{
		const double* row = R.rows + (size_t) L.row * 2 * (size_t) J.nbins;
		L.pf_dt = row[b];
		L.pf_v2 = row[J.nbins + b];
	}
	else
	{
		L.pf_dt = 0.0;
		L.pf_v2 = 0.0;
	}
	L.pf_cnt = R.counts ? R.counts[(size_t) L.row * (size_t) J.nbins + b] : 0u;
}

/* One step of the bincount into the bin of prev_r_km (dms_bin_anchor read it
 * and its stored values when the anchor was set). */
DMS_HD void dms_accumulate_bincount(DmsLane& L, const DmsRows& R, const DmsJobView& J, double v2, double dt_sec)
{
	if(!dm_isfinite(dt_sec) | (dt_sec <= 0.0))
		return;
	const int bin = L.pf_bin;
	if(bin < 0)
		return;
	if(bin != L.cur_bin)
	{
		dms_bin_flush(L, R, J.nbins);
		L.cur_bin = bin;
		L.acc_dt = L.pf_dt;
		L.acc_v2 = L.pf_v2;
		L.acc_cnt = L.pf_cnt;
		if(R.per_lane && L.acc_cnt == 0u)
			R.touched[(size_t) L.row * (size_t) J.nbins + L.n_touched++] = (uint16_t) bin;
	}
	const double v2dt = v2 * dt_sec;
	L.acc_dt += dt_sec;
	L.acc_v2 += v2dt;
	L.acc_cnt += 1u;
}

/* commit_accepted_event lambda, :842-874 (segment-local event) */
DMS_HD int dms_commit(DmsLane& L, const DmsRows& R, const DmsJobView& J, const DmsSolarView& S, double dm_mass,
                      double lt, const double* lr, const double* lv)
{
	const double abs_t = lt + L.time_origin;
	const double t_now = dm_div_const(lt, DMS_sec, 1.0 / DMS_sec);
	if(!J.terminate_on_capture)
	{
		const double r_now = dm_div_const(dms_norm(lr), DMS_km, 1.0 / DMS_km);
		const double v_now = dm_div_const(dms_norm(lv), DMS_km / DMS_sec, 1.0 / (DMS_km / DMS_sec));
		const double v2 = v_now * v_now;
		if(L.prev_time_sec >= 0.0)
		{
			const double dts = t_now - L.prev_time_sec;
			dms_accumulate_bincount(L, R, J, L.prev_v2, dts);
		}
		L.prev_time_sec = t_now;
		L.prev_r_km = r_now;
		L.prev_v2 = v2;
		dms_bin_anchor(L, R, J);
	}
	return dms_update_capture_state(L, S, dm_mass, J.terminate_on_capture, dms_norm(lr), dms_norm(lv), abs_t, 0);
}

/* Propagate_Freely entry (:764-805).  Returns -1 to continue with PREP or the
 * termination reason the reference returns before the step loop. */
#include "dms_engine_mixed.h"

DMS_HD int dms_segment_start(DmsLane& L, const DmsSolarView& S, double dm_mass, const DmsJobView& J)
{
	if(dms_outward_escaping(S, L.ev_r, L.ev_v, J.max_distance))
		return DMS_OUTWARD_ESCAPE;
	if(dms_uncaptured_bound(L, S, dm_mass, L.ev_r, L.ev_v))
		return DMS_NUMERICAL_FAILURE;
	L.time_origin = L.ev_t;
	dms_prop_init(L.prop, 0.0, L.ev_r, L.ev_v);
	L.eb_valid = 0;
	L.mlx = -dm_log(dm_max(dm_uniform(&L.rng), DM_DBL_MIN));
	L.mlx_f = (float) L.mlx;
	L.mlx_c = 0.0f;
	L.time_steps = 0;
	L.od_retries = 0;
	if(!J.terminate_on_capture && L.prev_time_sec >= 0.0)
		L.prev_time_sec = 0.0;
	return -1;
}

/* Event_In_3D of a propagator state and the speed the step post takes with it
 * (dms_prop_event_speed, or dms_prop_event_mx in mixed precision). */
struct DmsStepEvent
{
	double t, r[3], v[3], speed;
};

template <bool MX = false>
DMS_HD void dms_step_event(const DmsProp& p, DmsStepEvent& ev)
{
	ev.speed = MX ? dms_prop_event_mx(p, ev.t, ev.r, ev.v) : dms_prop_event_speed(p, ev.t, ev.r, ev.v);
}

/* Top of the step loop (:876-918).  Returns -1 when an RKF45 step follows,
 * otherwise the termination reason. */
template <bool MX = false>
DMS_HD int dms_step_prep(DmsLane& L, const DmsPointView& P, const DmsSolarView& S, const DmsJobView& J)
{
	if(__builtin_expect(!(L.time_steps < J.max_time_steps), 0))
		return DMS_MAX_FREE_STEPS;
	if(__builtin_expect((J.rk_budget > 0) & (L.rk_all > J.rk_budget), 0))
	{
		double t;
		dms_prop_event(L.prop, t, L.ev_r, L.ev_v);
		L.ev_t = t + L.time_origin;
		return DMS_WALL_TIME_LIMIT;
	}
	if(__builtin_expect(!L.eb_valid, 0)) /* first step of a segment */
	{
		dms_prop_event(L.prop, L.eb_t, L.eb_r, L.eb_v);
		L.r_before = dms_prop_radius(L.prop);
		L.v_before = dms_prop_speed(L.prop);
		L.eb_valid = 1;
		L.rb_cached = 0;
	}
	const double cached_rate = L.rate_before;
	L.rate_before = 0.0;
	if constexpr(MX)
	{
		if(L.r_before >= DMS_rSun)
			L.rate_before_f = 0.0f;
	}
	if(L.r_before >= DMS_rSun)
	{
		const double cap = dms_free_propagation_cap(L.r_before, L.v_before, J.max_distance);
		L.prop.ts = dm_min(dms_rk45_sanitized(L.prop.ts), cap);
		const double rvb = (L.r_before > 0.0) ? dms_dot(L.eb_r, L.eb_v) / L.r_before : 0.0;
		if((L.r_before > DMS_rSun) & (rvb < 0.0))
		{
			const double est = (L.r_before - DMS_rSun) / (-rvb);
			if(dm_isfinite(est) & (est > 0.0))
				L.prop.ts = dm_min(L.prop.ts, dm_max(est, 1.0e-8 * DMS_sec));
		}
	}
	else
	{
		L.prop.ts = dms_rk45_sanitized(L.prop.ts);
		const double od_limit = J.terminate_on_capture ? DMS_CAPTURE_MODE_MAX_OPTICAL_DEPTH_STEP : DMS_MAX_OPTICAL_DEPTH_STEP;
		if constexpr(MX)
		{
			const float rb = L.rb_cached ? L.rate_before_f
			                             : dms_rate_total_mx(P, S, (float) (L.r_before * DMS_MX_INV_RSUN), (float) L.v_before);
			L.rate_before_f = rb;
			if(__builtin_expect(!mx_isfinite(rb) || rb < 0.0f, 0))
			{
				double t;
				dms_prop_event(L.prop, t, L.ev_r, L.ev_v);
				L.ev_t = t + L.time_origin;
				return DMS_NUMERICAL_FAILURE;
			}
			if(rb > 0.0f)
				L.prop.ts = dm_min(L.prop.ts, (double) ((float) od_limit / rb) * DMS_sec);
		}
		else
		{
			L.rate_before = L.rb_cached ? cached_rate : dms_rate_total(P, S, L.r_before, L.v_before);
			if(__builtin_expect(!dm_isfinite(L.rate_before) | (L.rate_before < 0.0), 0))
			{
				double t;
				dms_prop_event(L.prop, t, L.ev_r, L.ev_v);
				L.ev_t = t + L.time_origin;
				return DMS_NUMERICAL_FAILURE;
			}
			const double od_ts = od_limit / L.rate_before;
			L.prop.ts = (L.rate_before > 0.0) ? dm_min(L.prop.ts, od_ts) : L.prop.ts;
		}
	}
	L.pb_time = L.prop.time;
	L.pb_radius = L.prop.radius;
	L.pb_phi = L.prop.phi;
	L.pb_vr = L.prop.vr;
	L.pb_ts = L.prop.ts;
	/* Runge_Kutta_45_Step entry */
	L.prop.ts = dms_rk45_sanitized(L.prop.ts);
	L.inner_iter = 0;
	return -1;
}

/* Loop body after the RKF45 step (:921-1115).  Returns DMS_POST_NEXT when the
 * next step starts from the propagator state the attempt left (the common
 * exit), -2 to run PREP again from another state (optical-depth rejection,
 * Kepler return), -3 when the segment ended in a scattering (Scatter
 * follows), or the Propagate_Freely return value for every other exit. */
#define DMS_POST_NEXT (-4)
template <bool MX = false>
DMS_HD int dms_step_post(DmsLane& L, const DmsPointView& P, const DmsSolarView& S, const DmsJobView& J, const DmsRows& R,
                         const DmsStepEvent* pre = nullptr, const double* pre_rates = nullptr)
{
	const double actual_dt = L.prop.time - L.pb_time;
	const double r_after = dms_prop_radius(L.prop);
	double ea_t, ea_r[3], ea_v[3];
	double v_after;
	if(pre) /* the event of this propagator state, computed beforehand (dms_step_event) */
	{
		ea_t = pre->t;
		for(int c = 0; c < 3; c++)
		{
			ea_r[c] = pre->r[c];
			ea_v[c] = pre->v[c];
		}
		v_after = pre->speed;
	}
	else if constexpr(MX)
		v_after = dms_prop_event_mx(L.prop, ea_t, ea_r, ea_v);
	else
		v_after = dms_prop_event_speed(L.prop, ea_t, ea_r, ea_v);
#define DMS_RETURN_AT(tt, rr, vv, code) \
	do                                  \
	{                                   \
		L.ev_t = (tt) + L.time_origin;  \
		for(int c_ = 0; c_ < 3; c_++)   \
		{                               \
			L.ev_r[c_] = (rr)[c_];      \
			L.ev_v[c_] = (vv)[c_];      \
		}                               \
		return (code);                  \
	} while(0)
	{
		/* the reference's four checks in its order, under one rarely taken branch */
		const bool failed = !L.rk_ok;
		const bool non_finite = !dm_isfinite(r_after) | !dm_isfinite(v_after) | !dm_isfinite(actual_dt) | (actual_dt <= 0.0);
		const bool too_fast = v_after > DMS_V_MAX;
		const bool budget = (J.rk_budget > 0) & (L.rk_all > J.rk_budget);
		if(__builtin_expect(failed | non_finite | too_fast | budget, 0))
			DMS_RETURN_AT(ea_t, ea_r, ea_v,
			              failed ? DMS_NUMERICAL_FAILURE
			                     : non_finite ? DMS_NON_FINITE_STATE : too_fast ? DMS_SPEED_LIMIT : DMS_WALL_TIME_LIMIT);
	}
	const int fast = J.terminate_on_capture;
	const double od_limit = fast ? DMS_CAPTURE_MODE_MAX_OPTICAL_DEPTH_STEP : DMS_MAX_OPTICAL_DEPTH_STEP;
	const int od_count = fast ? DMS_CAPTURE_MODE_OPTICAL_DEPTH_PIECES : DMS_MAX_OPTICAL_DEPTH_PIECES;
	static_assert((DMS_MAX_OPTICAL_DEPTH_PIECES & (DMS_MAX_OPTICAL_DEPTH_PIECES - 1)) == 0
	                  && (DMS_CAPTURE_MODE_OPTICAL_DEPTH_PIECES & (DMS_CAPTURE_MODE_OPTICAL_DEPTH_PIECES - 1)) == 0,
	              "piece counts are powers of two, so x / od_count == x * inv_pieces exactly");
	const double inv_pieces = fast ? 1.0 / DMS_CAPTURE_MODE_OPTICAL_DEPTH_PIECES : 1.0 / DMS_MAX_OPTICAL_DEPTH_PIECES;
	int scattering = 0, reflection = 0;
	bool want_next = false;
	double rate_next = 0.0;
	float rate_next_f = 0.0f;
	double acc_t = ea_t, acc_r[3] = {ea_r[0], ea_r[1], ea_r[2]}, acc_v[3] = {ea_v[0], ea_v[1], ea_v[2]};
	double is = 0.0, ie = 0.0;
	if(dms_interior_interval(L.eb_r, ea_r, L.r_before, r_after, is, ie))
	{
		L.n_int++;
		if((r_after < DMS_rSun) & (v_after < 0.0))
			return DMS_MAX_FREE_STEPS; /* unreachable 'break' of the reference, :964-968 */
		if constexpr(MX)
		{
			/* FP32 optical depth: rates in 1/s, times in s, Kahan-compensated sums */
			float frf[DMS_MAX_OPTICAL_DEPTH_PIECES + 1], rtf[DMS_MAX_OPTICAL_DEPTH_PIECES + 1];
			const bool use_cached = L.r_before < DMS_rSun && is <= 1.0e-12;
			float tau2 = 0.0f, dtau = 0.0f;
			int invalid = 0;
			const float dt_s = (float) (actual_dt * DMS_MX_INV_SEC);
			if(ie > is && actual_dt > 0.0)
			{
				const float isf = (float) is, span = (float) (ie - is);
#pragma unroll
				for(int i = 0; i <= DMS_MAX_OPTICAL_DEPTH_PIECES; i++)
					frf[i] = (i <= od_count) ? isf + (float) i * (float) inv_pieces * span : 0.0f;
				const float b[6] = {(float) (L.eb_r[0] * DMS_MX_INV_RSUN), (float) (L.eb_r[1] * DMS_MX_INV_RSUN),
				                    (float) (L.eb_r[2] * DMS_MX_INV_RSUN), (float) L.eb_v[0], (float) L.eb_v[1], (float) L.eb_v[2]};
				const float a[6] = {(float) (ea_r[0] * DMS_MX_INV_RSUN), (float) (ea_r[1] * DMS_MX_INV_RSUN),
				                    (float) (ea_r[2] * DMS_MX_INV_RSUN), (float) ea_v[0], (float) ea_v[1], (float) ea_v[2]};
				want_next = r_after < DMS_rSun;
				if(pre_rates)
				{
					/* FP32 rates evaluated beforehand, held as doubles (exact) */
#pragma unroll
					for(int i = 0; i <= DMS_MAX_OPTICAL_DEPTH_PIECES; i++)
						rtf[i] = (float) pre_rates[i];
					if(want_next)
						rate_next_f = (float) pre_rates[DMS_MAX_OPTICAL_DEPTH_PIECES + 1];
				}
				else
					dms_rates_at_fractions_mx(P, S, b, a, frf, od_count + 1, use_cached, L.rate_before_f, rtf, want_next,
					                          (float) (r_after * DMS_MX_INV_RSUN), (float) v_after, &rate_next_f);
#pragma unroll
				for(int i = 0; i <= DMS_MAX_OPTICAL_DEPTH_PIECES; i++)
					if(i <= od_count && (!mx_isfinite(rtf[i]) || rtf[i] < 0.0f))
						invalid = 1;
				if(!invalid)
				{
					/* equal pieces of width w: composite trapezoid rules on 4 and 2 pieces */
					const float w = dt_s * span * (float) inv_pieces;
					if(od_count >= 4)
					{
						const float ends = 0.5f * (rtf[0] + rtf[4]);
						dtau = w * (ends + (rtf[1] + rtf[2] + rtf[3]));
						tau2 = (2.0f * w) * (ends + rtf[2]);
					}
					else
					{
						dtau = w * (0.5f * (rtf[0] + rtf[od_count]) + (od_count == 2 ? rtf[1] : 0.0f));
						tau2 = dtau;
					}
				}
			}
			if(__builtin_expect(invalid || !mx_isfinite(dtau) || dtau < 0.0f, 0))
				DMS_RETURN_AT(ea_t, ea_r, ea_v, DMS_NUMERICAL_FAILURE);
			const float scale = fmaxf(dtau, (float) DMS_OPTICAL_DEPTH_ABSOLUTE_TOLERANCE);
			const float relerr = fabsf(dtau - tau2) * mx_rcp(scale);
			const bool rej_od = dtau > (float) od_limit;
			const bool rej_acc = !fast && relerr > (float) DMS_OPTICAL_DEPTH_RELATIVE_TOLERANCE;
			if(rej_od || rej_acc)
			{
				L.n_rej++;
				L.od_retries++;
				if(__builtin_expect(L.od_retries > DMS_MAX_OPTICAL_DEPTH_RETRIES, 0))
					DMS_RETURN_AT(L.eb_t, L.eb_r, L.eb_v, DMS_NUMERICAL_FAILURE);
				L.prop.time = L.pb_time;
				L.prop.radius = L.pb_radius;
				L.prop.phi = L.pb_phi;
				L.prop.vr = L.pb_vr;
				L.prop.ts = L.pb_ts;
				L.rb_cached = 1;
				double factor = DMS_RK45_MAX_STEP_FACTOR;
				if(rej_od && dtau > 0.0f)
					factor = dm_min(factor, 0.8 * od_limit / (double) dtau);
				if(rej_acc && relerr > 0.0f)
					factor = dm_min(factor, 0.8 * (double) mx_sqrt((float) DMS_OPTICAL_DEPTH_RELATIVE_TOLERANCE * mx_rcp(relerr)));
				factor = dm_max(DMS_RK45_MIN_STEP_FACTOR, dm_min(factor, 1.0));
				L.prop.ts = dms_rk45_sanitized(factor * actual_dt);
				return -2;
			}
			const float remaining_total = L.mlx_f - L.mlx_c;
			if(__builtin_expect(dtau >= remaining_total, 0)) /* a scattering inside this step */
			{
				float remaining = remaining_total;
				int hit = -1;
				float ps = 0.0f, pe = 0.0f, rs0 = 0.0f, re0 = 0.0f;
#pragma unroll
				for(int i = 0; i < DMS_MAX_OPTICAL_DEPTH_PIECES; i++)
				{
					if(hit < 0 && i < od_count && frf[i + 1] > frf[i])
					{
						const float piece_tau = 0.5f * (rtf[i] + rtf[i + 1]) * (dt_s * (frf[i + 1] - frf[i]));
						if(piece_tau >= remaining)
						{
							hit = i;
							ps = frf[i];
							pe = frf[i + 1];
							rs0 = rtf[i];
							re0 = rtf[i + 1];
						}
						else
							remaining -= piece_tau;
					}
				}
				double fraction = ie;
				if(hit >= 0)
				{
					const float lf = dms_solve_linear_rate_mx(rs0, re0, dt_s * (pe - ps), remaining);
					fraction = (double) ps + (double) lf * (double) (pe - ps);
				}
				dms_interpolate_event(L.eb_t, L.eb_r, L.eb_v, ea_t, ea_r, ea_v, fraction, acc_t, acc_r, acc_v);
				scattering = 1;
			}
			else
				dms_kahan_sub_mx(L.mlx_f, L.mlx_c, dtau);
		}
		else
		{
			/* Build_Optical_Depth_Pieces, :347-414 */
			double fr[DMS_MAX_OPTICAL_DEPTH_PIECES + 1], rt[DMS_MAX_OPTICAL_DEPTH_PIECES + 1];
			const bool use_cached = (L.r_before < DMS_rSun) & (is <= 1.0e-12);
			double tau2 = 0.0, dtau = 0.0;
			int invalid = 0, pc = 0;
			if((ie > is) & (actual_dt > 0.0))
			{
				const double span = ie - is;
	#pragma unroll
				for(int i = 0; i <= DMS_MAX_OPTICAL_DEPTH_PIECES; i++)
					fr[i] = (i <= od_count) ? is + (double) i * inv_pieces * span : 0.0;
				/* Scattering_Rate_At_Fraction for all fractions, grid loads issued together
				 * (or their values, evaluated beforehand by dms_rate_fraction_slot) */
				want_next = r_after < DMS_rSun;
				if(pre_rates)
				{
	#pragma unroll
					for(int i = 0; i <= DMS_MAX_OPTICAL_DEPTH_PIECES; i++)
						rt[i] = pre_rates[i];
					if(want_next)
						rate_next = pre_rates[DMS_MAX_OPTICAL_DEPTH_PIECES + 1];
				}
				else
					dms_rates_at_fractions(P, S, L.eb_r, L.eb_v, ea_r, ea_v, fr, od_count + 1, use_cached, L.rate_before, rt,
					                       want_next, r_after, v_after, &rate_next);
	#pragma unroll
				for(int i = 0; i <= DMS_MAX_OPTICAL_DEPTH_PIECES; i++)
					invalid |= (int) ((i <= od_count) & (!dm_isfinite(rt[i]) | (rt[i] < 0.0)));
				if(!invalid)
				{
	#define DMS_TRAP(s, e) (0.5 * (rt[s] + rt[e]) * (actual_dt * (fr[e] - fr[s])))
					if(od_count >= 4)
						tau2 = DMS_TRAP(0, 2) + DMS_TRAP(2, 4);
					else
					{
	#pragma unroll
						for(int i = 0; i < DMS_MAX_OPTICAL_DEPTH_PIECES; i++)
							if(i < od_count)
								tau2 += DMS_TRAP(i, i + 1);
					}
	#pragma unroll
					for(int i = 0; i < DMS_MAX_OPTICAL_DEPTH_PIECES; i++)
						if(i < od_count && fr[i + 1] > fr[i])
						{
							pc++;
							dtau += DMS_TRAP(i, i + 1);
						}
	#undef DMS_TRAP
				}
			}
			if(__builtin_expect((invalid != 0) | !dm_isfinite(dtau) | (dtau < 0.0), 0))
				DMS_RETURN_AT(ea_t, ea_r, ea_v, DMS_NUMERICAL_FAILURE);
			const double scale = dm_max(dtau, DMS_OPTICAL_DEPTH_ABSOLUTE_TOLERANCE);
			const double relerr = dm_fabs(dtau - tau2) / scale;
			const bool rej_od = dtau > od_limit;
			const bool rej_acc = !fast & (relerr > DMS_OPTICAL_DEPTH_RELATIVE_TOLERANCE);
			if(rej_od | rej_acc)
			{
				L.n_rej++;
				L.od_retries++;
				if(__builtin_expect(L.od_retries > DMS_MAX_OPTICAL_DEPTH_RETRIES, 0))
					DMS_RETURN_AT(L.eb_t, L.eb_r, L.eb_v, DMS_NUMERICAL_FAILURE);
				L.prop.time = L.pb_time;
				L.prop.radius = L.pb_radius;
				L.prop.phi = L.pb_phi;
				L.prop.vr = L.pb_vr;
				L.prop.ts = L.pb_ts;
				/* the step is retried from the same state: rate_before stays valid */
				L.rb_cached = 1;
				double factor = DMS_RK45_MAX_STEP_FACTOR;
				if(dtau > od_limit && dtau > 0.0)
					factor = dm_min(factor, 0.8 * od_limit / dtau);
				if(rej_acc && relerr > 0.0)
					factor = dm_min(factor, 0.8 * dm_sqrt(DMS_OPTICAL_DEPTH_RELATIVE_TOLERANCE / relerr));
				factor = dm_max(DMS_RK45_MIN_STEP_FACTOR, dm_min(factor, 1.0));
				L.prop.ts = dms_rk45_sanitized(factor * actual_dt);
				return -2;
			}
			if(__builtin_expect(dtau >= L.mlx, 0)) /* a scattering inside this step */
			{
				/* locate the scattering piece; pieces with end > start only */
				double remaining = L.mlx;
				int hit = -1;
	#pragma unroll
				for(int i = 0; i < DMS_MAX_OPTICAL_DEPTH_PIECES; i++)
				{
					if(hit < 0 && i < od_count && fr[i + 1] > fr[i])
					{
						const double piece_dt = actual_dt * (fr[i + 1] - fr[i]);
						const double piece_tau = 0.5 * (rt[i] + rt[i + 1]) * piece_dt;
						if(piece_tau >= remaining)
							hit = i;
						else
							remaining -= piece_tau;
					}
				}
				double fraction = ie;
				if(hit >= 0)
				{
					double ps = 0.0, pe = 0.0, rs0 = 0.0, re0 = 0.0;
	#pragma unroll
					for(int i = 0; i < DMS_MAX_OPTICAL_DEPTH_PIECES; i++)
						if(i == hit)
						{
							ps = fr[i];
							pe = fr[i + 1];
							rs0 = rt[i];
							re0 = rt[i + 1];
						}
					const double piece_dt = actual_dt * (pe - ps);
					const double lf = dms_solve_linear_rate(rs0, re0, piece_dt, remaining);
					fraction = ps + lf * (pe - ps);
				}
				dms_interpolate_event(L.eb_t, L.eb_r, L.eb_v, ea_t, ea_r, ea_v, fraction, acc_t, acc_r, acc_v);
				scattering = 1;
			}
			else
				L.mlx -= dtau;
		}
	}
	else
		L.n_ext++;
	if(!scattering)
	{
		double bt = ea_t, br[3] = {ea_r[0], ea_r[1], ea_r[2]}, bv[3] = {ea_v[0], ea_v[1], ea_v[2]};
		double rb = r_after, vb = v_after;
		if(__builtin_expect((L.r_before < J.max_distance) & (r_after >= J.max_distance) & (r_after > L.r_before), 0))
		{
			const double bf = (J.max_distance - L.r_before) / (r_after - L.r_before);
			dms_interpolate_event(L.eb_t, L.eb_r, L.eb_v, ea_t, ea_r, ea_v, bf, bt, br, bv);
			rb = dms_norm(br);
			vb = dms_norm(bv);
		}
		const bool beyond = (rb >= J.max_distance) & (rb >= L.r_before);
		if(__builtin_expect(beyond, 0) && vb > dms_local_escape_speed(S, rb))
		{
			acc_t = bt;
			for(int c = 0; c < 3; c++)
			{
				acc_r[c] = br[c];
				acc_v[c] = bv[c];
			}
			reflection = 1;
		}
		else if(__builtin_expect(beyond, 0))
		{
			if(dms_uncaptured_bound(L, S, P.mass, br, bv))
				DMS_RETURN_AT(bt, br, bv, DMS_NUMERICAL_FAILURE);
			double it, ir[3], iv[3];
			if(dms_bound_kepler_return(bt, br, bv, it, ir, iv))
			{
				L.time_steps++;
				L.od_retries = 0;
				if constexpr(MX)
					dms_commit_mx(L, R, J, S, P.mass, bt, dms_norm(br), dms_dot(bv, bv));
				else
					dms_commit(L, R, J, S, P.mass, bt, br, bv);
				L.ev_t = it + L.time_origin;
				for(int c = 0; c < 3; c++)
				{
					L.ev_r[c] = ir[c];
					L.ev_v[c] = iv[c];
				}
				/* reset_propagator_at_absolute_event, :814-822 */
				L.time_origin = L.ev_t;
				dms_prop_init(L.prop, 0.0, L.ev_r, L.ev_v);
				L.eb_valid = 0;
				if(!J.terminate_on_capture && L.prev_time_sec >= 0.0)
					L.prev_time_sec = 0.0;
				return -2;
			}
		}
	}
	L.time_steps++;
	L.od_retries = 0;
	int captured_now = 0;
	if constexpr(MX)
	{
		/* at the step's end point the propagator's FP64 radius and speed; an
		 * interpolated (scattering) or boundary (reflection) event its own */
		double cr = r_after, cv2 = v_after * v_after;
		if(__builtin_expect(scattering || reflection, 0))
		{
			cr = dms_norm(acc_r);
			cv2 = dms_dot(acc_v, acc_v);
		}
		dms_commit_mx(L, R, J, S, P.mass, acc_t, cr, cv2);
	}
	else
		captured_now = dms_commit(L, R, J, S, P.mass, acc_t, acc_r, acc_v);
	L.ev_t = acc_t + L.time_origin;
	for(int c = 0; c < 3; c++)
	{
		L.ev_r[c] = acc_r[c];
		L.ev_v[c] = acc_v[c];
	}
	{
		/* abort_if_uncaptured_bound on the committed event: the energy was just
		 * computed for the same (|r|, |v|) by Update_Capture_State */
		double E = L.last_E;
		if(dm_isnan(E))
			E = MX ? dms_capture_energy_mx(S, P.mass, dms_norm(L.ev_r), dms_dot(L.ev_v, L.ev_v))
			       : dms_capture_energy_eV(S, P.mass, dms_norm(L.ev_r), dms_norm(L.ev_v));
		if(!L.captured & !(dm_isfinite(E) & (E >= 0.0)))
			return DMS_NUMERICAL_FAILURE;
	}
	{
		const bool capture_stop = (J.terminate_on_capture != 0) & (captured_now != 0);
		if(__builtin_expect(capture_stop | (reflection != 0) | (scattering != 0), 0))
			return capture_stop ? DMS_CAPTURE_MODE : (reflection ? DMS_OUTWARD_ESCAPE : -3);
	}
	/* the next step starts from this propagator state: Event_In_3D is ea */
	L.eb_t = ea_t;
	for(int c = 0; c < 3; c++)
	{
		L.eb_r[c] = ea_r[c];
		L.eb_v[c] = ea_v[c];
	}
	L.r_before = r_after;
	L.v_before = v_after;
	L.eb_valid = 1;
	/* Total_DM_Scattering_Rate at the new r_before, read with this step's gathers */
	L.rb_cached = want_next ? 1 : 0;
	L.rate_before = rate_next;
	L.rate_before_f = rate_next_f;
	return DMS_POST_NEXT;
#undef DMS_RETURN_AT
}

/* Sample_Target_Velocity (Romano & Walsh), :1160-1226; false where it throws */
DMS_HD bool dms_sample_target_velocity(DmsLane& L, double temperature, double target_mass, const double* vel, double* out)
{
	if(!dm_isfinite(temperature) || temperature <= 0.0 || !dm_isfinite(target_mass) || target_mass <= 0.0)
		return false;
	const double kappa = dm_sqrt(target_mass / 2.0 / temperature);
	const double vDM = dms_norm(vel);
	if(!dm_isfinite(kappa) || kappa <= 0.0 || !dm_isfinite(vDM) || vDM <= 0.0)
		return false;
	const double y = kappa * vDM;
	double x = y, mu = 1.0;
	unsigned long attempts = 0;
	for(;;)
	{
		L.n_tv++;
		const double xs = x * x, ys = y * y;
		const double raw = xs + ys - 2.0 * x * y * mu;
		const double den = x + y;
		const double roundoff = dm_max(1.0, xs + ys + 2.0 * dm_fabs(x * y));
		if(!dm_isfinite(x) || x < 0.0 || !dm_isfinite(y) || y <= 0.0 || !dm_isfinite(mu) || mu < -1.0 || mu > 1.0
		   || !dm_isfinite(raw) || raw < -1.0e-12 * roundoff || !dm_isfinite(den) || den <= 0.0)
			return false;
		const double ratio = dm_sqrt(dm_max(0.0, raw)) / den;
		if(!dm_isfinite(ratio) || ratio < 0.0 || ratio > 1.0 + 1.0e-12)
			return false;
		const double accept = dm_min(1.0, ratio);
		if(!(accept < dm_uniform(&L.rng)))
			break;
		attempts++;
		if(attempts > DMS_TARGET_VELOCITY_MAX_REJECTION_ATTEMPTS)
			return false;
		mu = -1.0 + 2.0 * dm_uniform(&L.rng);
		const double xi_1 = dm_uniform(&L.rng);
		if(xi_1 < 2.0 / (DM_SQRT_PI * y + 2.0))
		{
			const double xi_2 = dm_max(dm_uniform(&L.rng), DM_DBL_MIN);
			const double xi_3 = dm_max(dm_uniform(&L.rng), DM_DBL_MIN);
			const double z = -dm_log(xi_2) - dm_log(xi_3);
			x = dm_sqrt(z);
		}
		else
		{
			const double xi_2 = dm_max(dm_uniform(&L.rng), DM_DBL_MIN);
			const double xi_3 = dm_uniform(&L.rng);
			const double xi_4 = dm_max(dm_uniform(&L.rng), DM_DBL_MIN);
			const double c = dm_cos(DM_PI_2 * xi_3);
			const double z = -dm_log(xi_2) - c * c * dm_log(xi_4);
			x = dm_sqrt(z);
		}
	}
	const double vT = x / kappa;
	const double phi = 2.0 * DM_PI * dm_uniform(&L.rng);
	double u[3];
	if(!dms_unit_vector_at_angle(mu, phi, vel, u))
		return false;
	for(int c = 0; c < 3; c++)
		out[c] = vT * u[c];
	return dm_isfinite(dms_norm(out));
}

/* Scatter after the target choice (:1249-1269): target velocity, CM angle
 * (low-mass contact 2 xi - 1; dark photon src/Dark_Photon.cpp:286-325) and
 * New_DM_Velocity (:1228-1247).  target: -1 electron, >= 0 isotope. */
DMS_HD bool dms_scatter_rest(DmsLane& L, const DmsPointView& P, const DmsSolarView& S, int target)
{
	const double r = dms_norm(L.ev_r);
	const double v = dms_norm(L.ev_v);
	const double mT = (target == -1) ? DMS_mElectron : S.iso_mass[target];
	double vT[3];
	if(!dms_sample_target_velocity(L, dms_temperature(S, r), mT, L.ev_v, vT))
		return false;
	const double xi = dm_uniform(&L.rng);
	double cos_a;
	if(P.interaction == 2 && P.ff == 1)
	{
		const double m2 = P.m_dark_photon * P.m_dark_photon;
		const double mu = (target < 0) ? P.mu_e : P.mass * S.iso_mass[target] / (P.mass + S.iso_mass[target]);
		const double q2max = 4.0 * ((mu * v) * (mu * v));
		cos_a = (m2 * (2.0 * xi - 1.0) + q2max * xi) / (m2 + q2max * xi);
	}
	else
		cos_a = 2.0 * xi - 1.0;
	const double mDM = P.mass;
	if(!dm_isfinite(cos_a) || cos_a < -1.0 - 1.0e-12 || cos_a > 1.0 + 1.0e-12 || !dm_isfinite(mDM) || mDM <= 0.0
	   || !dm_isfinite(mT) || mT <= 0.0 || !dm_isfinite(dms_norm(L.ev_v)) || !dm_isfinite(dms_norm(vT)))
		return false;
	cos_a = dm_clamp_cos(cos_a);
	const double phi = 2.0 * DM_PI * dm_uniform(&L.rng);
	double n[3];
	if(!dms_unit_vector_at_angle(cos_a, phi, L.ev_v, n))
		return false;
	const double d[3] = {vT[0] - L.ev_v[0], vT[1] - L.ev_v[1], vT[2] - L.ev_v[2]};
	const double rel = dms_norm(d);
	const double sum_m = mT + mDM;
	const double f = mT * rel / sum_m;
	double out[3];
	for(int c = 0; c < 3; c++)
		out[c] = f * n[c] + (mDM * L.ev_v[c] + mT * vT[c]) / (mT + mDM);
	if(!dm_isfinite(dms_norm(out)))
		return false;
	for(int c = 0; c < 3; c++)
		L.ev_v[c] = out[c];
	return true;
}

/* Target choice of Sample_Target (:1140-1157) from the per-target rates
 * rates[0..n_iso-1] (nuclei) and rates[63] (electrons); -2 on an invalid rate. */
DMS_HD int dms_choose_target(const double* rates, int n_iso, double xi)
{
	for(int i = 0; i < n_iso; i++)
		if(!dm_isfinite(rates[i]) || rates[i] < 0.0)
			return -2;
	const double re = rates[DMS_MAX_TARGETS - 1];
	if(!dm_isfinite(re) || re < 0.0)
		return -2;
	double total = re;
	for(int i = 0; i < n_iso; i++)
		total = total + rates[i];
	if(!dm_isfinite(total) || total <= 0.0)
		return -2;
	double sum = re / total;
	if(sum > xi)
		return -1;
	for(int i = 0; i < n_iso; i++)
	{
		sum += rates[i] / total;
		if(sum > xi || i == n_iso - 1)
			return i;
	}
	return -2;
}

/* Simulate start for one trajectory (:1284-1302). */
DMS_HD void dms_trajectory_begin(DmsLane& L, const DmsSolarView& S, double dm_mass, const DmsJobView& J, uint64_t id,
                                 const dms_event& ic)
{
	dm_rng_init(&L.rng, J.seed, J.point_id, id, DMS_STREAM_SIM);
	L.ev_t = ic.t;
	for(int c = 0; c < 3; c++)
	{
		L.ev_r[c] = ic.r[c];
		L.ev_v[c] = ic.v[c];
	}
	L.n_scat = 0;
	L.reason = DMS_UNKNOWN;
	L.captured = 0;
	L.bound = 0;
	L.event_observed = 0;
	L.boundary_escape_observed = 0;
	L.survival_valid = 1;
	L.nie = 0;
	L.t_capture = -1.0;
	L.t_last_bound = -1.0;
	L.t_unbind = dm_nan();
	L.t_boundary = dm_nan();
	L.r_first_neg = -1.0;
	L.E_first_neg = dm_nan();
	L.dE_first_neg = dm_nan();
	L.drift_eV = 0.0;
	L.drift_rel = 0.0;
	L.prev_time_sec = -1.0;
	L.prev_r_km = 0.0;
	L.prev_v2 = 0.0;
	L.prev_cap_E = dms_capture_energy_eV(S, dm_mass, dms_norm(ic.r), dms_norm(ic.v));
	L.ff_ref_E = dm_nan();
	L.cur_bin = -1;
	L.acc_dt = 0.0;
	L.acc_v2 = 0.0;
	L.acc_cnt = 0;
	L.pf_bin = -1;
	L.n_touched = 0;
	L.rk_all = 0;
	L.n_int = 0;
	L.n_ext = 0;
	L.n_rej = 0;
	L.n_tv = 0;
}

/* After Propagate_Freely returned pr (:1313-1347).  Returns the next phase. */
DMS_HD int dms_after_propagation(DmsLane& L, int pr)
{
	if(pr == DMS_SCATTER && dms_norm(L.ev_r) < DMS_rSun)
		return PH_SCAT;
	L.reason = pr;
	return PH_FIN;
}

/* Scatter bookkeeping after a successful Scatter (:1327-1341). Returns next phase. */
DMS_HD int dms_after_scatter(DmsLane& L, const DmsSolarView& S, double dm_mass, const DmsJobView& J, const DmsRows& R)
{
	L.n_scat++;
	const int cap = dms_update_capture_state(L, S, dm_mass, J.terminate_on_capture, dms_norm(L.ev_r), dms_norm(L.ev_v), L.ev_t, 1);
	if(!J.terminate_on_capture)
	{
		/* Reset_Bincount_Anchor, :671-680 */
		L.prev_time_sec = L.ev_t / DMS_sec;
		L.prev_r_km = dms_norm(L.ev_r) / DMS_km;
		const double v = dms_norm(L.ev_v) / (DMS_km / DMS_sec);
		L.prev_v2 = v * v;
		dms_bin_anchor(L, R, J);
	}
	if(J.terminate_on_capture && cap)
	{
		L.reason = DMS_CAPTURE_MODE;
		return PH_FIN;
	}
	return PH_SEG;
}

DMS_HD bool dms_invalidates_survival(int reason)
{
	switch(reason)
	{
		case DMS_ENERGY_DRIFT_ESCAPE:
		case DMS_NUMERICAL_FAILURE:
		case DMS_NON_FINITE_STATE:
		case DMS_SPEED_LIMIT:
		case DMS_WALL_TIME_LIMIT:
		case DMS_MAX_FREE_STEPS:
		case DMS_MAX_SCATTERINGS:
		case DMS_UNKNOWN: return true;
		default: return false;
	}
}

/* Post-loop classification of Simulate (:1349-1397) plus the
 * Particle_Free / Particle_Reflected flags (:529-542) and the 1 AU shift of
 * reflected particles (src/Data_Generation.cpp:829).  Fills the record. */
DMS_HD void dms_trajectory_end(DmsLane& L, const DmsSolarView& S, double dm_mass, const DmsJobView& J, uint64_t id,
                               const dms_event& ic, dms_traj_record* rec, DmsSummary* sum)
{
	int reason = L.reason;
	if(reason == DMS_UNKNOWN && L.n_scat >= J.max_scatterings)
		reason = DMS_MAX_SCATTERINGS;
	const double t_term = L.ev_t / DMS_sec;
	double t_boundary = dm_nan();
	double t_unbind = L.t_unbind;
	if(L.captured)
	{
		const double rf = dms_norm(L.ev_r), vf = dms_norm(L.ev_v);
		const double vesc = dms_local_escape_speed(S, rf);
		const double Ef = 0.5 * dm_mass * (vf * vf - vesc * vesc);
		const double Ef_eV = Ef / DMS_eV;
		if(reason == DMS_OUTWARD_ESCAPE && Ef_eV >= 0.0)
		{
			if(dm_isfinite(t_unbind))
			{
				L.boundary_escape_observed = 1;
				t_boundary = t_term;
				L.event_observed = 1;
			}
			else
			{
				reason = DMS_ENERGY_DRIFT_ESCAPE;
				L.survival_valid = 0;
				L.nie = 1;
				L.boundary_escape_observed = 0;
				L.event_observed = 0;
			}
		}
	}
	if(dms_invalidates_survival(reason))
		L.survival_valid = 0;
	if(L.captured && !L.event_observed)
	{
		t_unbind = dm_nan();
		if(reason != DMS_CAPTURE_MODE)
			L.survival_valid = 0;
	}
	uint32_t flags = 0;
	if(L.captured)
		flags |= DMS_F_CAPTURED;
	if(L.event_observed)
		flags |= DMS_F_EVENT_OBSERVED;
	if(L.boundary_escape_observed)
		flags |= DMS_F_BOUNDARY_ESCAPE_OBSERVED;
	if(L.survival_valid)
		flags |= DMS_F_SURVIVAL_VALID;
	if(L.nie)
		flags |= DMS_F_NUMERICALLY_INVALID_ESCAPE;
	if(reason == DMS_OUTWARD_ESCAPE)
	{
		if(L.n_scat == 0)
			flags |= DMS_F_FREE;
		const double rf = dms_norm(L.ev_r);
		const double vesc = dm_sqrt(2 * DMS_G_Newton * DMS_mSun / rf);
		if(rf > DMS_rSun && dms_norm(L.ev_v) > vesc && L.n_scat > 0)
		{
			flags |= DMS_F_REFLECTED;
			double r[3] = {L.ev_r[0], L.ev_r[1], L.ev_r[2]}, v[3] = {L.ev_v[0], L.ev_v[1], L.ev_v[2]};
			if(!dms_kepler_shift(r, v, 1.0 * DMS_AU))
				flags |= DMS_F_REFLECTION_SHIFT_FAILED;
		}
	}
	if(sum)
	{
		sum->reason = reason;
		sum->flags = flags;
		sum->n_scat = L.n_scat;
		sum->t_capture_sec = L.t_capture;
		sum->t_unbind_sec = t_unbind;
		sum->n_rk = L.rk_all;
		sum->n_int = L.n_int;
		sum->n_ext = L.n_ext;
		sum->n_rej = L.n_rej;
		sum->n_tv = L.n_tv;
	}
	if(rec)
	{
		rec->initial = ic;
		rec->final.t = L.ev_t;
		for(int c = 0; c < 3; c++)
		{
			rec->final.r[c] = L.ev_r[c];
			rec->final.v[c] = L.ev_v[c];
		}
		rec->trajectory_id = id;
		rec->n_scatterings = L.n_scat;
		rec->reason = reason;
		rec->flags = flags;
		rec->t_capture_sec = L.t_capture;
		rec->t_last_bound_sec = L.t_last_bound;
		rec->t_final_unbinding_scatter_sec = t_unbind;
		rec->t_boundary_escape_sec = t_boundary;
		rec->t_termination_sec = t_term;
		rec->r_first_negative_km = L.r_first_neg;
		rec->E_first_negative_eV = L.E_first_neg;
		rec->dE_first_negative_from_prev_eV = L.dE_first_neg;
		rec->max_free_energy_drift_eV = L.drift_eV;
		rec->max_free_energy_drift_rel = L.drift_rel;
		rec->n_rk_attempts = L.rk_all;
		rec->n_steps_interior = L.n_int;
		rec->n_steps_exterior = L.n_ext;
		rec->n_tau_rejections = L.n_rej;
		rec->n_tv_iterations = L.n_tv;
	}
}

/* Record of a trajectory whose initial Kepler shift failed
 * (src/Data_Generation.cpp:749-761). */
DMS_HD void dms_shift_failure_record(uint64_t id, const dms_event& ic, dms_traj_record* rec, DmsSummary* sum)
{
	if(sum)
	{
		sum->reason = DMS_NUMERICAL_FAILURE;
		sum->flags = DMS_F_SHIFT_FAILED;
		sum->n_scat = 0;
		sum->t_capture_sec = -1.0;
		sum->t_unbind_sec = dm_nan();
		sum->n_rk = 0;
		sum->n_int = sum->n_ext = sum->n_rej = sum->n_tv = 0;
	}
	if(rec)
	{
		rec->initial = ic;
		rec->final = ic;
		rec->trajectory_id = id;
		rec->n_scatterings = 0;
		rec->reason = DMS_NUMERICAL_FAILURE;
		rec->flags = DMS_F_SHIFT_FAILED;
		rec->t_capture_sec = -1.0;
		rec->t_last_bound_sec = -1.0;
		rec->t_final_unbinding_scatter_sec = dm_nan();
		rec->t_boundary_escape_sec = dm_nan();
		rec->t_termination_sec = -1.0;
		rec->r_first_negative_km = -1.0;
		rec->E_first_negative_eV = dm_nan();
		rec->dE_first_negative_from_prev_eV = dm_nan();
		rec->max_free_energy_drift_eV = 0.0;
		rec->max_free_energy_drift_rel = 0.0;
		rec->n_rk_attempts = 0;
		rec->n_steps_interior = rec->n_steps_exterior = rec->n_tau_rejections = rec->n_tv_iterations = 0;
	}
}


/* Per-lane mode: add the touched bins k = first, first + stride, ... < n of
 * scratch row `row` (a finished trajectory of job `job`) to the sums of its
 * class (1 captured, 2 completed outward escape, 0 none) and clear them.  One
 * lane walks the whole list (stride 1), or the 64 lanes of a wave share it. */
DMS_HD void dms_flush_rows(const DmsRows& R, int nbins, uint32_t row_index, uint32_t job, uint32_t n, int cls, uint32_t first,
                           uint32_t stride)
{
	DmsBinAcc* row = R.bins + (size_t) row_index * (size_t) nbins;
	const uint16_t* list = R.touched + (size_t) row_index * (size_t) nbins;
	double* h = R.hist + (size_t) job * 8 * (size_t) nbins + (cls == 2 ? 4 * (size_t) nbins : 0);
	unsigned long long* hc = R.hcounts + (size_t) job * 2 * (size_t) nbins + (cls == 2 ? (size_t) nbins : 0);
	for(uint32_t k = first; k < n; k += stride)
	{
		const int b = list[k];
		const double dt = row[b].dt;
		const double v2 = row[b].v2dt;
		const uint32_t cn = row[b].cnt;
		if(cls != 0)
		{
#if defined(__HIP_DEVICE_COMPILE__)
			unsafeAtomicAdd(h + b, dt);
			unsafeAtomicAdd(h + nbins + b, v2);
			unsafeAtomicAdd(h + 2 * nbins + b, dt * dt);
			unsafeAtomicAdd(h + 3 * nbins + b, v2 * v2);
			atomicAdd(hc + b, (unsigned long long) cn);
#else
			h[b] += dt;
			h[nbins + b] += v2;
			h[2 * nbins + b] += dt * dt;
			h[3 * nbins + b] += v2 * v2;
			hc[b] += cn;
#endif
		}
		row[b].dt = 0.0;
		row[b].v2dt = 0.0;
		row[b].cnt = 0u;
	}
}

DMS_HD void dms_flush_lane_rows(DmsLane& L, const DmsRows& R, int nbins, int cls)
{
	dms_flush_rows(R, nbins, L.row, L.job, L.n_touched, cls, 0u, 1u);
	L.n_touched = 0;
}

/* ---- phase drivers shared by the kernel and host replays ------------------ */
struct DmsLaneCtx
{
	DmsSolarView S;
	DmsPointView P;
	DmsJobView J;
	DmsRows R;
};

/* One RKF45 attempt (PH_RK; PH_POST once accepted). */
template <bool MX = false>
DMS_HD void dms_lane_advance_rk(DmsLane& L, const DmsLaneCtx& c)
{
	if(L.phase == PH_RK)
	{
		L.inner_iter++;
		L.rk_all++;
		const int r = MX ? dms_rk45_attempt_mx(L.prop, c.S, L.inner_iter) : dms_rk45_attempt(L.prop, c.S, L.inner_iter);
		L.rk_ok = (r >= 0) ? r : L.rk_ok;
		L.phase = (r >= 0) ? (int) PH_POST : L.phase;
	}
}

/* Step post-processing and the scatter hand-off (r, v and the Sample_Target
 * uniform are drawn here, before the wave-level target choice).  *post_ret
 * receives dms_step_post's value (0 when no post ran). */
template <bool MX = false>
DMS_HD void dms_lane_advance_post(DmsLane& L, const DmsLaneCtx& c, int* post_ret = nullptr, const DmsStepEvent* pre = nullptr,
                                  const double* pre_rates = nullptr)
{
	int pr = 0;
	if(L.phase == PH_POST)
	{
		pr = dms_step_post<MX>(L, c.P, c.S, c.J, c.R, pre, pre_rates);
		if((pr == -2) | (pr == DMS_POST_NEXT))
			L.phase = PH_PREP;
		else
			L.phase = dms_after_propagation(L, (pr == -3) ? (int) DMS_SCATTER : pr);
	}
	if(post_ret)
		*post_ret = pr;
	if(L.phase == PH_SCAT)
	{
		L.sc_r = dms_norm(L.ev_r);
		L.sc_v = dms_norm(L.ev_v);
		L.sc_xi = dm_uniform(&L.rng);
	}
}

/* RK attempt, step post-processing and the scatter hand-off. */
template <bool MX = false>
DMS_HD void dms_lane_advance_a(DmsLane& L, const DmsLaneCtx& c)
{
	dms_lane_advance_rk<MX>(L, c);
	dms_lane_advance_post<MX>(L, c);
}

/* Rest of Scatter once sc_target holds the Sample_Target result. */
template <bool MX = false>
DMS_HD void dms_lane_finish_scatter(DmsLane& L, const DmsLaneCtx& c)
{
	if(L.phase != PH_SCAT)
		return;
	if(L.sc_target < -1 || !(MX ? dms_scatter_rest_mx(L, c.P, c.S, L.sc_target) : dms_scatter_rest(L, c.P, c.S, L.sc_target)))
	{
		L.reason = DMS_NUMERICAL_FAILURE;
		L.phase = PH_FIN;
	}
	else
		L.phase = dms_after_scatter(L, c.S, c.P.mass, c.J, c.R);
}

/* Trajectory end up to the per-lane row flush.  Returns the class of the
 * scratch row still to be flushed (dms_flush_rows), or -1 when there is none;
 * the trajectory's flags and termination reason go to flags_out/reason_out. */
DMS_HD int dms_lane_finish_rows_pending(DmsLane& L, const DmsLaneCtx& c, const dms_event& ic, dms_traj_record* rec, DmsSummary* sum,
                                        uint32_t* flags_out = nullptr, int* reason_out = nullptr)
{
	dms_bin_flush(L, c.R, c.J.nbins);
	uint32_t flags = 0;
	int reason = 0;
	{
		DmsSummary tmp;
		DmsSummary* s = sum ? sum : &tmp;
		dms_trajectory_end(L, c.S, c.P.mass, c.J, L.id, ic, rec, s);
		flags = s->flags;
		reason = s->reason;
	}
	L.phase = PH_NEED;
	if(flags_out)
		*flags_out = flags;
	if(reason_out)
		*reason_out = reason;
	if(c.R.per_lane && L.n_touched > 0)
		return (flags & DMS_F_CAPTURED) ? 1 : ((reason == DMS_OUTWARD_ESCAPE) ? 2 : 0);
	return -1;
}

DMS_HD void dms_lane_finish(DmsLane& L, const DmsLaneCtx& c, const dms_event& ic, dms_traj_record* rec, DmsSummary* sum)
{
	const int cls = dms_lane_finish_rows_pending(L, c, ic, rec, sum);
	if(cls >= 0)
		dms_flush_lane_rows(L, c.R, c.J.nbins, cls);
}

DMS_HD void dms_lane_start(DmsLane& L, const DmsLaneCtx& c, const dms_event& ic, bool shift_ok, dms_traj_record* rec, DmsSummary* sum)
{
	if(!shift_ok)
	{
		dms_shift_failure_record(L.id, ic, rec, sum);
		L.phase = PH_NEED;
		return;
	}
	dms_trajectory_begin(L, c.S, c.P.mass, c.J, L.id, ic);
	L.phase = PH_SEG;
}

/* Segment start and step preparation. */
template <bool MX = false>
DMS_HD void dms_lane_advance_b(DmsLane& L, const DmsLaneCtx& c)
{
	if(L.phase == PH_SEG)
	{
		if(!(L.n_scat < c.J.max_scatterings))
		{
			L.reason = DMS_UNKNOWN;
			L.phase = PH_FIN;
		}
		else
		{
			const int pr = dms_segment_start(L, c.S, c.P.mass, c.J);
			L.phase = (pr < 0) ? (int) PH_PREP : dms_after_propagation(L, pr);
		}
	}
	if(L.phase == PH_PREP)
	{
		const int pr = dms_step_prep<MX>(L, c.P, c.S, c.J);
		L.phase = (pr < 0) ? (int) PH_RK : dms_after_propagation(L, pr);
	}
}


/* ---- a lone trajectory split over two waves --------------------------------
 * A trajectory that runs alone in its wave executes one instruction stream:
 * each step is the step preparation, the RKF45 attempts, then the step's
 * post-processing (optical depth, bincount, capture state).  The preparation
 * of the next step depends on the post only through the state the post leaves
 * when it continues (DMS_POST_NEXT): the committed event is the propagator's,
 * r_before and v_before are its radius and speed, and the next step's rate is
 * Total_DM_Scattering_Rate at that point.  A second wave of the block (the
 * helper) therefore computes the next step's preparation, RKF45 attempts and
 * end event from the propagator state alone, assuming the post continues,
 * while the trajectory's own wave runs the post; the post then takes the
 * helper's step (dms_pipe_take) instead of computing it, and its end event.
 * Every other post exit (an optical-depth rejection, a Kepler return, a
 * scattering, an end) discards the helper's steps and the trajectory's wave
 * goes on from its own state.  The helper runs the same functions on the same
 * values, so the trajectory is the serial one bit for bit. */
#define DMS_PIPE_STEP 0     /* an entry holds a prepared, accepted step */
#define DMS_PIPE_PREP_END 1 /* the step preparation ended the segment */

/* Session start: the propagator after the accepted RKF45 attempt of the step
 * whose post the trajectory's wave is about to run, and its event (which that
 * post takes too).  A redo session (redo = 1) starts instead from a step the
 * post rejected on optical depth: the propagator is back at the step's start
 * with the reduced step size, and the helper's first step is that step's
 * preparation and attempts from the lane's own start point (eb, r_before,
 * v_before) and cached start rate. */
struct DmsPipeSeed
{
	DmsProp prop;
	uint64_t time_steps; /* before that post */
	uint64_t rk_all;
	double time_origin;
	uint32_t job;
	uint32_t redo;
	DmsStepEvent ev;
	/* redo only: the step's start */
	DmsStepEvent eb; /* speed = v_before */
	double r_before, rate_before;
	float rate_before_f;
	uint32_t pad;
};

/* One step prepared by the helper: what dms_step_prep and the RKF45 attempts
 * change in the lane state, and the event of the propagator they leave (the
 * step post's end event). */
struct DmsPipeEntry
{
	double time, radius, phi, vr, ts;                /* propagator after the accepted attempt */
	double pb_time, pb_radius, pb_phi, pb_vr, pb_ts; /* the step's start, as the preparation left it */
	double rate_before;
	uint64_t rk_all;
	float rate_before_f;
	int32_t rk_ok, inner_iter;
	uint32_t session;
	int32_t kind;
	uint32_t job;
	DmsStepEvent ev;
	/* the step's start: the event and radius the post takes as eb and r_before */
	DmsStepEvent eb;
	double r_before;
	/* the post's rates at the optical-depth fractions and the end point
	 * (dms_rate_fraction_slot; mixed precision: dms_rate_fraction_slot_mx,
	 * widened), when have_rates, filled by a second helper wave */
	int32_t have_rates;
	int32_t pad;
	double rt[DMS_MAX_OPTICAL_DEPTH_PIECES + 2];
};

template <bool MX = false>
DMS_HD void dms_pipe_make_seed(const DmsLane& L, uint32_t job, DmsPipeSeed& s)
{
	s.prop = L.prop;
	s.time_steps = L.time_steps;
	s.rk_all = L.rk_all;
	s.time_origin = L.time_origin;
	s.job = job;
	s.redo = 0;
	dms_step_event<MX>(L.prop, s.ev);
}

/* The seed of a redo session: the lane after dms_step_post rejected its step
 * on optical depth (returned -2 with od_retries > 0), before its own
 * dms_step_prep. */
DMS_HD void dms_pipe_make_redo_seed(const DmsLane& L, uint32_t job, DmsPipeSeed& s)
{
	s.prop = L.prop;
	s.time_steps = L.time_steps;
	s.rk_all = L.rk_all;
	s.time_origin = L.time_origin;
	s.job = job;
	s.redo = 1;
	s.eb.t = L.eb_t;
	for(int c = 0; c < 3; c++)
	{
		s.eb.r[c] = L.eb_r[c];
		s.eb.v[c] = L.eb_v[c];
	}
	s.eb.speed = L.v_before;
	s.r_before = L.r_before;
	s.rate_before = L.rate_before;
	s.rate_before_f = L.rate_before_f;
	s.pad = 0;
}

/* A helper's lane state from a seed: the fields dms_step_prep and the RKF45
 * attempts read (for a redo seed, the step's start too: eb_valid and
 * rb_cached as the rejection left them). */
DMS_HD void dms_pipe_from_seed(DmsLane& A, const DmsPipeSeed& s)
{
	A.prop = s.prop;
	A.time_steps = s.time_steps;
	A.rk_all = s.rk_all;
	A.time_origin = s.time_origin;
	if(s.redo)
	{
		A.eb_t = s.eb.t;
		for(int c = 0; c < 3; c++)
		{
			A.eb_r[c] = s.eb.r[c];
			A.eb_v[c] = s.eb.v[c];
		}
		A.r_before = s.r_before;
		A.v_before = s.eb.speed;
		A.eb_valid = 1;
		A.rb_cached = 1;
		A.rate_before = s.rate_before;
		A.rate_before_f = s.rate_before_f;
	}
}

/* dms_step_post's optical-depth rejection: the lane goes on with
 * dms_step_prep from the same step start. */
DMS_HD bool dms_post_rejected(const DmsLane& L, int post_ret) { return (post_ret == -2) & (L.od_retries > 0) & (L.eb_valid != 0); }

/* Helper side: the updates of dms_step_post's DMS_POST_NEXT exit that the next
 * dms_step_prep reads, from the event ev of the propagator state (the rate at
 * the new point is recomputed by the preparation: Total_DM_Scattering_Rate
 * there is the value the post's end-point query caches). */
DMS_HD void dms_pipe_assume_next(DmsLane& A, const DmsStepEvent& ev)
{
	A.eb_t = ev.t;
	for(int c = 0; c < 3; c++)
	{
		A.eb_r[c] = ev.r[c];
		A.eb_v[c] = ev.v[c];
	}
	A.r_before = dms_prop_radius(A.prop);
	A.v_before = ev.speed;
	A.eb_valid = 1;
	A.rb_cached = 0;
	A.time_steps++;
	A.od_retries = 0;
}

/* Helper side: the entry of a step whose attempts ended with r >= 0. */
template <bool MX = false>
DMS_HD void dms_pipe_fill_entry(const DmsLane& A, int r, DmsPipeEntry& e)
{
	e.time = A.prop.time;
	e.radius = A.prop.radius;
	e.phi = A.prop.phi;
	e.vr = A.prop.vr;
	e.ts = A.prop.ts;
	e.pb_time = A.pb_time;
	e.pb_radius = A.pb_radius;
	e.pb_phi = A.pb_phi;
	e.pb_vr = A.pb_vr;
	e.pb_ts = A.pb_ts;
	e.rate_before = A.rate_before;
	e.rate_before_f = A.rate_before_f;
	e.rk_all = A.rk_all;
	e.rk_ok = r;
	e.inner_iter = A.inner_iter;
	e.kind = DMS_PIPE_STEP;
	e.have_rates = 0;
	dms_step_event<MX>(A.prop, e.ev);
}

/* Helper side: the next step's preparation and RKF45 attempts (as
 * dms_lane_advance_b and dms_lane_advance_a run them, the attempts with paired
 * stages as the helper wave runs them) into e.  Returns false when the
 * preparation ends the segment (e.kind DMS_PIPE_PREP_END).  This is the host
 * replay's helper: on the device the paired attempts need lanes 0 and 1 of a
 * wave running them together, as the helper wave in dms_kernels.hip does. */
template <bool MX = false>
DMS_HD bool dms_pipe_prepare_step(DmsLane& A, const DmsPointView& P, const DmsSolarView& S, const DmsJobView& J, DmsPipeEntry& e)
{
	if(dms_step_prep<MX>(A, P, S, J) >= 0)
	{
		e.kind = DMS_PIPE_PREP_END;
		return false;
	}
	int r;
	do
	{
		A.inner_iter++;
		A.rk_all++;
		r = MX ? dms_rk45_attempt_mx<true>(A.prop, S, A.inner_iter) : dms_rk45_attempt<true>(A.prop, S, A.inner_iter);
	} while(r < 0);
	dms_pipe_fill_entry<MX>(A, r, e);
	return true;
}

/* The inputs of the FP64 post's rate queries for the step of entry e (the
 * post's own steps up to dms_rates_at_fractions: interior interval, fractions,
 * cached start rate, end-point query).  Returns false when the post makes no
 * rate query for the step; query k is then dms_rate_fraction_slot(..., k, ...)
 * with these inputs. */
struct DmsPipeRateInputs
{
	double fr[DMS_MAX_OPTICAL_DEPTH_PIECES + 1];
	double r_after;
	int n;
	bool use_cached, want_next;
};

DMS_HD bool dms_pipe_rate_inputs(const DmsPipeEntry& e, const DmsJobView& J, DmsPipeRateInputs& q)
{
	q.r_after = (e.radius < 0.0) ? 0.0 : e.radius; /* dms_prop_radius */
	const double actual_dt = e.time - e.pb_time;
	double is = 0.0, ie = 0.0;
	if(!dms_interior_interval(e.eb.r, e.ev.r, e.r_before, q.r_after, is, ie))
		return false;
	if(!((ie > is) & (actual_dt > 0.0)))
		return false;
	const int fast = J.terminate_on_capture;
	const int od_count = fast ? DMS_CAPTURE_MODE_OPTICAL_DEPTH_PIECES : DMS_MAX_OPTICAL_DEPTH_PIECES;
	const double inv_pieces = fast ? 1.0 / DMS_CAPTURE_MODE_OPTICAL_DEPTH_PIECES : 1.0 / DMS_MAX_OPTICAL_DEPTH_PIECES;
	const double span = ie - is;
	for(int i = 0; i <= DMS_MAX_OPTICAL_DEPTH_PIECES; i++)
		q.fr[i] = (i <= od_count) ? is + (double) i * inv_pieces * span : 0.0;
	q.n = od_count + 1;
	q.use_cached = (e.r_before < DMS_rSun) & (is <= 1.0e-12);
	q.want_next = q.r_after < DMS_rSun;
	return true;
}

/* The mixed-precision post's rate-query inputs for the step of entry e (its
 * own steps up to dms_rates_at_fractions_mx: interior interval, FP32
 * fractions and chord end points, cached start rate, end-point query).
 * Returns false when the post makes no rate query for the step. */
struct DmsPipeRateInputsMx
{
	float fr[DMS_MAX_OPTICAL_DEPTH_PIECES + 1];
	float b[6], a[6];
	float xr, xv;
	int n;
	bool use_cached, want_next;
};

DMS_HD bool dms_pipe_rate_inputs_mx(const DmsPipeEntry& e, const DmsJobView& J, DmsPipeRateInputsMx& q)
{
	const double r_after = (e.radius < 0.0) ? 0.0 : e.radius; /* dms_prop_radius */
	const double actual_dt = e.time - e.pb_time;
	double is = 0.0, ie = 0.0;
	if(!dms_interior_interval(e.eb.r, e.ev.r, e.r_before, r_after, is, ie))
		return false;
	if(!(ie > is && actual_dt > 0.0))
		return false;
	const int fast = J.terminate_on_capture;
	const int od_count = fast ? DMS_CAPTURE_MODE_OPTICAL_DEPTH_PIECES : DMS_MAX_OPTICAL_DEPTH_PIECES;
	const double inv_pieces = fast ? 1.0 / DMS_CAPTURE_MODE_OPTICAL_DEPTH_PIECES : 1.0 / DMS_MAX_OPTICAL_DEPTH_PIECES;
	const float isf = (float) is, span = (float) (ie - is);
	for(int i = 0; i <= DMS_MAX_OPTICAL_DEPTH_PIECES; i++)
		q.fr[i] = (i <= od_count) ? isf + (float) i * (float) inv_pieces * span : 0.0f;
	for(int c = 0; c < 3; c++)
	{
		q.b[c] = (float) (e.eb.r[c] * DMS_MX_INV_RSUN);
		q.b[3 + c] = (float) e.eb.v[c];
		q.a[c] = (float) (e.ev.r[c] * DMS_MX_INV_RSUN);
		q.a[3 + c] = (float) e.ev.v[c];
	}
	q.xr = (float) (r_after * DMS_MX_INV_RSUN);
	q.xv = (float) e.ev.speed;
	q.n = od_count + 1;
	q.use_cached = e.r_before < DMS_rSun && is <= 1.0e-12;
	q.want_next = r_after < DMS_rSun;
	return true;
}

/* Query k of entry e in the mixed-precision kernel, widened to double. */
DMS_HD double dms_pipe_rate_mx(const DmsPipeEntry& e, const DmsPipeRateInputsMx& q, const DmsPointView& P, const DmsSolarView& S,
                               int k)
{
	return (double) dms_rate_fraction_slot_mx(P, S, k, q.b, q.a, q.fr, q.n, q.use_cached, e.rate_before_f, q.want_next, q.xr,
	                                          q.xv);
}

/* Query k of entry e (inputs from dms_pipe_rate_inputs). */
DMS_HD double dms_pipe_rate(const DmsPipeEntry& e, const DmsPipeRateInputs& q, const DmsPointView& P, const DmsSolarView& S, int k)
{
	return dms_rate_fraction_slot(P, S, k, e.eb.r, e.eb.v, e.ev.r, e.ev.v, q.fr, q.n, q.use_cached, e.rate_before,
	                              q.want_next, q.r_after, e.ev.speed);
}

/* Trajectory side, after its post returned DMS_POST_NEXT: take the helper's
 * step as if dms_step_prep and the RKF45 attempts had run here.  The fields the
 * preparation would also write (eb, r_before, v_before, rb_cached) already hold
 * the values it would write, from the post. */
DMS_HD void dms_pipe_take(DmsLane& L, const DmsPipeEntry& e)
{
	L.prop.time = e.time;
	L.prop.radius = e.radius;
	L.prop.phi = e.phi;
	L.prop.vr = e.vr;
	L.prop.ts = e.ts;
	L.pb_time = e.pb_time;
	L.pb_radius = e.pb_radius;
	L.pb_phi = e.pb_phi;
	L.pb_vr = e.pb_vr;
	L.pb_ts = e.pb_ts;
	L.rate_before = e.rate_before;
	L.rate_before_f = e.rate_before_f;
	L.rk_all = e.rk_all;
	L.rk_ok = e.rk_ok;
	L.inner_iter = e.inner_iter;
}

#endif
