/* CPU ORACLE -- TEST INFRASTRUCTURE ONLY (see dms_oracle.h).
 * Trajectory engine restated from src/Simulation_Trajectory.cpp:
 * Free_Particle_Propagator (:1403-1742), Propagate_Freely (:762-1119),
 * Scatter and its samplers (:1122-1269), capture bookkeeping (:640-760) and
 * Simulate (:1276-1400).  One scalar FP64 thread, same operation order as the
 * reference source. */
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

/* constants, src/Simulation_Trajectory.cpp:29-43 */
#define RK45_MIN_STEP_FACTOR 0.1
#define RK45_MAX_STEP_FACTOR 4.0
#define RK45_MAX_INNER_RETRIES 256
#define FREE_ENERGY_DRIFT_REL_E_SCALE_EV 1.0e-30
#define MAX_OPTICAL_DEPTH_STEP 0.05
#define CAPTURE_MODE_MAX_OPTICAL_DEPTH_STEP 0.10
#define OPTICAL_DEPTH_RELATIVE_TOLERANCE 1.0e-2
#define OPTICAL_DEPTH_ABSOLUTE_TOLERANCE (1.0e-12 * MAX_OPTICAL_DEPTH_STEP)
#define MAX_OPTICAL_DEPTH_RETRIES 100u
#define MAX_OPTICAL_DEPTH_PIECES 4
#define CAPTURE_MODE_OPTICAL_DEPTH_PIECES 2
#define TARGET_VELOCITY_MAX_REJECTION_ATTEMPTS 10000UL
#define OR_V_MAX 0.75

/* ---- Free_Particle_Propagator ------------------------------------------- */
struct or_prop
{
	double time, radius, phi, v_radial, J;
	double ax[3], ay[3], az[3];
	int radial_mode;
	double time_step;
	double tol[3];
	uint64_t attempts;
};

static double rk45_sanitized(double step)
{
	if(!dm_isfinite(step) || step <= 0.0)
		return 0.1 * OR_sec;
	return dm_min(step, 1.0e6 * OR_sec);
}

/* Build_Stable_Orthonormal_Basis, :130-142 */
static void stable_basis(const double* axis, double* e1, double* e2, double* e3)
{
	const double n = or_norm(axis);
	double t[3] = {axis[0] / n, axis[1] / n, axis[2] / n};
	const double ref[3] = {(dm_fabs(t[2]) < 0.9) ? 0.0 : 1.0, 0.0, (dm_fabs(t[2]) < 0.9) ? 1.0 : 0.0};
	double c[3];
	e1[0] = t[0];
	e1[1] = t[1];
	e1[2] = t[2];
	or_cross(ref, e1, c);
	or_normalize(c, e2);
	or_cross(e1, e2, c);
	or_normalize(c, e3);
}

/* constructor, :1403-1464 */
static void prop_init(struct or_prop* p, const dms_event* e)
{
	memset(p, 0, sizeof(*p));
	p->time = e->t;
	p->radius = or_norm(e->r);
	p->phi = 0.0;
	p->radial_mode = 0;
	p->time_step = 0.1 * OR_sec;
	const double speed = or_norm(e->v);
	const double radius_eps = 1.0e-12 * OR_km;
	const double angular_scale = dm_max(p->radius * speed, 1.0 * OR_km * OR_km / OR_sec);
	const double angular_eps = 1.0e-12 * angular_scale;
	double L[3];
	or_cross(e->r, e->v, L);
	const double Lnorm = or_norm(L);
	if(dm_isfinite(p->radius) && p->radius > radius_eps)
	{
		for(int c = 0; c < 3; c++)
			p->ax[c] = e->r[c] / p->radius;
		p->v_radial = or_dot(e->r, e->v) / p->radius;
	}
	else if(dm_isfinite(speed) && speed > 0.0)
	{
		for(int c = 0; c < 3; c++)
			p->ax[c] = e->v[c] / speed;
		p->v_radial = speed;
		p->radius = 0.0;
		p->radial_mode = 1;
	}
	else
	{
		p->ax[0] = 1.0;
		p->ax[1] = 0.0;
		p->ax[2] = 0.0;
		p->v_radial = 0.0;
		p->radius = 0.0;
		p->radial_mode = 1;
	}
	if(!p->radial_mode && dm_isfinite(Lnorm) && Lnorm > angular_eps)
	{
		for(int c = 0; c < 3; c++)
			p->az[c] = L[c] / Lnorm;
		or_cross(p->az, p->ax, p->ay);
		const double ayn = or_norm(p->ay);
		if(dm_isfinite(ayn) && ayn > 0.0)
		{
			for(int c = 0; c < 3; c++)
				p->ay[c] = p->ay[c] / ayn;
			double t[3];
			or_cross(p->ay, p->az, t);
			or_normalize(t, p->ax);
			p->J = Lnorm;
		}
		else
			p->radial_mode = 1;
	}
	else
		p->radial_mode = 1;
	if(p->radial_mode)
	{
		double axis[3] = {p->ax[0], p->ax[1], p->ax[2]};
		stable_basis(axis, p->ax, p->ay, p->az);
		p->J = 0.0;
	}
	p->tol[0] = 1.0 * OR_km;
	p->tol[1] = 1.0e-3 * OR_km / OR_sec;
	p->tol[2] = 1.0e-7;
}

static double prop_dv_dt(const struct or_prop* p, double r, double mass)
{
	if(r < 1.0 * OR_km)
		r = 1.0 * OR_km;
	return p->J * p->J / r / r / r - OR_G_Newton * mass / r / r;
}

static double prop_dphi_dt(const struct or_prop* p, double r)
{
	if(p->radial_mode || p->J == 0.0 || r <= 0.0)
		return 0.0;
	return p->J / r / r;
}

/* RK45_Next_Step_Size / RK45_Errors_Within_Tolerance, :466-502 */
static int rk45_within(const double* e, const double* tol)
{
	for(int i = 0; i < 3; i++)
		if(!dm_isfinite(e[i]) || e[i] > tol[i])
			return 0;
	return 1;
}

static double rk45_next(double step, const double* e, const double* tol)
{
	for(int i = 0; i < 3; i++)
		if(!dm_isfinite(e[i]))
			return RK45_MIN_STEP_FACTOR * step;
	double factor = RK45_MAX_STEP_FACTOR;
	for(int i = 0; i < 3; i++)
	{
		if(e[i] <= 0.0)
			continue;
		const double candidate = 0.84 * dm_pow_quarter(tol[i] / e[i]);
		if(dm_isfinite(candidate))
			factor = dm_min(factor, candidate);
	}
	factor = dm_max(RK45_MIN_STEP_FACTOR, dm_min(factor, RK45_MAX_STEP_FACTOR));
	return rk45_sanitized(factor * step);
}

/* Runge_Kutta_45_Step, :1487-1602 (S == NULL: constant-mass overload :1605-1692) */
static int prop_rk45(struct or_prop* p, const or_solar* S, double constant_mass)
{
	p->time_step = rk45_sanitized(p->time_step);
	int inner_iter = 0;
	for(;;)
	{
		inner_iter++;
		p->attempts++;
		const double h = p->time_step;
		double kr[6], kv[6], kp[6], ri;
#define MASS(x) ((S != NULL) ? or_mass(S, (x)) : constant_mass)
		kr[0] = h * p->v_radial;
		kv[0] = h * prop_dv_dt(p, p->radius, MASS(p->radius));
		kp[0] = h * prop_dphi_dt(p, p->radius);

		ri = p->radius + kr[0] / 4.0;
		kr[1] = h * (p->v_radial + kv[0] / 4.0);
		kv[1] = h * prop_dv_dt(p, ri, MASS(ri));

		ri = p->radius + 3.0 / 32.0 * kr[0] + 9.0 / 32.0 * kr[1];
		kr[2] = h * (p->v_radial + 3.0 / 32.0 * kv[0] + 9.0 / 32.0 * kv[1]);
		kv[2] = h * prop_dv_dt(p, ri, MASS(ri));
		kp[2] = h * prop_dphi_dt(p, ri);

		ri = p->radius + 1932.0 / 2197.0 * kr[0] - 7200.0 / 2197.0 * kr[1] + 7296.0 / 2197.0 * kr[2];
		kr[3] = h * (p->v_radial + 1932.0 / 2197.0 * kv[0] - 7200.0 / 2197.0 * kv[1] + 7296.0 / 2197.0 * kv[2]);
		kv[3] = h * prop_dv_dt(p, ri, MASS(ri));
		kp[3] = h * prop_dphi_dt(p, ri);

		ri = p->radius + 439.0 / 216.0 * kr[0] - 8.0 * kr[1] + 3680.0 / 513.0 * kr[2] - 845.0 / 4104.0 * kr[3];
		kr[4] = h * (p->v_radial + 439.0 / 216.0 * kv[0] - 8.0 * kv[1] + 3680.0 / 513.0 * kv[2] - 845.0 / 4104.0 * kv[3]);
		kv[4] = h * prop_dv_dt(p, ri, MASS(ri));
		kp[4] = h * prop_dphi_dt(p, ri);

		ri = p->radius - 8.0 / 27.0 * kr[0] + 2.0 * kr[1] - 3544.0 / 2565.0 * kr[2] + 1859.0 / 4104.0 * kr[3] - 11.0 / 40.0 * kr[4];
		kr[5] = h * (p->v_radial - 8.0 / 27.0 * kv[0] + 2.0 * kv[1] - 3544.0 / 2565.0 * kv[2] + 1859.0 / 4104.0 * kv[3] - 11.0 / 40.0 * kv[4]);
		kv[5] = h * prop_dv_dt(p, ri, MASS(ri));
		kp[5] = h * prop_dphi_dt(p, ri);
#undef MASS
		const double r4 = p->radius + 25.0 / 216.0 * kr[0] + 1408.0 / 2565.0 * kr[2] + 2197.0 / 4104.0 * kr[3] - 1.0 / 5.0 * kr[4];
		const double v4 = p->v_radial + 25.0 / 216.0 * kv[0] + 1408.0 / 2565.0 * kv[2] + 2197.0 / 4104.0 * kv[3] - 1.0 / 5.0 * kv[4];
		const double p4 = p->phi + 25.0 / 216.0 * kp[0] + 1408.0 / 2565.0 * kp[2] + 2197.0 / 4104.0 * kp[3] - 1.0 / 5.0 * kp[4];
		const double r5 = p->radius + 16.0 / 135.0 * kr[0] + 6656.0 / 12825.0 * kr[2] + 28561.0 / 56430.0 * kr[3] - 9.0 / 50.0 * kr[4] + 2.0 / 55.0 * kr[5];
		const double v5 = p->v_radial + 16.0 / 135.0 * kv[0] + 6656.0 / 12825.0 * kv[2] + 28561.0 / 56430.0 * kv[3] - 9.0 / 50.0 * kv[4] + 2.0 / 55.0 * kv[5];
		const double p5 = p->phi + 16.0 / 135.0 * kp[0] + 6656.0 / 12825.0 * kp[2] + 28561.0 / 56430.0 * kp[3] - 9.0 / 50.0 * kp[4] + 2.0 / 55.0 * kp[5];
		const double err[3] = {dm_fabs(r5 - r4), dm_fabs(v5 - v4), dm_fabs(p5 - p4)};
		const double h_new = rk45_next(h, err, p->tol);
		if(rk45_within(err, p->tol))
		{
			p->time = p->time + h;
			p->radius = r4;
			if(p->radius < 0.0)
				p->radius = 0.0;
			p->v_radial = v4;
			p->phi = p4;
			p->time_step = h_new;
			return 1;
		}
		const double abs_min_step = 1.0e-8 * OR_sec;
		if(h_new < abs_min_step || inner_iter >= RK45_MAX_INNER_RETRIES)
		{
			const int state_ok = dm_isfinite(r4) && dm_isfinite(v4) && dm_isfinite(p4);
			p->time = p->time + abs_min_step;
			if(state_ok)
			{
				p->radius = r4;
				if(p->radius < 0.0)
					p->radius = 0.0;
				p->v_radial = v4;
				p->phi = p4;
			}
			p->time_step = abs_min_step;
			return 0;
		}
		p->time_step = h_new;
	}
}

static double prop_radius(const struct or_prop* p) { return (p->radius < 0.0) ? 0.0 : p->radius; }

static double prop_speed(const struct or_prop* p)
{
	const double r = (p->radius < 0.0) ? 0.0 : p->radius;
	if(r == 0 || p->J == 0)
		return dm_fabs(p->v_radial);
	return dm_sqrt(p->v_radial * p->v_radial + p->J * p->J / r / r);
}

/* Event_In_3D, :1715-1742 */
static void prop_event(const struct or_prop* p, dms_event* e)
{
	const double safe_radius = (dm_isfinite(p->radius) && p->radius > 0.0) ? p->radius : 0.0;
	e->t = p->time;
	if(p->radial_mode || p->J == 0.0 || safe_radius == 0.0)
	{
		for(int c = 0; c < 3; c++)
		{
			e->r[c] = safe_radius * p->ax[c];
			e->v[c] = p->v_radial * p->ax[c];
		}
		return;
	}
	const double cos_phi = dm_cos(p->phi);
	const double sin_phi = dm_sin(p->phi);
	const double tangential_speed = p->J / safe_radius;
	const double vx = p->v_radial * cos_phi - tangential_speed * sin_phi;
	const double vy = p->v_radial * sin_phi + tangential_speed * cos_phi;
	for(int c = 0; c < 3; c++)
	{
		e->r[c] = safe_radius * (cos_phi * p->ax[c] + sin_phi * p->ay[c]);
		e->v[c] = vx * p->ax[c] + vy * p->ay[c];
	}
}

or_prop* or_prop_create(const dms_event* e)
{
	struct or_prop* p = (struct or_prop*) malloc(sizeof(struct or_prop));
	prop_init(p, e);
	return p;
}
void or_prop_free(or_prop* p) { free(p); }
int or_prop_step_solar(or_prop* p, const or_solar* S) { return prop_rk45(p, S, 0.0); }
int or_prop_step_mass(or_prop* p, double m) { return prop_rk45(p, NULL, m); }
void or_prop_set_time_step(or_prop* p, double dt) { p->time_step = dt; }
void or_prop_state(const or_prop* p, double* time, double* radius, double* speed, double* time_step, dms_event* ev)
{
	*time = p->time;
	*radius = prop_radius(p);
	*speed = prop_speed(p);
	*time_step = p->time_step;
	prop_event(p, ev);
}

/* ---- helpers, :53-464 ---------------------------------------------------- */
static double free_propagation_cap(double radius, double speed, double maximum_distance)
{
	double cap = 1.0e6 * OR_sec;
	const double safe_speed = dm_max(dm_fabs(speed), 1.0e-12 * OR_km / OR_sec);
	const double crossing_scale = dm_max(0.25 * maximum_distance, 10.0 * OR_km);
	cap = dm_min(cap, crossing_scale / safe_speed);
	const double safe_radius = dm_max(radius, 1.0 * OR_km);
	const double dynamical_time = dm_sqrt(safe_radius * safe_radius * safe_radius / (OR_G_Newton * OR_mSun));
	if(dm_isfinite(dynamical_time) && dynamical_time > 0.0)
		cap = dm_min(cap, 0.1 * dynamical_time);
	return dm_max(cap, 1.0e-8 * OR_sec);
}

static int outward_escaping(const dms_event* e, const or_solar* S, double boundary)
{
	const double radius = or_norm(e->r);
	if(radius < boundary)
		return 0;
	const double vr = (radius > 0.0) ? or_dot(e->r, e->v) / radius : 0.0;
	return vr > 0.0 && or_norm(e->v) > or_local_escape_speed(S, radius);
}

/* Unit_Vector_At_Angle_From_Axis, :109-128; returns 0 on the reference's throw */
static int unit_vector_at_angle(double cos_theta, double phi, const double* axis, double* out)
{
	if(!dm_isfinite(cos_theta) || cos_theta < -1.0 - 1.0e-12 || cos_theta > 1.0 + 1.0e-12 || !dm_isfinite(phi))
		return 0;
	const double axis_norm = or_norm(axis);
	if(!dm_isfinite(axis_norm) || axis_norm <= 0.0)
		return 0;
	const double e3[3] = {axis[0] / axis_norm, axis[1] / axis_norm, axis[2] / axis_norm};
	const double ref[3] = {(dm_fabs(e3[2]) < 0.9) ? 0.0 : 1.0, 0.0, (dm_fabs(e3[2]) < 0.9) ? 1.0 : 0.0};
	double t[3], e1[3], e2[3];
	or_cross(ref, e3, t);
	or_normalize(t, e1);
	or_cross(e3, e1, t);
	or_normalize(t, e2);
	cos_theta = dm_clamp_cos(cos_theta);
	const double sin_theta = dm_sqrt(dm_max(0.0, 1.0 - cos_theta * cos_theta));
	const double cp = dm_cos(phi), sp = dm_sin(phi);
	for(int c = 0; c < 3; c++)
		out[c] = cos_theta * e3[c] + sin_theta * cp * e1[c] + sin_theta * sp * e2[c];
	return 1;
}

/* Bound_Kepler_Return_At_Same_Radius, :152-201 */
static int bound_kepler_return(const dms_event* o, dms_event* in)
{
	const double radius = or_norm(o->r);
	const double speed = or_norm(o->v);
	const double vr = (radius <= 0.0) ? 0.0 : or_dot(o->r, o->v) / radius;
	if(!dm_isfinite(radius) || !dm_isfinite(speed) || !dm_isfinite(vr) || radius <= OR_rSun || speed <= 0.0 || vr <= 0.0)
		return 0;
	const double mu = OR_G_Newton * OR_mSun;
	const double specific_energy = 0.5 * speed * speed - mu / radius;
	if(!dm_isfinite(specific_energy) || specific_energy >= 0.0)
		return 0;
	double L[3];
	or_cross(o->r, o->v, L);
	const double J = or_norm(L);
	if(!dm_isfinite(J))
		return 0;
	const double a = -mu / (2.0 * specific_energy);
	if(!dm_isfinite(a) || a <= 0.0)
		return 0;
	double e2 = 1.0 + 2.0 * specific_energy * J * J / (mu * mu);
	if(!dm_isfinite(e2) || e2 < 0.0)
		return 0;
	e2 = dm_max(0.0, e2);
	const double ecc = dm_sqrt(e2);
	if(ecc <= 1.0e-14 || ecc >= 1.0)
		return 0;
	const double cosE = dm_clamp_cos((1.0 - radius / a) / ecc);
	const double E = dm_acos(cosE);
	const double sinE = dm_sqrt(dm_max(0.0, 1.0 - cosE * cosE));
	const double n = dm_sqrt(mu / (a * a * a));
	if(!dm_isfinite(n) || n <= 0.0)
		return 0;
	const double return_time = (2.0 * DM_PI - 2.0 * E + 2.0 * ecc * sinE) / n;
	if(!dm_isfinite(return_time) || return_time <= 0.0)
		return 0;
	*in = *o;
	in->t += return_time;
	for(int c = 0; c < 3; c++)
	{
		const double ru = o->r[c] / radius;
		in->v[c] = o->v[c] - 2.0 * vr * ru;
	}
	return in->t > o->t && dm_isfinite(or_norm(in->v));
}

int or_bound_kepler_return(const dms_event* o, dms_event* in) { return bound_kepler_return(o, in); }

/* Interpolate_Event, :203-216 */
static void interpolate_event(const dms_event* b, const dms_event* a, double f, dms_event* out)
{
	f = dm_clamp_unit(f);
	for(int c = 0; c < 3; c++)
	{
		out->r[c] = b->r[c] + f * (a->r[c] - b->r[c]);
		out->v[c] = b->v[c] + f * (a->v[c] - b->v[c]);
	}
	out->t = b->t + f * (a->t - b->t);
}

static double radius_at_fraction(const dms_event* b, const dms_event* a, double f)
{
	f = dm_clamp_unit(f);
	double rs = 0.0;
	for(int c = 0; c < 3; c++)
	{
		const double v = b->r[c] + f * (a->r[c] - b->r[c]);
		rs += v * v;
	}
	return dm_sqrt(dm_max(0.0, rs));
}

static int surface_crossings(const dms_event* b, const dms_event* a, double* first, double* second)
{
	double qa = 0.0, qb = 0.0, qc = -OR_rSun * OR_rSun;
	for(int c = 0; c < 3; c++)
	{
		const double d = a->r[c] - b->r[c];
		qa += d * d;
		qb += 2.0 * b->r[c] * d;
		qc += b->r[c] * b->r[c];
	}
	if(qa <= 0.0)
		return 0;
	const double disc = qb * qb - 4.0 * qa * qc;
	if(!dm_isfinite(disc) || disc < 0.0)
		return 0;
	const double sd = dm_sqrt(dm_max(0.0, disc));
	*first = (-qb - sd) / (2.0 * qa);
	*second = (-qb + sd) / (2.0 * qa);
	if(*first > *second)
	{
		const double t = *first;
		*first = *second;
		*second = t;
	}
	return 1;
}

static int fraction_in_unit(double f) { return f >= -1.0e-12 && f <= 1.0 + 1.0e-12; }

static double find_surface_crossing(const dms_event* b, const dms_event* a, double lower, double upper)
{
	double r_lower = radius_at_fraction(b, a, lower) - OR_rSun;
	for(int it = 0; it < 60; it++)
	{
		const double mid = 0.5 * (lower + upper);
		const double r_mid = radius_at_fraction(b, a, mid) - OR_rSun;
		if((r_lower <= 0.0 && r_mid <= 0.0) || (r_lower >= 0.0 && r_mid >= 0.0))
		{
			lower = mid;
			r_lower = r_mid;
		}
		else
			upper = mid;
	}
	return dm_clamp_unit(0.5 * (lower + upper));
}

/* Solar_Interior_Fraction_Interval, :282-325 */
static int interior_interval(const dms_event* b, const dms_event* a, double r_before, double r_after, double* start, double* end)
{
	const int before_inside = r_before < OR_rSun;
	const int after_inside = r_after < OR_rSun;
	double first = 0.0, second = 0.0;
	const int has = surface_crossings(b, a, &first, &second);
	if(before_inside && after_inside)
	{
		*start = 0.0;
		*end = 1.0;
		return 1;
	}
	if(!before_inside && after_inside)
	{
		if(has && fraction_in_unit(first))
			*start = dm_clamp_unit(first);
		else if(has && fraction_in_unit(second))
			*start = dm_clamp_unit(second);
		else
			*start = find_surface_crossing(b, a, 0.0, 1.0);
		*end = 1.0;
		return *end > *start;
	}
	if(before_inside && !after_inside)
	{
		*start = 0.0;
		if(has && fraction_in_unit(second))
			*end = dm_clamp_unit(second);
		else if(has && fraction_in_unit(first))
			*end = dm_clamp_unit(first);
		else
			*end = find_surface_crossing(b, a, 0.0, 1.0);
		return *end > *start;
	}
	if(has && fraction_in_unit(first) && fraction_in_unit(second))
	{
		*start = dm_clamp_unit(first);
		*end = dm_clamp_unit(second);
		return *end > *start;
	}
	return 0;
}

static double rate_at_fraction(const or_point* P, const dms_event* b, const dms_event* a, double f)
{
	f = dm_clamp_unit(f);
	double rs = 0.0, vs = 0.0;
	for(int c = 0; c < 3; c++)
	{
		const double pos = b->r[c] + f * (a->r[c] - b->r[c]);
		const double vel = b->v[c] + f * (a->v[c] - b->v[c]);
		rs += pos * pos;
		vs += vel * vel;
	}
	const double radius = dm_sqrt(dm_max(0.0, rs));
	if(radius >= OR_rSun)
		return 0.0;
	return or_rate_total(P, radius, dm_sqrt(dm_max(0.0, vs)));
}

typedef struct
{
	double start, end, rate_start, rate_end;
} od_piece;

/* Build_Optical_Depth_Pieces, :347-414 */
static int build_pieces(const or_point* P, const dms_event* b, const dms_event* a, double is, double ie, double actual_dt,
                        int use_cached, double cached_rate, int count, od_piece* pieces, int* piece_count,
                        double* tau_two, double* tau_four, int* invalid)
{
	*piece_count = 0;
	*tau_two = 0.0;
	*tau_four = 0.0;
	*invalid = 0;
	if(!(ie > is) || actual_dt <= 0.0)
		return 0;
	if(count > MAX_OPTICAL_DEPTH_PIECES)
		count = MAX_OPTICAL_DEPTH_PIECES;
	if(count < 1)
		count = 1;
	const double span = ie - is;
	double fr[MAX_OPTICAL_DEPTH_PIECES + 1], rt[MAX_OPTICAL_DEPTH_PIECES + 1];
	for(int i = 0; i <= count; i++)
		fr[i] = is + (double) i / (double) count * span;
	for(int i = 0; i <= count; i++)
	{
		if(i == 0 && use_cached)
			rt[i] = cached_rate;
		else
			rt[i] = rate_at_fraction(P, b, a, fr[i]);
		if(!dm_isfinite(rt[i]) || rt[i] < 0.0)
		{
			*invalid = 1;
			return 1;
		}
	}
#define TRAP(s, e) (0.5 * (rt[s] + rt[e]) * (actual_dt * (fr[e] - fr[s])))
	if(count >= 4)
		*tau_two = TRAP(0, 2) + TRAP(2, 4);
	else
		for(int i = 0; i < count; i++)
			*tau_two += TRAP(i, i + 1);
	for(int i = 0; i < count; i++)
	{
		if(fr[i + 1] > fr[i])
		{
			pieces[*piece_count].start = fr[i];
			pieces[*piece_count].end = fr[i + 1];
			pieces[*piece_count].rate_start = rt[i];
			pieces[*piece_count].rate_end = rt[i + 1];
			(*piece_count)++;
			*tau_four += TRAP(i, i + 1);
		}
	}
#undef TRAP
	return *piece_count > 0;
}

/* Solve_Linear_Rate_Optical_Depth_Fraction, :416-464 */
static double solve_linear_rate(double rs, double re, double dt, double target_tau)
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
		const double roots[2] = {(-qb + sd) / (2.0 * qa), (-qb - sd) / (2.0 * qa)};
		for(int k = 0; k < 2; k++)
			if(dm_isfinite(roots[k]) && roots[k] >= 0.0 && roots[k] <= 1.0)
				return dm_clamp_unit(roots[k]);
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

/* ---- Trajectory_Simulator --------------------------------------------------- */
typedef struct
{
	const or_sim_cfg* cfg;
	const or_point* P;
	dm_rng rng;
	double* dt_hist;
	double* v2dt_hist;
	uint32_t* counts;
	dms_traj_record* rec; /* bincount scalars live here */
	int is_captured, event_observed, boundary_escape_observed, survival_valid, numerically_invalid_escape;
	double prev_time_sec, prev_r_km, prev_v2, prev_dt_sec;
	double previous_capture_energy_eV, free_flight_reference_energy_eV;
	int bound_state;
	uint64_t rk_total;
} or_traj;

static double capture_energy_eV(const or_traj* T, double radius, double speed)
{
	const double vesc = or_local_escape_speed(T->P->S, radius);
	const double E = 0.5 * T->P->dm.mass * (speed * speed - vesc * vesc);
	return E / OR_eV;
}

/* Accumulate_Bincount_Step, :640-669 */
static void accumulate_bincount(or_traj* T, double r_km, double v2, double dt_sec)
{
	if(!dm_isfinite(dt_sec) || dt_sec <= 0.0)
		return;
	const double bin_max = 2.0 * OR_R_SUN_KM;
	int bin = -1;
	if(dm_isfinite(r_km) && r_km >= 0.0 && r_km < bin_max)
	{
		bin = (int) (r_km / T->cfg->bin_width_km);
		if(bin < 0)
			bin = 0;
		if(bin >= T->cfg->nbins)
			bin = -1;
	}
	if(bin >= 0)
	{
		const double v2dt = v2 * dt_sec;
		if(T->dt_hist)
		{
			T->dt_hist[bin] += dt_sec;
			T->v2dt_hist[bin] += v2dt;
		}
		if(T->counts)
			T->counts[bin] += 1u;
	}
}

/* Reset_Bincount_Anchor, :671-680 */
static void reset_anchor(or_traj* T, const dms_event* e)
{
	if(T->cfg->terminate_on_capture)
		return;
	T->prev_time_sec = e->t / OR_sec;
	T->prev_r_km = or_norm(e->r) / OR_km;
	const double v = or_norm(e->v) / (OR_km / OR_sec);
	T->prev_v2 = v * v;
}

/* Update_Capture_State, :689-760 */
static int update_capture_state(or_traj* T, double radius, double speed, double time, int allow_new)
{
	if(T->cfg->terminate_on_capture && !allow_new && !T->is_captured)
		return 0;
	const double E = capture_energy_eV(T, radius, speed);
	double dE = dm_nan();
	if(dm_isfinite(T->previous_capture_energy_eV))
		dE = E - T->previous_capture_energy_eV;
	T->previous_capture_energy_eV = E;
	const int negative = E < 0.0;
	const double t_now = time / OR_sec;
	dms_traj_record* R = T->rec;
	if(!allow_new)
	{
		if(T->is_captured && dm_isfinite(T->free_flight_reference_energy_eV))
		{
			const double drift = dm_fabs(E - T->free_flight_reference_energy_eV);
			if(dm_isfinite(drift) && drift > R->max_free_energy_drift_eV)
				R->max_free_energy_drift_eV = drift;
			const double scale = dm_max(dm_fabs(T->free_flight_reference_energy_eV), FREE_ENERGY_DRIFT_REL_E_SCALE_EV);
			const double rel = drift / scale;
			if(dm_isfinite(rel) && rel > R->max_free_energy_drift_rel)
				R->max_free_energy_drift_rel = rel;
		}
		if(T->bound_state && T->is_captured)
			R->t_last_bound_sec = t_now;
		return 0;
	}
	const int was_bound = T->bound_state;
	T->bound_state = negative;
	T->free_flight_reference_energy_eV = E;
	if(negative)
	{
		if(!T->is_captured)
		{
			T->is_captured = 1;
			R->t_capture_sec = t_now;
			R->t_last_bound_sec = t_now;
			R->r_first_negative_km = radius / OR_km;
			R->E_first_negative_eV = E;
			R->dE_first_negative_from_prev_eV = dE;
		}
		else if(!was_bound)
			R->t_final_unbinding_scatter_sec = dm_nan();
		R->t_last_bound_sec = t_now;
		return 1;
	}
	if(was_bound && T->is_captured)
		R->t_final_unbinding_scatter_sec = t_now;
	return 0;
}

/* Propagate_Freely, :762-1119.  *ev is the absolute event (in/out). */
static int propagate_freely(or_traj* T, dms_event* ev)
{
	const or_sim_cfg* cfg = T->cfg;
	const or_point* P = T->P;
	const or_solar* S = P->S;
	if(outward_escaping(ev, S, cfg->max_distance))
		return DMS_OUTWARD_ESCAPE;
#define ABORT_IF_UNCAPTURED_BOUND(evp)                                                 \
	(!T->is_captured && !(dm_isfinite(capture_energy_eV(T, or_norm((evp)->r), or_norm((evp)->v))) \
	                      && capture_energy_eV(T, or_norm((evp)->r), or_norm((evp)->v)) >= 0.0))
	if(ABORT_IF_UNCAPTURED_BOUND(ev))
		return DMS_NUMERICAL_FAILURE;
	double time_origin = ev->t;
	dms_event local = *ev;
	local.t = 0.0;
	struct or_prop prop;
	prop_init(&prop, &local);
	double minus_log_xi = -dm_log(dm_max(dm_uniform(&T->rng), DM_DBL_MIN));
	int outcome = DMS_MAX_FREE_STEPS;
	uint64_t time_steps = 0;
	unsigned int od_retries = 0;
	const int fast = cfg->terminate_on_capture;
	const double od_limit = fast ? CAPTURE_MODE_MAX_OPTICAL_DEPTH_STEP : MAX_OPTICAL_DEPTH_STEP;
	const int od_count = fast ? CAPTURE_MODE_OPTICAL_DEPTH_PIECES : MAX_OPTICAL_DEPTH_PIECES;
	if(!cfg->terminate_on_capture && T->prev_time_sec >= 0.0)
		T->prev_time_sec = 0.0;

#define TO_ABS(src, dst) \
	do                   \
	{                    \
		*(dst) = *(src); \
		(dst)->t += time_origin; \
	} while(0)
#define WALL_EXCEEDED() (cfg->rk_budget > 0 && (T->rk_total + prop.attempts) > cfg->rk_budget)

	while(time_steps < cfg->max_time_steps && outcome == DMS_MAX_FREE_STEPS)
	{
		OR_OPS_STEP_BEGIN();
		if(WALL_EXCEEDED())
		{
			dms_event e3;
			prop_event(&prop, &e3);
			TO_ABS(&e3, ev);
			T->rk_total += prop.attempts;
			return DMS_WALL_TIME_LIMIT;
		}
		dms_event eb, ea;
		prop_event(&prop, &eb);
		const double r_before = prop_radius(&prop);
		const double v_before = prop_speed(&prop);
		double rate_before = 0.0;
		if(r_before >= OR_rSun)
		{
			const double cap = free_propagation_cap(r_before, v_before, cfg->max_distance);
			prop.time_step = dm_min(rk45_sanitized(prop.time_step), cap);
			const double rvb = (r_before > 0.0) ? or_dot(eb.r, eb.v) / r_before : 0.0;
			if(r_before > OR_rSun && rvb < 0.0)
			{
				const double est = (r_before - OR_rSun) / (-rvb);
				if(dm_isfinite(est) && est > 0.0)
					prop.time_step = dm_min(prop.time_step, dm_max(est, 1.0e-8 * OR_sec));
			}
		}
		else
		{
			prop.time_step = rk45_sanitized(prop.time_step);
			rate_before = or_rate_total(P, r_before, v_before);
			if(!dm_isfinite(rate_before) || rate_before < 0.0)
			{
				dms_event e3;
				prop_event(&prop, &e3);
				TO_ABS(&e3, ev);
				T->rk_total += prop.attempts;
				return DMS_NUMERICAL_FAILURE;
			}
			if(rate_before > 0.0)
				prop.time_step = dm_min(prop.time_step, od_limit / rate_before);
		}
		const struct or_prop prop_before = prop;
		const double t_before = prop.time;
		OR_OPS_PUSH(OR_OPS_RK);
		const int ok = prop_rk45(&prop, S, 0.0);
		OR_OPS_POP();
		const double actual_dt = prop.time - t_before;
		const double r_after = prop_radius(&prop);
		const double v_after = prop_speed(&prop);
		prop_event(&prop, &ea);
		if(!ok)
		{
			TO_ABS(&ea, ev);
			T->rk_total += prop.attempts;
			return DMS_NUMERICAL_FAILURE;
		}
		if(!dm_isfinite(r_after) || !dm_isfinite(v_after) || !dm_isfinite(actual_dt) || actual_dt <= 0.0)
		{
			TO_ABS(&ea, ev);
			T->rk_total += prop.attempts;
			return DMS_NON_FINITE_STATE;
		}
		if(v_after > OR_V_MAX)
		{
			TO_ABS(&ea, ev);
			T->rk_total += prop.attempts;
			return DMS_SPEED_LIMIT;
		}
		if(WALL_EXCEEDED())
		{
			TO_ABS(&ea, ev);
			T->rk_total += prop.attempts;
			return DMS_WALL_TIME_LIMIT;
		}
		int scattering = 0, reflection = 0;
		dms_event accepted = ea;
		double is = 0.0, ie = 0.0;
		if(interior_interval(&eb, &ea, r_before, r_after, &is, &ie))
		{
			T->rec->n_steps_interior++;
			OR_OPS_STEP_CLASS(OR_OPS_INTERIOR);
			if(r_after < OR_rSun && v_after < 0.0)
				break;
			od_piece pieces[MAX_OPTICAL_DEPTH_PIECES];
			int pc = 0, invalid = 0;
			double tau2 = 0.0, dtau = 0.0;
			const int use_cached = r_before < OR_rSun && is <= 1.0e-12;
			build_pieces(P, &eb, &ea, is, ie, actual_dt, use_cached, rate_before, od_count, pieces, &pc, &tau2, &dtau, &invalid);
			if(invalid || !dm_isfinite(dtau) || dtau < 0.0)
			{
				TO_ABS(&ea, ev);
				T->rk_total += prop.attempts;
				return DMS_NUMERICAL_FAILURE;
			}
			const double scale = dm_max(dtau, OPTICAL_DEPTH_ABSOLUTE_TOLERANCE);
			const double relerr = dm_fabs(dtau - tau2) / scale;
			const int rej_od = dtau > od_limit;
			const int rej_acc = !fast && relerr > OPTICAL_DEPTH_RELATIVE_TOLERANCE;
			if(rej_od || rej_acc)
			{
				T->rec->n_tau_rejections++;
				od_retries++;
				if(od_retries > MAX_OPTICAL_DEPTH_RETRIES)
				{
					TO_ABS(&eb, ev);
					T->rk_total += prop.attempts;
					return DMS_NUMERICAL_FAILURE;
				}
				const uint64_t attempts_so_far = prop.attempts;
				prop = prop_before;
				prop.attempts = attempts_so_far;
				double factor = RK45_MAX_STEP_FACTOR;
				if(dtau > od_limit && dtau > 0.0)
					factor = dm_min(factor, 0.8 * od_limit / dtau);
				if(rej_acc && relerr > 0.0)
					factor = dm_min(factor, 0.8 * dm_sqrt(OPTICAL_DEPTH_RELATIVE_TOLERANCE / relerr));
				factor = dm_max(RK45_MIN_STEP_FACTOR, dm_min(factor, 1.0));
				prop.time_step = rk45_sanitized(factor * actual_dt);
				continue;
			}
			if(dtau >= minus_log_xi)
			{
				double remaining = minus_log_xi;
				for(int k = 0; k < pc; k++)
				{
					const double piece_dt = actual_dt * (pieces[k].end - pieces[k].start);
					const double piece_tau = 0.5 * (pieces[k].rate_start + pieces[k].rate_end) * piece_dt;
					if(piece_tau >= remaining)
					{
						const double lf = solve_linear_rate(pieces[k].rate_start, pieces[k].rate_end, piece_dt, remaining);
						const double fraction = pieces[k].start + lf * (pieces[k].end - pieces[k].start);
						interpolate_event(&eb, &ea, fraction, &accepted);
						scattering = 1;
						outcome = DMS_SCATTER;
						break;
					}
					remaining -= piece_tau;
				}
				if(!scattering)
				{
					interpolate_event(&eb, &ea, ie, &accepted);
					scattering = 1;
					outcome = DMS_SCATTER;
				}
			}
			else
				minus_log_xi -= dtau;
		}
		else
		{
			T->rec->n_steps_exterior++;
			OR_OPS_STEP_CLASS(OR_OPS_EXTERIOR);
		}
		if(!scattering)
		{
			dms_event be = ea;
			double rb = r_after, vb = v_after;
			if(r_before < cfg->max_distance && r_after >= cfg->max_distance && r_after > r_before)
			{
				const double bf = (cfg->max_distance - r_before) / (r_after - r_before);
				interpolate_event(&eb, &ea, bf, &be);
				rb = or_norm(be.r);
				vb = or_norm(be.v);
			}
			if(rb >= cfg->max_distance && rb >= r_before && vb > or_local_escape_speed(S, rb))
			{
				accepted = be;
				reflection = 1;
				outcome = DMS_OUTWARD_ESCAPE;
			}
			else if(rb >= cfg->max_distance && rb >= r_before)
			{
				dms_event abs_b;
				TO_ABS(&be, &abs_b);
				if(ABORT_IF_UNCAPTURED_BOUND(&abs_b))
				{
					*ev = abs_b;
					T->rk_total += prop.attempts;
					return DMS_NUMERICAL_FAILURE;
				}
				dms_event inbound;
				if(bound_kepler_return(&be, &inbound))
				{
					time_steps++;
					od_retries = 0;
					/* commit_accepted_event(boundary_event) */
					{
						dms_event abs_e;
						TO_ABS(&be, &abs_e);
						const double t_now = be.t / OR_sec;
						if(!cfg->terminate_on_capture)
						{
							const double r_now = or_norm(be.r) / OR_km;
							const double v_now = or_norm(be.v) / (OR_km / OR_sec);
							const double v2 = v_now * v_now;
							if(T->prev_time_sec >= 0.0)
							{
								const double dts = t_now - T->prev_time_sec;
								accumulate_bincount(T, T->prev_r_km, T->prev_v2, dts);
								T->prev_dt_sec = dts;
							}
							T->prev_time_sec = t_now;
							T->prev_r_km = r_now;
							T->prev_v2 = v2;
						}
						update_capture_state(T, or_norm(abs_e.r), or_norm(abs_e.v), abs_e.t, 0);
					}
					TO_ABS(&inbound, ev);
					/* reset_propagator_at_absolute_event */
					T->rk_total += prop.attempts;
					time_origin = ev->t;
					local = *ev;
					local.t = 0.0;
					prop_init(&prop, &local);
					if(!cfg->terminate_on_capture && T->prev_time_sec >= 0.0)
						T->prev_time_sec = 0.0;
					continue;
				}
			}
		}
		time_steps++;
		od_retries = 0;
		int captured_now;
		{
			dms_event abs_e;
			TO_ABS(&accepted, &abs_e);
			const double t_now = accepted.t / OR_sec;
			if(!cfg->terminate_on_capture)
			{
				const double r_now = or_norm(accepted.r) / OR_km;
				const double v_now = or_norm(accepted.v) / (OR_km / OR_sec);
				const double v2 = v_now * v_now;
				if(T->prev_time_sec >= 0.0)
				{
					const double dts = t_now - T->prev_time_sec;
					accumulate_bincount(T, T->prev_r_km, T->prev_v2, dts);
					T->prev_dt_sec = dts;
				}
				T->prev_time_sec = t_now;
				T->prev_r_km = r_now;
				T->prev_v2 = v2;
			}
			captured_now = update_capture_state(T, or_norm(abs_e.r), or_norm(abs_e.v), abs_e.t, 0);
		}
		TO_ABS(&accepted, ev);
		if(ABORT_IF_UNCAPTURED_BOUND(ev))
		{
			T->rk_total += prop.attempts;
			return DMS_NUMERICAL_FAILURE;
		}
		if(cfg->terminate_on_capture && captured_now)
		{
			T->rk_total += prop.attempts;
			return DMS_CAPTURE_MODE;
		}
		if(reflection || scattering)
			break;
	}
	T->rk_total += prop.attempts;
	return outcome;
#undef TO_ABS
#undef WALL_EXCEEDED
#undef ABORT_IF_UNCAPTURED_BOUND
}

/* Sample_Target, :1122-1158.  Returns -1 electron, >= 0 isotope, -2 on error. */
static int sample_target(const or_point* P, dm_rng* rng, double r, double v)
{
	if(r > OR_rSun)
		return -2;
	double rates[OR_MAX_ISO];
	const int n = P->S->n_iso;
	for(int i = 0; i < n; i++)
	{
		rates[i] = or_rate_nucleus(P, r, v, i);
		if(!dm_isfinite(rates[i]) || rates[i] < 0.0)
			return -2;
	}
	const double re = or_rate_electron(P, r, v);
	if(!dm_isfinite(re) || re < 0.0)
		return -2;
	double total = re;
	for(int i = 0; i < n; i++)
		total = total + rates[i];
	if(!dm_isfinite(total) || total <= 0.0)
		return -2;
	const double xi = dm_uniform(rng);
	double sum = re / total;
	if(sum > xi)
		return -1;
	for(int i = 0; i < n; i++)
	{
		sum += rates[i] / total;
		if(sum > xi || i == n - 1)
			return i;
	}
	return -2;
}

/* Sample_Target_Velocity (Romano & Walsh), :1160-1226 */
static int sample_target_velocity(dm_rng* rng, double temperature, double target_mass, const double* vel_DM, double* out,
                                  uint64_t* iterations)
{
	if(!dm_isfinite(temperature) || temperature <= 0.0 || !dm_isfinite(target_mass) || target_mass <= 0.0)
		return 0;
	const double kappa = dm_sqrt(target_mass / 2.0 / temperature);
	const double vDM = or_norm(vel_DM);
	if(!dm_isfinite(kappa) || kappa <= 0.0 || !dm_isfinite(vDM) || vDM <= 0.0)
		return 0;
	const double y = kappa * vDM;
	double x = y, mu = 1.0;
	unsigned long attempts = 0;
	for(;;)
	{
		(*iterations)++;
		const double xs = x * x, ys = y * y;
		const double raw = xs + ys - 2.0 * x * y * mu;
		const double den = x + y;
		const double roundoff = dm_max(1.0, xs + ys + 2.0 * dm_fabs(x * y));
		if(!dm_isfinite(x) || x < 0.0 || !dm_isfinite(y) || y <= 0.0 || !dm_isfinite(mu) || mu < -1.0 || mu > 1.0
		   || !dm_isfinite(raw) || raw < -1.0e-12 * roundoff || !dm_isfinite(den) || den <= 0.0)
			return 0;
		const double rel = dm_max(0.0, raw);
		const double ratio = dm_sqrt(rel) / den;
		if(!dm_isfinite(ratio) || ratio < 0.0 || ratio > 1.0 + 1.0e-12)
			return 0;
		const double accept = dm_min(1.0, ratio);
		if(!(accept < dm_uniform(rng)))
			break;
		attempts++;
		if(attempts > TARGET_VELOCITY_MAX_REJECTION_ATTEMPTS)
			return 0;
		mu = -1.0 + 2.0 * dm_uniform(rng);
		const double xi_1 = dm_uniform(rng);
		if(xi_1 < 2.0 / (DM_SQRT_PI * y + 2.0))
		{
			const double xi_2 = dm_max(dm_uniform(rng), DM_DBL_MIN);
			const double xi_3 = dm_max(dm_uniform(rng), DM_DBL_MIN);
			const double z = -dm_log(xi_2) - dm_log(xi_3);
			x = dm_sqrt(z);
		}
		else
		{
			const double xi_2 = dm_max(dm_uniform(rng), DM_DBL_MIN);
			const double xi_3 = dm_uniform(rng);
			const double xi_4 = dm_max(dm_uniform(rng), DM_DBL_MIN);
			const double c = dm_cos(DM_PI_2 * xi_3);
			const double z = -dm_log(xi_2) - c * c * dm_log(xi_4);
			x = dm_sqrt(z);
		}
	}
	const double vT = x / kappa;
	const double phi = 2.0 * DM_PI * dm_uniform(rng);
	double u[3];
	if(!unit_vector_at_angle(mu, phi, vel_DM, u))
		return 0;
	for(int c = 0; c < 3; c++)
		out[c] = vT * u[c];
	return dm_isfinite(or_norm(out));
}

/* DM CM-angle samplers: low-mass contact 2 xi - 1 (obscura); dark photon
 * src/Dark_Photon.cpp:286-325 */
static double sample_angle(const or_point* P, dm_rng* rng, int target, double v)
{
	const or_dm* dm = &P->dm;
	const double xi = dm_uniform(rng);
	if(dm->interaction == DMS_DARK_PHOTON && dm->ff == DMS_FF_GENERAL)
	{
		const double m2 = dm->m_dark_photon * dm->m_dark_photon;
		const double mu = (target < 0) ? dm->mu_e : dm->mass * P->S->m[target] / (dm->mass + P->S->m[target]);
		const double q2max = 4.0 * ((mu * v) * (mu * v));
		return (m2 * (2.0 * xi - 1.0) + q2max * xi) / (m2 + q2max * xi);
	}
	return 2.0 * xi - 1.0;
}

/* New_DM_Velocity, :1228-1247 */
static int new_dm_velocity(dm_rng* rng, double cos_a, double mDM, double mT, double* vel, const double* vT)
{
	if(!dm_isfinite(cos_a) || cos_a < -1.0 - 1.0e-12 || cos_a > 1.0 + 1.0e-12 || !dm_isfinite(mDM) || mDM <= 0.0
	   || !dm_isfinite(mT) || mT <= 0.0 || !dm_isfinite(or_norm(vel)) || !dm_isfinite(or_norm(vT)))
		return 0;
	cos_a = dm_clamp_cos(cos_a);
	const double phi = 2.0 * DM_PI * dm_uniform(rng);
	double n[3];
	if(!unit_vector_at_angle(cos_a, phi, vel, n))
		return 0;
	const double d[3] = {vT[0] - vel[0], vT[1] - vel[1], vT[2] - vel[2]};
	const double rel = or_norm(d);
	const double sum_m = mT + mDM;
	const double f = mT * rel / sum_m;
	double out[3];
	for(int c = 0; c < 3; c++)
		out[c] = f * n[c] + (mDM * vel[c] + mT * vT[c]) / (mT + mDM);
	if(!dm_isfinite(or_norm(out)))
		return 0;
	for(int c = 0; c < 3; c++)
		vel[c] = out[c];
	return 1;
}

/* Scatter, :1249-1269.  Returns 0 on success, -1 where the reference throws. */
int or_scatter_impl(const or_point* P, dms_event* e, dm_rng* rng, uint64_t* tv_iterations)
{
	const double r = or_norm(e->r);
	const double v = or_norm(e->v);
	const int target = sample_target(P, rng, r, v);
	if(target < -1)
		return -1;
	const double mT = (target == -1) ? OR_mElectron : P->S->m[target];
	double vT[3];
	if(!sample_target_velocity(rng, or_temperature(P->S, r), mT, e->v, vT, tv_iterations))
		return -1;
	const double cos_a = sample_angle(P, rng, target, v);
	if(!new_dm_velocity(rng, cos_a, P->dm.mass, mT, e->v, vT))
		return -1;
	return 0;
}

int or_scatter(const or_point* P, dms_event* e, uint32_t seed, uint32_t point_id, uint64_t traj_id)
{
	dm_rng rng;
	dm_rng_init(&rng, seed, point_id, traj_id, DMS_STREAM_SIM);
	uint64_t it = 0;
	return or_scatter_impl(P, e, &rng, &it);
}

static int invalidates_survival(int reason)
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
		case DMS_UNKNOWN: return 1;
		default: return 0;
	}
}

/* Simulate, :1276-1400 (plus the Particle_Free / Particle_Reflected and 1 AU
 * shift classification of Data_Generation.cpp:812-846 as flags). */
int or_simulate_one(const or_sim_cfg* cfg, const dms_event* ic, uint32_t seed, uint32_t point_id, uint64_t traj_id,
                    dms_traj_record* rec, double* dt_hist, double* v2dt_hist, uint32_t* counts)
{
	const or_point* P = cfg->P;
	memset(rec, 0, sizeof(*rec));
	rec->initial = *ic;
	rec->final = *ic;
	rec->trajectory_id = traj_id;
	rec->t_capture_sec = -1.0;
	rec->t_last_bound_sec = -1.0;
	rec->t_final_unbinding_scatter_sec = dm_nan();
	rec->t_boundary_escape_sec = dm_nan();
	rec->t_termination_sec = -1.0;
	rec->r_first_negative_km = -1.0;
	rec->E_first_negative_eV = dm_nan();
	rec->dE_first_negative_from_prev_eV = dm_nan();
	const double r0 = or_norm(ic->r), v0 = or_norm(ic->v);
	if(!dm_isfinite(ic->t) || ic->t < 0.0 || !dm_isfinite(r0) || r0 <= 0.0 || !dm_isfinite(v0) || v0 <= 0.0)
		return DMS_ERR_ARG;
	or_traj T;
	memset(&T, 0, sizeof(T));
	T.cfg = cfg;
	T.P = P;
	T.rec = rec;
	T.dt_hist = dt_hist;
	T.v2dt_hist = v2dt_hist;
	T.counts = counts;
	T.survival_valid = 1;
	dm_rng_init(&T.rng, seed, point_id, traj_id, DMS_STREAM_SIM);
	T.prev_time_sec = -1.0;
	T.prev_r_km = 0.0;
	T.prev_v2 = 0.0;
	T.prev_dt_sec = 0.0;
	T.previous_capture_energy_eV = capture_energy_eV(&T, r0, v0);
	T.free_flight_reference_energy_eV = dm_nan();
	T.bound_state = 0;

	dms_event cur = *ic;
	uint64_t n_scat = 0;
	int reason = DMS_UNKNOWN;
	while(n_scat < cfg->max_scatterings)
	{
		const int pr = propagate_freely(&T, &cur);
		OR_OPS_STEP_END();
		if(pr == DMS_SCATTER && or_norm(cur.r) < OR_rSun)
		{
			OR_OPS_PUSH(OR_OPS_SCATTER);
			const int sr = or_scatter_impl(P, &cur, &T.rng, &rec->n_tv_iterations);
			OR_OPS_POP();
			if(sr != 0)
			{
				reason = DMS_NUMERICAL_FAILURE;
				break;
			}
			n_scat++;
			const int cap = update_capture_state(&T, or_norm(cur.r), or_norm(cur.v), cur.t, 1);
			reset_anchor(&T, &cur);
			if(cfg->terminate_on_capture && cap)
			{
				reason = DMS_CAPTURE_MODE;
				break;
			}
		}
		else
		{
			reason = pr;
			break;
		}
	}
	if(reason == DMS_UNKNOWN && n_scat >= cfg->max_scatterings)
		reason = DMS_MAX_SCATTERINGS;
	rec->n_scatterings = n_scat;
	rec->t_termination_sec = cur.t / OR_sec;
	if(T.is_captured)
	{
		const double rf = or_norm(cur.r), vf = or_norm(cur.v);
		const double vesc = or_local_escape_speed(P->S, rf);
		const double Ef = 0.5 * P->dm.mass * (vf * vf - vesc * vesc);
		const double Ef_eV = Ef / OR_eV;
		if(reason == DMS_OUTWARD_ESCAPE && Ef_eV >= 0.0)
		{
			if(dm_isfinite(rec->t_final_unbinding_scatter_sec))
			{
				T.boundary_escape_observed = 1;
				rec->t_boundary_escape_sec = rec->t_termination_sec;
				T.event_observed = 1;
			}
			else
			{
				reason = DMS_ENERGY_DRIFT_ESCAPE;
				T.survival_valid = 0;
				T.numerically_invalid_escape = 1;
				T.boundary_escape_observed = 0;
				T.event_observed = 0;
			}
		}
	}
	if(invalidates_survival(reason))
		T.survival_valid = 0;
	if(T.is_captured)
	{
		if(!T.event_observed)
		{
			rec->t_final_unbinding_scatter_sec = dm_nan();
			if(reason != DMS_CAPTURE_MODE)
				T.survival_valid = 0;
		}
	}
	rec->final = cur;
	rec->reason = reason;
	rec->n_rk_attempts = T.rk_total;
	uint32_t flags = 0;
	if(T.is_captured)
		flags |= DMS_F_CAPTURED;
	if(T.event_observed)
		flags |= DMS_F_EVENT_OBSERVED;
	if(T.boundary_escape_observed)
		flags |= DMS_F_BOUNDARY_ESCAPE_OBSERVED;
	if(T.survival_valid)
		flags |= DMS_F_SURVIVAL_VALID;
	if(T.numerically_invalid_escape)
		flags |= DMS_F_NUMERICALLY_INVALID_ESCAPE;
	/* Particle_Free / Particle_Reflected, :529-542 */
	if(reason == DMS_OUTWARD_ESCAPE)
	{
		if(n_scat == 0)
			flags |= DMS_F_FREE;
		const double rf = or_norm(cur.r);
		const double vesc = dm_sqrt(2 * OR_G_Newton * OR_mSun / rf);
		if(rf > OR_rSun && or_norm(cur.v) > vesc && n_scat > 0)
		{
			flags |= DMS_F_REFLECTED;
			dms_event far = cur;
			if(!or_kepler_shift_impl(&far, 1.0 * OR_AU))
				flags |= DMS_F_REFLECTION_SHIFT_FAILED;
		}
	}
	rec->flags = flags;
	return DMS_OK;
}

/* ---- sampler hooks for the closed-form checks of the test suite ------------
 * Each draws n samples from one Philox stream (key {seed, 0}, trajectory 0). */
void or_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { dm_philox4x32_10(ctr, key, out); }

void or_sample_angles(const or_point* P, uint32_t seed, int target, double v, size_t n, double* out)
{
	dm_rng rng;
	dm_rng_init(&rng, seed, 0, 0, DMS_STREAM_SIM);
	for(size_t i = 0; i < n; i++)
		out[i] = sample_angle(P, &rng, target, v);
}

void or_sample_targets(const or_point* P, uint32_t seed, double r, double v, size_t n, int32_t* out)
{
	dm_rng rng;
	dm_rng_init(&rng, seed, 0, 0, DMS_STREAM_SIM);
	for(size_t i = 0; i < n; i++)
		out[i] = sample_target(P, &rng, r, v);
}

/* out: 3n components; returns the number of failed draws */
size_t or_sample_target_velocities(uint32_t seed, double temperature, double target_mass, const double* vel_DM, size_t n,
                                   double* out)
{
	dm_rng rng;
	dm_rng_init(&rng, seed, 0, 0, DMS_STREAM_SIM);
	size_t failed = 0;
	uint64_t it = 0;
	for(size_t i = 0; i < n; i++)
		if(!sample_target_velocity(&rng, temperature, target_mass, vel_DM, out + 3 * i, &it))
			failed++;
	return failed;
}

int or_new_dm_velocity(uint32_t seed, double cos_a, double mDM, double mT, double* vel, const double* vT)
{
	dm_rng rng;
	dm_rng_init(&rng, seed, 0, 0, DMS_STREAM_SIM);
	return new_dm_velocity(&rng, cos_a, mDM, mT, vel, vT);
}

/* ---- test hooks: engine helpers exposed for the oracle's known-answer checks -- */
double or_solve_linear_rate(double rs, double re, double dt, double target_tau)
{
	return solve_linear_rate(rs, re, dt, target_tau);
}

int or_interior_interval(const dms_event* b, const dms_event* a, double r_before, double r_after, double* start, double* end)
{
	*start = 0.0;
	*end = 0.0;
	return interior_interval(b, a, r_before, r_after, start, end);
}
