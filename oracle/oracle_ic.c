/* CPU ORACLE -- TEST INFRASTRUCTURE ONLY (see dms_oracle.h).
 * Standard Halo Model (obscura, un-vendored: the truncated Maxwellian with
 * N_esc = erf(z) - 2 z exp(-z^2)/sqrt(pi)), initial conditions and the
 * hyperbolic Kepler shift, restated from src/Simulation_Utilities.cpp:123-321
 * and src/Reflection_Spectrum.cpp:66-73. */
#include <string.h>

#include "oracle_internal.h"

typedef struct
{
	double v0, vesc, vobs[3], vobs_norm, rho, Nesc;
} or_shm;

static void shm_init(or_shm* H, const dms_halo* h)
{
	const double kms = OR_km / OR_sec;
	H->v0 = h->v0_kms * kms;
	H->vesc = h->vesc_kms * kms;
	for(int c = 0; c < 3; c++)
		H->vobs[c] = h->vobs_kms[c] * kms;
	H->vobs_norm = or_norm(H->vobs);
	H->rho = h->rho_GeV_cm3 * OR_GeV / (OR_cm * OR_cm * OR_cm);
	const double z = H->vesc / H->v0;
	H->Nesc = dm_erf(z) - 2.0 * z * dm_exp(-z * z) / DM_SQRT_PI;
}

static double shm_pdf_speed(const or_shm* H, double v)
{
	const double vo = H->vobs_norm;
	if(!(v >= 0.0) || v >= H->vesc + vo)
		return 0.0;
	const double v02 = H->v0 * H->v0;
	if(vo == 0.0)
	{
		if(v >= H->vesc)
			return 0.0;
		const double K = DM_PI * v02;
		return 4.0 * DM_PI * v * v * dm_exp(-(v * v) / v02) / (H->Nesc * (K * dm_sqrt(K)));
	}
	const double pref = v / (DM_SQRT_PI * H->v0 * vo * H->Nesc);
	const double a = dm_exp(-((v - vo) * (v - vo)) / v02);
	const double b = (v < H->vesc - vo) ? dm_exp(-((v + vo) * (v + vo)) / v02) : dm_exp(-(H->vesc * H->vesc) / v02);
	return pref * (a - b);
}

static double shm_pdf_velocity(const or_shm* H, const double* vel)
{
	const double g[3] = {vel[0] + H->vobs[0], vel[1] + H->vobs[1], vel[2] + H->vobs[2]};
	const double g2 = or_dot(g, g);
	if(dm_sqrt(g2) >= H->vesc)
		return 0.0;
	const double K = DM_PI * H->v0 * H->v0;
	return dm_exp(-g2 / (H->v0 * H->v0)) / (H->Nesc * (K * dm_sqrt(K)));
}

/* <v> and <1/v> of the speed distribution: 10-point Gauss-Legendre on 64
 * panels of [0, vesc - vobs] and 64 panels of [vesc - vobs, vesc + vobs]. */
static double shm_moment(const or_shm* H, int power)
{
	const double vo = H->vobs_norm;
	const double edges[3] = {0.0, H->vesc - vo, H->vesc + vo};
	double total = 0.0;
	for(int s = 0; s < 2; s++)
	{
		const double a = edges[s], b = edges[s + 1];
		if(!(b > a))
			continue;
		const int panels = 64;
		const double w = (b - a) / panels;
		for(int p = 0; p < panels; p++)
		{
			const double lo = a + p * w;
			const double mid = lo + 0.5 * w;
			const double half = 0.5 * w;
			double sum = 0.0;
			for(int k = 0; k < 10; k++)
			{
				const double v = mid + half * DM_GL10_X[k];
				const double f = shm_pdf_speed(H, v);
				sum = sum + DM_GL10_W[k] * ((power > 0) ? f * v : f / v);
			}
			total = total + half * sum;
		}
	}
	return total;
}

double or_shm_pdf_speed(const dms_halo* h, double v)
{
	or_shm H;
	shm_init(&H, h);
	return shm_pdf_speed(&H, v);
}
double or_shm_average_speed(const dms_halo* h)
{
	or_shm H;
	shm_init(&H, h);
	return shm_moment(&H, 1);
}
double or_shm_inverse_average_speed(const dms_halo* h)
{
	or_shm H;
	shm_init(&H, h);
	return shm_moment(&H, -1);
}

/* DM_Entering_Rate, src/Reflection_Spectrum.cpp:66-73 */
double or_dm_entering_rate(const or_solar* S, const dms_halo* h, double mDM_GeV)
{
	or_shm H;
	shm_init(&H, h);
	const double n = H.rho / (mDM_GeV * OR_GeV);
	const double ua = shm_moment(&H, 1), ui = shm_moment(&H, -1);
	const double vesc = or_local_escape_speed(S, OR_rSun);
	return OR_rSun * OR_rSun * DM_PI * n * (ua + vesc * vesc * ui);
}

/* libphysica::Spherical_Coordinates(r, theta, phi, axis): polar angle theta
 * from `axis`, azimuth phi in a basis built like Unit_Vector_At_Angle_From_Axis. */
static void spherical(double r, double theta, double phi, const double* axis, double* out)
{
	const double an = or_norm(axis);
	const double e3[3] = {axis[0] / an, axis[1] / an, axis[2] / an};
	const double ref[3] = {(dm_fabs(e3[2]) < 0.9) ? 0.0 : 1.0, 0.0, (dm_fabs(e3[2]) < 0.9) ? 1.0 : 0.0};
	double t[3], e1[3], e2[3];
	or_cross(ref, e3, t);
	or_normalize(t, e1);
	or_cross(e3, e1, t);
	or_normalize(t, e2);
	const double st = dm_sin(theta), ct = dm_cos(theta), cp = dm_cos(phi), sp = dm_sin(phi);
	for(int c = 0; c < 3; c++)
		out[c] = r * ((st * cp * e1[c] + st * sp * e2[c]) + ct * e3[c]);
}

typedef struct
{
	const or_shm* H;
	double norm;   /* <v> + vesc(R)^2 <1/v> */
	double vescR;
} ic_ctx;

/* PDF_Initial_Speed, src/Simulation_Utilities.cpp:123-143 */
static double pdf_initial_speed(const ic_ctx* C, double v)
{
	if(v == 0.0)
		return 0.0;
	const double sp = shm_pdf_speed(C->H, v);
	return (sp * v + (sp / v) * C->vescR * C->vescR) / C->norm;
}

/* PDF_Cos_Theta, src/Simulation_Utilities.cpp:146-174 */
static double pdf_cos_theta(const or_shm* H, double c, double v)
{
	const double sp = shm_pdf_speed(H, v);
	const double normalization = 2.0 * DM_PI * v * v / sp;
	const double zaxis[3] = {0.0, 0.0, 1.0};
	const double* axis = (H->vobs_norm > 0.0) ? H->vobs : zaxis;
	double vel[3];
	spherical(v, dm_acos(c), 0.0, axis, vel);
	return normalization * shm_pdf_velocity(H, vel);
}

/* Initial_Conditions, src/Simulation_Utilities.cpp:176-249 */
static void ic_ctx_init(ic_ctx* C, const or_solar* S, const or_shm* H)
{
	C->H = H;
	C->vescR = or_local_escape_speed(S, OR_rSun);
	C->norm = shm_moment(H, 1) + C->vescR * C->vescR * shm_moment(H, -1);
}

static void make_ic(const or_solar* S, const ic_ctx* Cp, dm_rng* rng, dms_event* out)
{
	const ic_ctx C = *Cp;
	const or_shm* H = C.H;
	const double vmin = 0.0, vmax = H->vesc + H->vobs_norm;
	double u;
	for(;;)
	{
		const double x = vmin + (vmax - vmin) * dm_uniform(rng);
		const double y = 1200.0 * dm_uniform(rng);
		if(y <= pdf_initial_speed(&C, x))
		{
			u = x;
			break;
		}
	}
	const double vsun = H->vobs_norm;
	const double vgal = vmax - vsun;
	const double cos_max = (vsun == 0.0) ? 1.0 : dm_min(1.0, (vgal * vgal - vsun * vsun - u * u) / (2.0 * u * vsun));
	const double ymax = pdf_cos_theta(H, -1.0, u);
	double ct;
	for(;;)
	{
		const double x = -1.0 + (cos_max - -1.0) * dm_uniform(rng);
		const double y = ymax * dm_uniform(rng);
		if(y <= pdf_cos_theta(H, x, u))
		{
			ct = x;
			break;
		}
	}
	const double phi = 2.0 * DM_PI * dm_uniform(rng);
	const double zaxis[3] = {0.0, 0.0, 1.0};
	const double* axis = (vsun > 0.0) ? H->vobs : zaxis;
	double vel[3];
	spherical(u, dm_acos(dm_clamp_cos(ct)), phi, axis, vel);
	const double D = 1000.0 * OR_AU;
	const double vesc_asym = or_local_escape_speed(S, D);
	const double v = dm_sqrt(u * u + vesc_asym * vesc_asym);
	const double vn = or_norm(vel);
	for(int c = 0; c < 3; c++)
		vel[c] = v * (vel[c] / vn);
	const double vescR = or_local_escape_speed(S, OR_rSun);
	const double bmax = dm_sqrt(u * u + vescR * vescR) / v * OR_rSun;
	const double vn2 = or_norm(vel);
	const double ez[3] = {-1.0 * (vel[0] / vn2), -1.0 * (vel[1] / vn2), -1.0 * (vel[2] / vn2)};
	const double ref[3] = {(dm_fabs(ez[2]) < 0.9) ? 0.0 : 1.0, 0.0, (dm_fabs(ez[2]) < 0.9) ? 1.0 : 0.0};
	double t[3], ex[3], ey[3];
	or_cross(ref, ez, t);
	or_normalize(t, ex);
	or_cross(ez, ex, t);
	or_normalize(t, ey);
	const double phi_disk = 2.0 * DM_PI * dm_uniform(rng);
	const double xi = dm_uniform(rng);
	const double b = dm_sqrt(xi) * bmax;
	const double cp = dm_cos(phi_disk), sp = dm_sin(phi_disk);
	out->t = 0.0;
	for(int c = 0; c < 3; c++)
	{
		out->r[c] = D * ez[c] + b * (cp * ex[c] + sp * ey[c]);
		out->v[c] = vel[c];
	}
}

/* Hyperbolic_Kepler_Shift, src/Simulation_Utilities.cpp:252-321 */
int or_kepler_shift_impl(dms_event* e, double R_final)
{
	const double mu = OR_G_Newton * OR_mSun;
	const double Ri = or_norm(e->r);
	const double speed = or_norm(e->v);
	const double tol = 1.0e-10;
	if(R_final < OR_rSun || Ri < OR_rSun)
		return 0;
	if(!(dm_isfinite(R_final) && R_final > 0.0) || !(dm_isfinite(Ri) && Ri > 0.0) || !(dm_isfinite(speed) && speed > 0.0))
		return 0;
	if(dm_fabs(R_final - Ri) <= tol * dm_max(Ri, R_final))
		return 1;
	const double vr = or_dot(e->r, e->v) / Ri;
	if(!dm_isfinite(vr) || vr == 0.0)
		return 0;
	const double dir = (vr < 0.0) ? -1.0 : 1.0;
	if(dir < 0.0 && R_final > Ri)
		return 0;
	if(dir > 0.0 && R_final < Ri)
		return 0;
	const double energy = 0.5 * speed * speed - mu / Ri;
	if(!(dm_isfinite(energy) && energy > 0.0))
		return 0;
	const double rhat[3] = {e->r[0] / Ri, e->r[1] / Ri, e->r[2] / Ri};
	double hv[3];
	or_cross(e->r, e->v, hv);
	const double h = or_norm(hv);
	const double h_scale = Ri * speed;
	if(!dm_isfinite(h) || h <= tol * h_scale)
	{
		const double fss = 2.0 * (energy + mu / R_final);
		if(!(dm_isfinite(fss) && fss > 0.0))
			return 0;
		const double s = dir * dm_sqrt(fss);
		for(int c = 0; c < 3; c++)
		{
			e->r[c] = R_final * rhat[c];
			e->v[c] = s * rhat[c];
		}
		return 1;
	}
	const double p = h * h / mu;
	double vxh[3];
	or_cross(e->v, hv, vxh);
	const double ev[3] = {vxh[0] / mu - rhat[0], vxh[1] / mu - rhat[1], vxh[2] / mu - rhat[2]};
	const double ecc = or_norm(ev);
	if(!(dm_isfinite(p) && p > 0.0) || !dm_isfinite(ecc) || ecc <= 1.0)
		return 0;
	const double cos_raw = (p / R_final - 1.0) / ecc;
	if(!dm_isfinite(cos_raw) || cos_raw > 1.0 + tol || cos_raw < -1.0 - tol)
		return 0;
	double ax[3] = {ev[0] / ecc, ev[1] / ecc, ev[2] / ecc};
	const double az[3] = {hv[0] / h, hv[1] / h, hv[2] / h};
	double ay[3];
	or_cross(az, ax, ay);
	const double ayn = or_norm(ay);
	if(!(dm_isfinite(ayn) && ayn > 0.0))
		return 0;
	for(int c = 0; c < 3; c++)
		ay[c] = ay[c] / ayn;
	double t[3];
	or_cross(ay, az, t);
	or_normalize(t, ax);
	const double theta = dir * dm_acos(dm_clamp_cos(cos_raw));
	const double ct = dm_cos(theta), st = dm_sin(theta);
	const double sv = dm_sqrt(mu / p);
	for(int c = 0; c < 3; c++)
	{
		e->r[c] = R_final * (ct * ax[c] + st * ay[c]);
		e->v[c] = sv * (-st * ax[c] + (ecc + ct) * ay[c]);
	}
	return 1;
}

int or_hyperbolic_kepler_shift(dms_event* e, double R_final) { return or_kepler_shift_impl(e, R_final); }

void or_make_initial_condition(const or_solar* S, const dms_halo* h, dm_rng* rng, dms_event* out)
{
	or_shm H;
	shm_init(&H, h);
	ic_ctx C;
	ic_ctx_init(&C, S, &H);
	make_ic(S, &C, rng, out);
}

int or_initial_conditions(const or_solar* S, const dms_halo* h, uint32_t seed, uint32_t point_id, uint64_t first_id,
                          size_t n, double r_final_rsun, dms_event* out, int32_t* shift_ok)
{
	or_shm H;
	shm_init(&H, h);
	ic_ctx C;
	ic_ctx_init(&C, S, &H);
	for(size_t i = 0; i < n; i++)
	{
		dm_rng rng;
		dm_rng_init(&rng, seed, point_id, first_id + i, DMS_STREAM_IC);
		make_ic(S, &C, &rng, &out[i]);
		if(r_final_rsun > 0.0)
		{
			const int ok = or_kepler_shift_impl(&out[i], r_final_rsun * OR_rSun);
			if(shift_ok)
				shift_ok[i] = ok;
		}
		else if(shift_ok)
			shift_ok[i] = 1;
	}
	return DMS_OK;
}
