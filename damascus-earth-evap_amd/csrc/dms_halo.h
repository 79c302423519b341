/*
 * Standard Halo Model and the initial-condition generator as
 * __host__ __device__ functions: Initial_Conditions, PDF_Initial_Speed and
 * PDF_Cos_Theta (src/Simulation_Utilities.cpp:123-249 of the reference) over
 * obscura's truncated Maxwellian (un-vendored; N_esc = erf(z) - 2 z e^{-z^2}/sqrt(pi)).
 */
#ifndef DMS_HALO_H
#define DMS_HALO_H

#include "dms_physics.h"
#include "philox.h"
#include "../../include/dms.h"

struct DmsHaloView
{
	double v0, vesc, vobs[3], vobs_norm, rho, Nesc;
	double norm;  /* <v> + v_esc(R_sun)^2 <1/v>  (PDF_Initial_Speed normalisation) */
	double vescR; /* Local_Escape_Speed(R_sun) */
	double vesc_asym; /* Local_Escape_Speed(1000 AU) */
};

DMS_HD void dms_halo_init(DmsHaloView& H, const dms_halo& h)
{
	const double kms = DMS_km / DMS_sec;
	H.v0 = h.v0_kms * kms;
	H.vesc = h.vesc_kms * kms;
	for(int c = 0; c < 3; c++)
		H.vobs[c] = h.vobs_kms[c] * kms;
	H.vobs_norm = dms_norm(H.vobs);
	H.rho = h.rho_GeV_cm3 * DMS_GeV / (DMS_cm * DMS_cm * DMS_cm);
	const double z = H.vesc / H.v0;
	H.Nesc = dm_erf(z) - 2.0 * z * dm_exp(-z * z) / DM_SQRT_PI;
	H.norm = 0.0;
	H.vescR = 0.0;
	H.vesc_asym = 0.0;
}

/* Standard_Halo_Model::PDF_Speed */
DMS_HD double dms_shm_pdf_speed(const DmsHaloView& H, double v)
{
	const double vo = H.vobs_norm;
	if(!(v >= 0.0) || v >= H.vesc + vo)
		return 0.0;
	const double v02 = H.v0 * H.v0;
	if(vo == 0.0)
	{
		if(v >= H.vesc)
			return 0.0;
		const double K = DM_PI * v02;
		return 4.0 * DM_PI * v * v * dm_exp(-(v * v) / v02) / (H.Nesc * (K * dm_sqrt(K)));
	}
	const double pref = v / (DM_SQRT_PI * H.v0 * vo * H.Nesc);
	const double a = dm_exp(-((v - vo) * (v - vo)) / v02);
	const double b = (v < H.vesc - vo) ? dm_exp(-((v + vo) * (v + vo)) / v02) : dm_exp(-(H.vesc * H.vesc) / v02);
	return pref * (a - b);
}

/* Standard_Halo_Model::PDF_Velocity (lab-frame velocity, galactic v + v_obs) */
DMS_HD double dms_shm_pdf_velocity(const DmsHaloView& H, const double* vel)
{
	const double g[3] = {vel[0] + H.vobs[0], vel[1] + H.vobs[1], vel[2] + H.vobs[2]};
	const double g2 = dms_dot(g, g);
	if(dm_sqrt(g2) >= H.vesc)
		return 0.0;
	const double K = DM_PI * H.v0 * H.v0;
	return dm_exp(-g2 / (H.v0 * H.v0)) / (H.Nesc * (K * dm_sqrt(K)));
}

/* <v> (power > 0) or <1/v>: 10-point Gauss-Legendre on 64 panels of each of
 * [0, vesc - vobs] and [vesc - vobs, vesc + vobs]. */
DMS_HD double dms_shm_moment(const DmsHaloView& H, int power)
{
	const double vo = H.vobs_norm;
	const double edges[3] = {0.0, H.vesc - vo, H.vesc + vo};
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
				const double f = dms_shm_pdf_speed(H, v);
				sum = sum + DM_GL10_W[k] * ((power > 0) ? f * v : f / v);
			}
			total = total + half * sum;
		}
	}
	return total;
}

/* libphysica::Spherical_Coordinates(r, theta, phi, axis) */
DMS_HD void dms_spherical(double r, double theta, double phi, const double* axis, double* out)
{
	const double an = dms_norm(axis);
	const double e3[3] = {axis[0] / an, axis[1] / an, axis[2] / an};
	double e1[3], e2[3];
	dms_frame(e3, e1, e2);
	const double st = dm_sin(theta), ct = dm_cos(theta), cp = dm_cos(phi), sp = dm_sin(phi);
	for(int c = 0; c < 3; c++)
		out[c] = r * ((st * cp * e1[c] + st * sp * e2[c]) + ct * e3[c]);
}

DMS_HD double dms_pdf_initial_speed(const DmsHaloView& H, double v)
{
	if(v == 0.0)
		return 0.0;
	const double sp = dms_shm_pdf_speed(H, v);
	return (sp * v + (sp / v) * H.vescR * H.vescR) / H.norm;
}

DMS_HD double dms_pdf_cos_theta(const DmsHaloView& H, double c, double v)
{
	const double sp = dms_shm_pdf_speed(H, v);
	const double normalization = 2.0 * DM_PI * v * v / sp;
	const double zaxis[3] = {0.0, 0.0, 1.0};
	const double* axis = (H.vobs_norm > 0.0) ? H.vobs : zaxis;
	double vel[3];
	dms_spherical(v, dm_acos(c), 0.0, axis, vel);
	return normalization * dms_shm_pdf_velocity(H, vel);
}

/* Initial_Conditions (src/Simulation_Utilities.cpp:176-249) on Philox stream 1. */
DMS_HD void dms_initial_condition(const DmsHaloView& H, dm_rng* rng, dms_event* out)
{
	const double vmin = 0.0, vmax = H.vesc + H.vobs_norm;
	double u;
	for(;;)
	{
		const double x = vmin + (vmax - vmin) * dm_uniform(rng);
		const double y = 1200.0 * dm_uniform(rng);
		if(y <= dms_pdf_initial_speed(H, x))
		{
			u = x;
			break;
		}
	}
	const double vsun = H.vobs_norm;
	const double vgal = vmax - vsun;
	const double cos_max = (vsun == 0.0) ? 1.0 : dm_min(1.0, (vgal * vgal - vsun * vsun - u * u) / (2.0 * u * vsun));
	const double ymax = dms_pdf_cos_theta(H, -1.0, u);
	double ct;
	for(;;)
	{
		const double x = -1.0 + (cos_max - -1.0) * dm_uniform(rng);
		const double y = ymax * dm_uniform(rng);
		if(y <= dms_pdf_cos_theta(H, x, u))
		{
			ct = x;
			break;
		}
	}
	const double phi = 2.0 * DM_PI * dm_uniform(rng);
	const double zaxis[3] = {0.0, 0.0, 1.0};
	const double* axis = (vsun > 0.0) ? H.vobs : zaxis;
	double vel[3];
	dms_spherical(u, dm_acos(dm_clamp_cos(ct)), phi, axis, vel);
	const double D = 1000.0 * DMS_AU;
	const double v = dm_sqrt(u * u + H.vesc_asym * H.vesc_asym);
	const double vn = dms_norm(vel);
	for(int c = 0; c < 3; c++)
		vel[c] = v * (vel[c] / vn);
	const double bmax = dm_sqrt(u * u + H.vescR * H.vescR) / v * DMS_rSun;
	const double vn2 = dms_norm(vel);
	const double ez[3] = {-1.0 * (vel[0] / vn2), -1.0 * (vel[1] / vn2), -1.0 * (vel[2] / vn2)};
	double ex[3], ey[3];
	dms_frame(ez, ex, ey);
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

#endif
