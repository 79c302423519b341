/* CPU ORACLE -- TEST INFRASTRUCTURE ONLY (see dms_oracle.h).
 * Solar model, cross sections, scattering rates and the rate grid, restated from
 * src/Solar_Model.cpp, src/Dark_Photon.cpp and obscura's low-mass SI formulas. */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

double or_constant(const char* name)
{
	static const struct
	{
		const char* name;
		double value;
	} table[] = {
	    {"cm", OR_cm}, {"meter", OR_meter}, {"km", OR_km}, {"sec", OR_sec}, {"gram", OR_gram},
	    {"kg", OR_kg}, {"Kelvin", OR_Kelvin}, {"eV", OR_eV}, {"MeV", OR_MeV}, {"GeV", OR_GeV},
	    {"mProton", OR_mProton}, {"mElectron", OR_mElectron}, {"mNucleon", OR_mNucleon},
	    {"aEM", OR_aEM}, {"G_Newton", OR_G_Newton}, {"mSun", OR_mSun}, {"rSun", OR_rSun},
	    {"AU", OR_AU}, {"pb", 1.0e-36 * OR_cm * OR_cm}};
	for(size_t i = 0; i < sizeof(table) / sizeof(table[0]); i++)
		if(strcmp(table[i].name, name) == 0)
			return table[i].value;
	return dm_nan();
}

/* ---- Steffen monotone cubic interpolation ---------------------------------
 * libphysica::Interpolation (used at src/Solar_Model.cpp:137-165) is a Steffen
 * spline (M. Steffen, A&A 239 (1990) 443): eqs. 11 (interior slopes) and
 * 26-27 (limited one-sided boundary slopes).  Evaluation in Horner form. */
static double or_sign(double v) { return (double) ((v > 0.0) - (v < 0.0)); }

static void or_spline_build(or_spline* s, int n, const double* x, const double* y, double pref)
{
	s->n = n;
	s->x = x;
	s->pref = pref;
	s->a = (double*) malloc(sizeof(double) * (size_t) (n - 1));
	s->b = (double*) malloc(sizeof(double) * (size_t) (n - 1));
	s->c = (double*) malloc(sizeof(double) * (size_t) (n - 1));
	s->d = (double*) malloc(sizeof(double) * (size_t) (n - 1));
	double* h = (double*) malloc(sizeof(double) * (size_t) (n - 1));
	double* sl = (double*) malloc(sizeof(double) * (size_t) (n - 1));
	double* yp = (double*) malloc(sizeof(double) * (size_t) n);
	for(int i = 0; i < n - 1; i++)
	{
		h[i] = x[i + 1] - x[i];
		sl[i] = (y[i + 1] - y[i]) / h[i];
	}
	for(int i = 1; i < n - 1; i++)
	{
		const double p = (sl[i - 1] * h[i] + sl[i] * h[i - 1]) / (h[i - 1] + h[i]);
		yp[i] = (or_sign(sl[i - 1]) + or_sign(sl[i])) * dm_min(dm_fabs(sl[i - 1]), dm_min(dm_fabs(sl[i]), 0.5 * dm_fabs(p)));
	}
	{
		const double p0 = sl[0] * (1.0 + h[0] / (h[0] + h[1])) - sl[1] * h[0] / (h[0] + h[1]);
		yp[0] = (p0 * sl[0] <= 0.0) ? 0.0 : ((dm_fabs(p0) > 2.0 * dm_fabs(sl[0])) ? 2.0 * sl[0] : p0);
		const int m = n - 2;
		const double pn = sl[m] * (1.0 + h[m] / (h[m] + h[m - 1])) - sl[m - 1] * h[m] / (h[m] + h[m - 1]);
		yp[n - 1] = (pn * sl[m] <= 0.0) ? 0.0 : ((dm_fabs(pn) > 2.0 * dm_fabs(sl[m])) ? 2.0 * sl[m] : pn);
	}
	for(int i = 0; i < n - 1; i++)
	{
		s->a[i] = (yp[i] + yp[i + 1] - 2.0 * sl[i]) / h[i] / h[i];
		s->b[i] = (3.0 * sl[i] - 2.0 * yp[i] - yp[i + 1]) / h[i];
		s->c[i] = yp[i];
		s->d[i] = y[i];
	}
	free(h);
	free(sl);
	free(yp);
}

static void or_spline_free(or_spline* s)
{
	free(s->a);
	free(s->b);
	free(s->c);
	free(s->d);
}

/* Interval j with x_j <= x < x_{j+1}, clamped to [0, n-2] (bisection). */
static int or_spline_interval(const or_spline* s, double x)
{
	int lo = 0, hi = s->n - 1;
	while(hi - lo > 1)
	{
		const int mid = (lo + hi) / 2;
		if(x < s->x[mid])
			hi = mid;
		else
			lo = mid;
	}
	return lo;
}

double or_spline_eval(const or_spline* s, double x)
{
	const int j = or_spline_interval(s, x);
	const double t = x - s->x[j];
	return s->pref * (((s->a[j] * t + s->b[j]) * t + s->c[j]) * t + s->d[j]);
}

/* ---- isotopes: src/Solar_Model.cpp:143-163 ---------------------------------
 * 10 explicit isotopes (H1 ... O18) plus all natural isotopes of Z = 10..28.
 * Masses follow obscura's Isotope convention (A == 1 ? m_p : A m_nucleon);
 * natural abundances are IUPAC representative values (obscura's own table is
 * not available offline: parity unpinned for Z >= 10 abundances). */
typedef struct
{
	int Z, A;
	double abundance;
} or_nat_iso;

static const or_nat_iso OR_NATURAL[] = {
    {10, 20, 0.9048}, {10, 21, 0.0027}, {10, 22, 0.0925},
    {11, 23, 1.0},
    {12, 24, 0.7899}, {12, 25, 0.1000}, {12, 26, 0.1101},
    {13, 27, 1.0},
    {14, 28, 0.92223}, {14, 29, 0.04685}, {14, 30, 0.03092},
    {15, 31, 1.0},
    {16, 32, 0.9499}, {16, 33, 0.0075}, {16, 34, 0.0425}, {16, 36, 0.0001},
    {17, 35, 0.7576}, {17, 37, 0.2424},
    {18, 36, 0.003336}, {18, 38, 0.000629}, {18, 40, 0.996035},
    {19, 39, 0.932581}, {19, 40, 0.000117}, {19, 41, 0.067302},
    {20, 40, 0.96941}, {20, 42, 0.00647}, {20, 43, 0.00135}, {20, 44, 0.02086}, {20, 46, 0.00004}, {20, 48, 0.00187},
    {21, 45, 1.0},
    {22, 46, 0.0825}, {22, 47, 0.0744}, {22, 48, 0.7372}, {22, 49, 0.0541}, {22, 50, 0.0518},
    {23, 50, 0.0025}, {23, 51, 0.9975},
    {24, 50, 0.04345}, {24, 52, 0.83789}, {24, 53, 0.09501}, {24, 54, 0.02365},
    {25, 55, 1.0},
    {26, 54, 0.05845}, {26, 56, 0.91754}, {26, 57, 0.02119}, {26, 58, 0.00282},
    {27, 59, 1.0},
    {28, 58, 0.68077}, {28, 60, 0.26223}, {28, 61, 0.011399}, {28, 62, 0.036346}, {28, 64, 0.009255}};

/* ---- Solar_Model::Solar_Model, src/Solar_Model.cpp:43-166 ------------------ */
or_solar* or_solar_create(const double* raw_rows, size_t n_rows)
{
	if(raw_rows == NULL || n_rows < 3)
		return NULL;
	or_solar* S = (or_solar*) calloc(1, sizeof(or_solar));
	const int n = (int) n_rows + 2;
	S->n = n;
	S->raw = (double*) malloc(sizeof(double) * (size_t) n * OR_NCOL);
	S->r = (double*) malloc(sizeof(double) * (size_t) n);
	/* Import_Raw_Data: units per column (src/Solar_Model.cpp:45-46) */
	double units[OR_NCOL];
	for(int c = 0; c < OR_NCOL; c++)
		units[c] = 1.0;
	units[0] = OR_mSun;
	units[1] = OR_rSun;
	units[2] = OR_Kelvin;
	units[3] = OR_gram / OR_cm / OR_cm / OR_cm;
	units[4] = OR_dyne / OR_cm / OR_cm;
	for(size_t i = 0; i < n_rows; i++)
		for(int c = 0; c < OR_NCOL; c++)
			S->raw[(i + 1) * OR_NCOL + c] = raw_rows[i * OR_NCOL + c] * units[c];
	/* padding rows (src/Solar_Model.cpp:47-65) */
	for(int c = 0; c < OR_NCOL; c++)
	{
		S->raw[c] = (c == 0 || c == 1) ? 0.0 : S->raw[1 * OR_NCOL + c];
		double last;
		if(c == 0)
			last = OR_mSun;
		else if(c == 1)
			last = OR_rSun;
		else if(c == 2)
			last = 5800 * OR_Kelvin;
		else if(c == 3)
			last = 1.0e-9 * OR_gram / OR_cm / OR_cm / OR_cm;
		else
			last = S->raw[(size_t) (n - 2) * OR_NCOL + c];
		S->raw[(size_t) (n - 1) * OR_NCOL + c] = last;
	}
	double* y = (double*) calloc((size_t) n, sizeof(double));
	for(int i = 0; i < n; i++)
		S->r[i] = S->raw[i * OR_NCOL + 1];
#define OR_COLUMN(col)                           \
	for(int i = 0; i < n; i++)                   \
		y[i] = S->raw[i * OR_NCOL + (col)];
	OR_COLUMN(0);
	or_spline_build(&S->mass, n, S->r, y, 1.0);
	OR_COLUMN(2);
	or_spline_build(&S->temperature, n, S->r, y, 1.0);
	/* Create_Escape_Speed_Table (src/Solar_Model.cpp:79-96): the integral of
	 * M(x)/x^2 from r to R_sun, here as 10-point Gauss-Legendre per profile
	 * interval, accumulated from the surface inward. */
	{
		double* Q = (double*) malloc(sizeof(double) * (size_t) n);
		for(int j = 0; j < n - 1; j++)
		{
			const double mid = 0.5 * (S->r[j] + S->r[j + 1]);
			const double half = 0.5 * (S->r[j + 1] - S->r[j]);
			double sum = 0.0;
			for(int k = 0; k < 10; k++)
			{
				const double xq = mid + half * DM_GL10_X[k];
				const double f = (xq == 0.0) ? 0.0 : or_mass(S, xq) / xq / xq;
				sum = sum + DM_GL10_W[k] * f;
			}
			Q[j] = half * sum;
		}
		double integral_above = 0.0;
		for(int i = n - 1; i >= 0; i--)
		{
			const double rr = S->r[i];
			double integral;
			if(rr < OR_rSun)
			{
				integral_above = Q[i] + integral_above;
				integral = integral_above;
			}
			else
				integral = OR_mSun * (1.0 / rr - 1.0 / OR_rSun);
			y[i] = 2.0 * OR_G_Newton * OR_mSun / OR_rSun * (1.0 + OR_rSun / OR_mSun * integral);
		}
		free(Q);
	}
	or_spline_build(&S->vesc2, n, S->r, y, 1.0);
	OR_COLUMN(3);
	or_spline_build(&S->rho, n, S->r, y, 1.0);
#undef OR_COLUMN
	/* isotopes and number densities (src/Solar_Model.cpp:98-110,143-163) */
	static const int Zs[29] = {1, 2, 2, 6, 6, 7, 7, 8, 8, 8, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28};
	static const int As[10] = {1, 4, 3, 12, 13, 14, 15, 16, 17, 18};
	int k = 0;
	for(int t = 0; t < 29; t++)
	{
		const int nat_count = (int) (sizeof(OR_NATURAL) / sizeof(OR_NATURAL[0]));
		for(int q = 0; q < (t < 10 ? 1 : nat_count); q++)
		{
			int Z, A;
			double ab;
			if(t < 10)
			{
				Z = Zs[t];
				A = As[t];
				ab = 1.0;
			}
			else
			{
				if(OR_NATURAL[q].Z != Zs[t])
					continue;
				Z = Zs[t];
				A = OR_NATURAL[q].A;
				ab = OR_NATURAL[q].abundance;
			}
			const double m = (A == 1) ? OR_mProton : A * OR_mNucleon;
			for(int i = 0; i < n; i++)
			{
				const double rho = S->raw[i * OR_NCOL + 3];
				const double f = S->raw[i * OR_NCOL + 6 + t];
				y[i] = f * rho / m;
			}
			S->Z[k] = Z;
			S->A[k] = A;
			S->m[k] = m;
			S->abundance[k] = ab;
			or_spline_build(&S->n_iso_spline[k], n, S->r, y, ab);
			k++;
		}
	}
	S->n_iso = k;
	/* Create_Number_Density_Table_Electron (src/Solar_Model.cpp:112-126) */
	for(int i = 0; i < n; i++)
		y[i] = 0.0;
	for(int q = 0; q < S->n_iso; q++)
		for(int i = 0; i < n; i++)
			y[i] += S->Z[q] * or_n_nucleus(S, S->r[i], q);
	or_spline_build(&S->n_e, n, S->r, y, 1.0);
	free(y);
	return S;
}

void or_solar_free(or_solar* S)
{
	if(S == NULL)
		return;
	or_spline_free(&S->mass);
	or_spline_free(&S->temperature);
	or_spline_free(&S->vesc2);
	or_spline_free(&S->rho);
	or_spline_free(&S->n_e);
	for(int i = 0; i < S->n_iso; i++)
		or_spline_free(&S->n_iso_spline[i]);
	free(S->raw);
	free(S->r);
	free(S);
}

/* src/Solar_Model.cpp:168-179 */
double or_mass(const or_solar* S, double r)
{
	if(!dm_isfinite(r))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(r > OR_rSun)
		return OR_mSun;
	return or_spline_eval(&S->mass, r);
}

/* src/Solar_Model.cpp:194-202 */
double or_temperature(const or_solar* S, double r)
{
	if(!dm_isfinite(r))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	return or_spline_eval(&S->temperature, r);
}

/* src/Solar_Model.cpp:204-218 */
double or_local_escape_speed(const or_solar* S, double r)
{
	if(!dm_isfinite(r))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(r > OR_rSun)
		return dm_sqrt(2 * OR_G_Newton * OR_mSun / r);
	return dm_sqrt(or_spline_eval(&S->vesc2, r));
}

static double or_mass_density(const or_solar* S, double r)
{
	if(!dm_isfinite(r))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(r > OR_rSun)
		return 0.0;
	return or_spline_eval(&S->rho, r);
}

/* Solar_Isotope::Number_Density, src/Solar_Model.cpp:28-39 */
double or_n_nucleus(const or_solar* S, double r, int i)
{
	if(!dm_isfinite(r))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(r > OR_rSun)
		return 0.0;
	return or_spline_eval(&S->n_iso_spline[i], r);
}

/* src/Solar_Model.cpp:250-261 */
double or_n_electron(const or_solar* S, double r)
{
	if(!dm_isfinite(r))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(r > OR_rSun)
		return 0.0;
	return or_spline_eval(&S->n_e, r);
}

double or_solar_eval(const or_solar* S, int which, double r)
{
	switch(which)
	{
		case 0: return or_mass(S, r);
		case 1: return or_temperature(S, r);
		case 2: return or_local_escape_speed(S, r);
		case 3: return or_mass_density(S, r);
		case 4: return or_n_electron(S, r);
		default:
			if(which >= 100 && which - 100 < S->n_iso)
				return or_n_nucleus(S, r, which - 100);
			return dm_nan();
	}
}

int or_solar_n_isotopes(const or_solar* S) { return S->n_iso; }

void or_solar_isotope(const or_solar* S, int i, int32_t* Z, int32_t* A, double* mass, double* abundance)
{
	*Z = S->Z[i];
	*A = S->A[i];
	*mass = S->m[i];
	*abundance = S->abundance[i];
}

/* Thermal_Averaged_Relative_Speed, src/Solar_Model.cpp:493-521 */
double or_thermal_averaged_relative_speed(double temperature, double mass_target, double v_DM)
{
	if(!dm_isfinite(temperature) || temperature <= 0.0 || !dm_isfinite(mass_target) || mass_target <= 0.0 || !dm_isfinite(v_DM) || v_DM < 0.0)
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

/* ---- DM particle -------------------------------------------------------------
 * SI low-mass contact (obscura DM_Particle_SI, un-vendored; the formula
 * sigma_N = sigma_p (mu_N/mu_p)^2 (Z + f_n/f_p (A-Z))^2 is the standard one,
 * parity unpinned) and the dark photon (src/Dark_Photon.cpp:123-213). */
static double or_reduced_mass(double a, double b) { return a * b / (a + b); }

static int or_dm_init(or_dm* dm, const or_solar* S, const dms_point_desc* d)
{
	memset(dm, 0, sizeof(*dm));
	if(!(d->mass_GeV > 0.0) || !dm_isfinite(d->mass_GeV) || d->low_mass != 1)
		return DMS_ERR_ARG;
	dm->interaction = d->interaction;
	dm->ff = d->form_factor;
	dm->mass = d->mass_GeV * OR_GeV;
	dm->fnfp = d->fn_over_fp;
	dm->mu_p = or_reduced_mass(dm->mass, OR_mProton);
	dm->mu_e = or_reduced_mass(dm->mass, OR_mElectron);
	if(d->interaction == DMS_SI)
	{
		if(d->form_factor != DMS_FF_CONTACT)
			return DMS_ERR_UNSUPPORTED;
		dm->sigma_p = d->sigma_p_cm2 * OR_cm * OR_cm;
		dm->sigma_e = d->sigma_e_cm2 * OR_cm * OR_cm;
		for(int i = 0; i < S->n_iso; i++)
		{
			const double mu_N = or_reduced_mass(dm->mass, S->m[i]);
			const double coupling = S->Z[i] + dm->fnfp * (S->A[i] - S->Z[i]);
			dm->sigma_N[i] = dm->sigma_p * mu_N * mu_N / dm->mu_p / dm->mu_p * coupling * coupling;
		}
		return DMS_OK;
	}
	if(d->interaction == DMS_DARK_PHOTON)
	{
		if(d->form_factor != DMS_FF_CONTACT && d->form_factor != DMS_FF_GENERAL)
			return DMS_ERR_UNSUPPORTED;
		const double alpha_dark = OR_aEM;
		dm->q_reference = OR_aEM * OR_mElectron;
		dm->m_dark_photon = OR_GeV;
		if(d->form_factor == DMS_FF_GENERAL)
		{
			if(!(d->mediator_mass_MeV > 0.0))
				return DMS_ERR_ARG;
			dm->m_dark_photon = d->mediator_mass_MeV * OR_MeV;
		}
		const double Q = dm->q_reference * dm->q_reference + dm->m_dark_photon * dm->m_dark_photon;
		const double sigma_e_in = d->sigma_e_cm2 * OR_cm * OR_cm;
		/* Set_Sigma_Electron (src/Dark_Photon.cpp:128-131) */
		const double epsilon = Q / 4.0 / dm->mu_e * dm_sqrt(sigma_e_in / DM_PI / OR_aEM / alpha_dark);
		/* Sigma_Proton / Sigma_Electron (src/Dark_Photon.cpp:159-169) */
		dm->sigma_p = 16.0 * DM_PI * OR_aEM * alpha_dark * epsilon * epsilon * dm->mu_p * dm->mu_p / (Q * Q);
		dm->sigma_e = 16.0 * DM_PI * OR_aEM * alpha_dark * epsilon * epsilon * dm->mu_e * dm->mu_e / (Q * Q);
		for(int i = 0; i < S->n_iso; i++)
		{
			const double mu_N = or_reduced_mass(dm->mass, S->m[i]);
			dm->sigma_N[i] = dm->sigma_p * mu_N * mu_N / dm->mu_p / dm->mu_p * S->Z[i] * S->Z[i];
		}
		return DMS_OK;
	}
	return DMS_ERR_UNSUPPORTED;
}

/* Sigma_Total_Nucleus (src/Dark_Photon.cpp:171-193 for DP) */
double or_sigma_nucleus(const or_dm* dm, const or_solar* S, int i, double v)
{
	double sigmatot = dm->sigma_N[i];
	if(dm->interaction == DMS_DARK_PHOTON && dm->ff == DMS_FF_GENERAL)
	{
		const double mu_N = or_reduced_mass(dm->mass, S->m[i]);
		const double q2max = 4.0 * ((mu_N * v) * (mu_N * v));
		const double md = dm->m_dark_photon;
		const double Q = dm->q_reference * dm->q_reference + md * md;
		sigmatot = sigmatot * ((Q * Q) / md / md / (md * md + q2max));
	}
	return sigmatot;
}

/* Sigma_Total_Electron (src/Dark_Photon.cpp:195-213 for DP) */
double or_sigma_electron(const or_dm* dm, double v)
{
	double sigmatot = dm->sigma_e;
	if(dm->interaction == DMS_DARK_PHOTON && dm->ff == DMS_FF_GENERAL)
	{
		const double q2max = 4.0 * ((dm->mu_e * v) * (dm->mu_e * v));
		const double md = dm->m_dark_photon;
		const double Q = dm->q_reference * dm->q_reference + md * md;
		sigmatot = sigmatot * ((Q * Q) / md / md / (md * md + q2max));
	}
	return sigmatot;
}

/* ---- rates: src/Solar_Model.cpp:263-393 ------------------------------------ */
double or_rate_electron(const or_point* P, double r, double v)
{
	if(!dm_isfinite(r) || !dm_isfinite(v))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(v < 0.0)
		v = 0.0;
	if(r > OR_rSun)
		return 0.0;
	const double v_rel = or_thermal_averaged_relative_speed(or_temperature(P->S, r), OR_mElectron, v);
	return or_n_electron(P->S, r) * or_sigma_electron(&P->dm, v) * v_rel;
}

double or_rate_nucleus(const or_point* P, double r, double v, int i)
{
	if(!dm_isfinite(r) || !dm_isfinite(v))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(v < 0.0)
		v = 0.0;
	if(r > OR_rSun)
		return 0.0;
	const double m_target = P->S->m[i];
	const double v_rel = or_thermal_averaged_relative_speed(or_temperature(P->S, r), m_target, v);
	return or_n_nucleus(P->S, r, i) * or_sigma_nucleus(&P->dm, P->S, i, v) * v_rel;
}

double or_rate_computed(const or_point* P, double r, double v)
{
	if(!dm_isfinite(r) || !dm_isfinite(v))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(v < 0.0)
		v = 0.0;
	if(r > OR_rSun)
		return 0.0;
	double total = or_rate_electron(P, r, v);
	for(int i = 0; i < P->S->n_iso; i++)
		total += or_rate_nucleus(P, r, v, i);
	return total;
}

static double or_rate_interpolated(const or_point* P, double r, double v)
{
	if(!dm_isfinite(r) || !dm_isfinite(v))
		return dm_nan();
	if(r > OR_rSun || r < 0.0)
		return 0.0;
	if(v < 0.0)
		return 0.0;
	if(v > P->vmax)
		return or_rate_computed(P, r, v);
	const double rc = r * P->inv_dr;
	const double vc = v * P->inv_dv;
	if(!dm_isfinite(rc) || !dm_isfinite(vc) || rc < 0.0 || vc < 0.0)
		return or_rate_computed(P, r, v);
	unsigned ri = (unsigned) rc;
	if(ri > (unsigned) (P->Nr - 2))
		ri = (unsigned) (P->Nr - 2);
	unsigned vi = (unsigned) vc;
	if(vi > (unsigned) (P->Nv - 2))
		vi = (unsigned) (P->Nv - 2);
	const double rf = rc - ri;
	const double vf = vc - vi;
	const size_t lo = (size_t) ri * (size_t) P->Nv + vi;
	const size_t up = lo + (size_t) P->Nv;
	const double* g = P->grid;
	const double lower = g[lo] + vf * (g[lo + 1] - g[lo]);
	const double upper = g[up] + vf * (g[up + 1] - g[up]);
	const double val = lower + rf * (upper - lower);
	return dm_isfinite(val) ? dm_max(0.0, val) : dm_nan();
}

double or_rate_total(const or_point* P, double r, double v)
{
	if(!dm_isfinite(r) || !dm_isfinite(v))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(v < 0.0)
		v = 0.0;
	if(P->use_grid && v <= P->vmax)
		return or_rate_interpolated(P, r, v);
	return or_rate_computed(P, r, v);
}

double or_point_rate(const or_point* P, int which, double r, double v)
{
	if(which == 0)
		return or_rate_total(P, r, v);
	if(which == 1)
		return or_rate_computed(P, r, v);
	if(which == 2)
		return or_rate_electron(P, r, v);
	if(which >= 3 && which - 3 < P->S->n_iso)
		return or_rate_nucleus(P, r, v, which - 3);
	return dm_nan();
}

/* Interpolate_Total_DM_Scattering_Rate, src/Solar_Model.cpp:395-465 */
typedef struct
{
	or_point* P;
	int row0, row1;
} or_grid_task;

static void* or_grid_worker(void* arg)
{
	or_grid_task* t = (or_grid_task*) arg;
	or_point* P = t->P;
	const double dv = (P->vmax - 0.0) / (double) (P->Nv - 1);
	for(int i = t->row0; i < t->row1; i++)
	{
		const double radius = OR_rSun * (double) i / (double) (P->Nr - 1);
		for(int j = 0; j < P->Nv; j++)
		{
			const double speed = 0.0 + j * dv;
			P->grid[(size_t) i * (size_t) P->Nv + j] = or_rate_computed(P, radius, speed);
		}
	}
	return NULL;
}

or_point* or_point_create(const or_solar* S, const dms_point_desc* desc, int n_threads)
{
	or_point* P = (or_point*) calloc(1, sizeof(or_point));
	P->S = S;
	if(or_dm_init(&P->dm, S, desc) != DMS_OK)
	{
		free(P);
		return NULL;
	}
	/* Interpolate_Total_DM_Scattering_Rate(DM, N_radius, N_speed),
	 * src/Solar_Model.cpp:395-465: either count below 2 disables the grid */
	const int Nr = desc->interpolation_points;
	const int Nv = (desc->speed_points == 0) ? Nr : desc->speed_points;
	if(Nr < 2 || Nv < 2)
	{
		P->use_grid = 0;
		return P;
	}
	P->Nr = Nr;
	P->Nv = Nv;
	P->vmax = 0.75;
	P->grid = (double*) malloc(sizeof(double) * (size_t) Nr * (size_t) Nv);
	if(desc->host_rate_grid != NULL)
		memcpy(P->grid, desc->host_rate_grid, sizeof(double) * (size_t) Nr * (size_t) Nv);
	else
	{
		if(n_threads < 1)
			n_threads = 1;
		if(n_threads > 64)
			n_threads = 64;
		pthread_t th[64];
		or_grid_task tasks[64];
		for(int k = 0; k < n_threads; k++)
		{
			tasks[k].P = P;
			tasks[k].row0 = (int) ((long long) Nr * k / n_threads);
			tasks[k].row1 = (int) ((long long) Nr * (k + 1) / n_threads);
			pthread_create(&th[k], NULL, or_grid_worker, &tasks[k]);
		}
		for(int k = 0; k < n_threads; k++)
			pthread_join(th[k], NULL);
	}
	P->inv_dr = (double) (Nr - 1) / OR_rSun;
	P->inv_dv = (double) (Nv - 1) / P->vmax;
	P->use_grid = 1;
	return P;
}

void or_point_free(or_point* P)
{
	if(P == NULL)
		return;
	free(P->grid);
	free(P);
}

int or_point_grid(const or_point* P, double* out, size_t n)
{
	if(!P->use_grid || n < (size_t) P->Nr * (size_t) P->Nv)
		return DMS_ERR_ARG;
	memcpy(out, P->grid, sizeof(double) * (size_t) P->Nr * (size_t) P->Nv);
	return DMS_OK;
}

void or_point_cross_sections(const or_point* P, double* sigma_p_cm2, double* sigma_e_cm2)
{
	*sigma_p_cm2 = P->dm.sigma_p / (OR_cm * OR_cm);
	*sigma_e_cm2 = P->dm.sigma_e / (OR_cm * OR_cm);
}

/* Host evaluation of the deterministic elementary functions (for comparing the
 * device build of the same functions). fn as dms_probe_math. */
void or_probe_math(int fn, const double* x, const double* y, double* out, size_t n)
{
	for(size_t i = 0; i < n; i++)
	{
		const double a = x[i];
		double r;
		switch(fn)
		{
			case 0: r = dm_sqrt(a); break;
			case 1: r = a / y[i]; break;
			case 2: r = dm_exp(a); break;
			case 3: r = dm_log(a); break;
			case 4: r = dm_sin(a); break;
			case 5: r = dm_cos(a); break;
			case 6: r = dm_erf(a); break;
			default: r = dm_acos(a); break;
		}
		out[i] = r;
	}
}

/* Test hooks: total cross sections of the point's DM model in cm^2 at DM speed
 * v (natural units, c = 1): nucleus i (0-based isotope index) or, for i < 0,
 * the electron. */
double or_point_sigma_target(const or_point* P, int i, double v)
{
	const double s = (i < 0) ? or_sigma_electron(&P->dm, v) : or_sigma_nucleus(&P->dm, P->S, i, v);
	return s / (OR_cm * OR_cm);
}
