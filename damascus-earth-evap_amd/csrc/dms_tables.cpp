/*
 * Solar model tables and DM particle parameters (host side of libdmsun).
 *
 * Restates Solar_Model construction (src/Solar_Model.cpp:43-166), the target
 * isotope list (:143-163) and the cross sections of obscura's low-mass SI model
 * and of src/Dark_Photon.cpp:123-213, producing the interleaved cubic tables
 * the kernels read.  Interpolation is Steffen's monotone cubic (M. Steffen,
 * A&A 239 (1990) 443; libphysica::Interpolation).
 */
#include <cmath>
#include <cstring>

#include "dms_host.hpp"

namespace dms
{

namespace
{

double sign_of(double v) { return (double) ((v > 0.0) - (v < 0.0)); }

/* Steffen coefficients for nodes x, values y (eqs. 11, 26-27 of Steffen 1990). */
void steffen(const std::vector<double>& x, const std::vector<double>& y, DmsCubic* out, size_t stride)
{
	const int n = (int) x.size();
	std::vector<double> h(n - 1), s(n - 1), yp(n);
	for(int i = 0; i < n - 1; i++)
	{
		h[i] = x[i + 1] - x[i];
		s[i] = (y[i + 1] - y[i]) / h[i];
	}
	for(int i = 1; i < n - 1; i++)
	{
		const double p = (s[i - 1] * h[i] + s[i] * h[i - 1]) / (h[i - 1] + h[i]);
		yp[i] = (sign_of(s[i - 1]) + sign_of(s[i])) * dm_min(dm_fabs(s[i - 1]), dm_min(dm_fabs(s[i]), 0.5 * dm_fabs(p)));
	}
	const double p0 = s[0] * (1.0 + h[0] / (h[0] + h[1])) - s[1] * h[0] / (h[0] + h[1]);
	yp[0] = (p0 * s[0] <= 0.0) ? 0.0 : ((dm_fabs(p0) > 2.0 * dm_fabs(s[0])) ? 2.0 * s[0] : p0);
	const int m = n - 2;
	const double pn = s[m] * (1.0 + h[m] / (h[m] + h[m - 1])) - s[m - 1] * h[m] / (h[m] + h[m - 1]);
	yp[n - 1] = (pn * s[m] <= 0.0) ? 0.0 : ((dm_fabs(pn) > 2.0 * dm_fabs(s[m])) ? 2.0 * s[m] : pn);
	for(int i = 0; i < n - 1; i++)
	{
		DmsCubic& c = out[(size_t) i * stride];
		c.a = (yp[i] + yp[i + 1] - 2.0 * s[i]) / h[i] / h[i];
		c.b = (3.0 * s[i] - 2.0 * yp[i] - yp[i + 1]) / h[i];
		c.c = yp[i];
		c.d = y[i];
	}
}

struct NaturalIsotope
{
	int Z, A;
	double abundance;
};

/* Natural isotopic abundances (IUPAC representative values) for Z = 10..28. */
const NaturalIsotope kNatural[] = {
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

} // namespace

DmsSolarView SolarTables::view() const
{
	DmsSolarView v;
	v.n = n;
	v.n_iso = n_iso;
	v.x1 = x1;
	v.inv_h = inv_h;
	v.x = x.data();
	v.mass = mass.data();
	v.temp = temp.data();
	v.vesc2 = vesc2.data();
	v.dens = dens.data();
	v.iso_pref = iso_pref.data();
	v.iso_mass = iso_mass.data();
	v.iso_Z = iso_Z.data();
	return v;
}

double SolarTables::mass_density(double r) const
{
	if(!dm_isfinite(r))
		return dm_nan();
	if(r < 0.0)
		r = 0.0;
	if(r > DMS_rSun)
		return 0.0;
	const DmsSolarView v = view();
	const int j = dms_interval(v, r);
	return 1.0 * dms_cubic(rho[j], r - x[j]);
}

std::string build_solar_tables(const double* raw_rows, size_t n_rows, SolarTables& T)
{
	if(raw_rows == nullptr || n_rows < 4)
		return "solar table needs at least 4 data rows";
	const int n = (int) n_rows + 2;
	T.n = n;
	T.raw.assign((size_t) n * kColumns, 0.0);
	double units[kColumns];
	for(int c = 0; c < kColumns; c++)
		units[c] = 1.0;
	units[0] = DMS_mSun;
	units[1] = DMS_rSun;
	units[2] = DMS_Kelvin;
	units[3] = DMS_gram / DMS_cm / DMS_cm / DMS_cm;
	units[4] = DMS_dyne / DMS_cm / DMS_cm;
	for(size_t i = 0; i < n_rows; i++)
		for(int c = 0; c < kColumns; c++)
		{
			const double v = raw_rows[i * kColumns + c];
			if(!std::isfinite(v))
				return "solar table contains a non-finite entry";
			T.raw[(i + 1) * kColumns + c] = v * units[c];
		}
	/* padding rows, src/Solar_Model.cpp:47-65 */
	for(int c = 0; c < kColumns; c++)
	{
		T.raw[c] = (c == 0 || c == 1) ? 0.0 : T.raw[1 * kColumns + c];
		double last;
		if(c == 0)
			last = DMS_mSun;
		else if(c == 1)
			last = DMS_rSun;
		else if(c == 2)
			last = 5800 * DMS_Kelvin;
		else if(c == 3)
			last = 1.0e-9 * DMS_gram / DMS_cm / DMS_cm / DMS_cm;
		else
			last = T.raw[(size_t) (n - 2) * kColumns + c];
		T.raw[(size_t) (n - 1) * kColumns + c] = last;
	}
	T.x.resize(n);
	for(int i = 0; i < n; i++)
		T.x[i] = T.raw[(size_t) i * kColumns + 1];
	for(int i = 0; i < n - 1; i++)
		if(!(T.x[i + 1] > T.x[i]))
			return "solar radius column is not strictly increasing";
	T.x1 = T.x[1];
	T.inv_h = 1.0 / ((T.x[n - 2] - T.x[1]) / (double) (n - 3));
	std::vector<double> y(n);
	auto column = [&](int c) {
		for(int i = 0; i < n; i++)
			y[i] = T.raw[(size_t) i * kColumns + c];
	};
	T.mass.resize(n - 1);
	T.temp.resize(n - 1);
	T.vesc2.resize(n - 1);
	T.rho.resize(n - 1);
	T.dens.assign((size_t) (n - 1) * DMS_MAX_TARGETS, DmsCubic{0.0, 0.0, 0.0, 0.0});
	column(0);
	steffen(T.x, y, T.mass.data(), 1);
	column(2);
	steffen(T.x, y, T.temp.data(), 1);
	/* isotope bookkeeping needed by view() before the escape-speed quadrature */
	T.iso_pref.assign(DMS_MAX_TARGETS, 0.0);
	T.iso_mass.assign(DMS_MAX_TARGETS, 0.0);
	T.iso_Z.assign(DMS_MAX_TARGETS, 0);
	T.iso_A.assign(DMS_MAX_TARGETS, 0);
	/* Create_Escape_Speed_Table (src/Solar_Model.cpp:79-96): integral of
	 * M(x)/x^2 from r to R_sun by 10-point Gauss-Legendre per profile interval,
	 * accumulated inward from the surface. */
	{
		const DmsSolarView V = T.view();
		std::vector<double> Q(n, 0.0);
		for(int j = 0; j < n - 1; j++)
		{
			const double mid = 0.5 * (T.x[j] + T.x[j + 1]);
			const double half = 0.5 * (T.x[j + 1] - T.x[j]);
			double sum = 0.0;
			for(int k = 0; k < 10; k++)
			{
				const double xq = mid + half * DM_GL10_X[k];
				const double f = (xq == 0.0) ? 0.0 : dms_mass(V, xq) / xq / xq;
				sum = sum + DM_GL10_W[k] * f;
			}
			Q[j] = half * sum;
		}
		double above = 0.0;
		for(int i = n - 1; i >= 0; i--)
		{
			const double rr = T.x[i];
			double integral;
			if(rr < DMS_rSun)
			{
				above = Q[i] + above;
				integral = above;
			}
			else
				integral = DMS_mSun * (1.0 / rr - 1.0 / DMS_rSun);
			y[i] = 2.0 * DMS_G_Newton * DMS_mSun / DMS_rSun * (1.0 + DMS_rSun / DMS_mSun * integral);
		}
	}
	steffen(T.x, y, T.vesc2.data(), 1);
	column(3);
	steffen(T.x, y, T.rho.data(), 1);
	/* target isotopes and number densities (src/Solar_Model.cpp:98-110,143-163) */
	static const int Zs[29] = {1, 2, 2, 6, 6, 7, 7, 8, 8, 8, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28};
	static const int As[10] = {1, 4, 3, 12, 13, 14, 15, 16, 17, 18};
	int k = 0;
	for(int t = 0; t < 29; t++)
	{
		const int nat = (int) (sizeof(kNatural) / sizeof(kNatural[0]));
		for(int q = 0; q < (t < 10 ? 1 : nat); q++)
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
				if(kNatural[q].Z != Zs[t])
					continue;
				Z = Zs[t];
				A = kNatural[q].A;
				ab = kNatural[q].abundance;
			}
			if(k >= kMaxIso)
				return "too many target isotopes";
			const double m = (A == 1) ? DMS_mProton : A * DMS_mNucleon;
			for(int i = 0; i < n; i++)
			{
				const double rho = T.raw[(size_t) i * kColumns + 3];
				const double f = T.raw[(size_t) i * kColumns + 6 + t];
				y[i] = f * rho / m;
			}
			T.iso_Z[k] = Z;
			T.iso_A[k] = A;
			T.iso_mass[k] = m;
			T.iso_pref[k] = ab;
			steffen(T.x, y, T.dens.data() + k, DMS_MAX_TARGETS);
			k++;
		}
	}
	T.n_iso = k;
	/* Create_Number_Density_Table_Electron (src/Solar_Model.cpp:112-126) */
	{
		const DmsSolarView V = T.view();
		for(int i = 0; i < n; i++)
			y[i] = 0.0;
		for(int q = 0; q < T.n_iso; q++)
			for(int i = 0; i < n; i++)
			{
				const double r = T.x[i];
				double nq = 0.0;
				if(!(r > DMS_rSun))
				{
					const int j = dms_interval(V, r);
					nq = dms_density_at(V, j, r - T.x[j], q);
				}
				y[i] += T.iso_Z[q] * nq;
			}
		steffen(T.x, y, T.dens.data() + (DMS_MAX_TARGETS - 1), DMS_MAX_TARGETS);
	}
	return "";
}

std::string build_point_params(const SolarTables& S, const dms_point_desc& d, PointParams& P)
{
	P = PointParams();
	if(!(d.mass_GeV > 0.0) || !std::isfinite(d.mass_GeV))
		return "DM mass must be finite and positive";
	if(d.low_mass != 1)
		return "only the low-mass mode (DM_light = true) is supported";
	if(!(d.sigma_p_cm2 >= 0.0) || !(d.sigma_e_cm2 >= 0.0) || !std::isfinite(d.sigma_p_cm2) || !std::isfinite(d.sigma_e_cm2))
		return "DM cross sections must be finite and non-negative";
	P.interaction = d.interaction;
	P.ff = d.form_factor;
	P.mass = d.mass_GeV * DMS_GeV;
	P.mu_p = dms_reduced_mass(P.mass, DMS_mProton);
	P.mu_e = dms_reduced_mass(P.mass, DMS_mElectron);
	if(d.interaction == DMS_SI)
	{
		if(d.form_factor != DMS_FF_CONTACT)
			return "SI interactions support the Contact form factor only";
		P.sigma_p = d.sigma_p_cm2 * DMS_cm * DMS_cm;
		P.sigma_e = d.sigma_e_cm2 * DMS_cm * DMS_cm;
		for(int i = 0; i < S.n_iso; i++)
		{
			const double mu_N = dms_reduced_mass(P.mass, S.iso_mass[i]);
			const double coupling = S.iso_Z[i] + d.fn_over_fp * (S.iso_A[i] - S.iso_Z[i]);
			P.sigma_N[i] = P.sigma_p * mu_N * mu_N / P.mu_p / P.mu_p * coupling * coupling;
		}
		return "";
	}
	if(d.interaction == DMS_DARK_PHOTON)
	{
		if(d.form_factor != DMS_FF_CONTACT && d.form_factor != DMS_FF_GENERAL)
			return "dark photon form factor must be Contact or General (others diverge in the IR)";
		const double alpha_dark = DMS_aEM;
		P.q_reference = DMS_aEM * DMS_mElectron;
		P.m_dark_photon = DMS_GeV;
		if(d.form_factor == DMS_FF_GENERAL)
		{
			if(!(d.mediator_mass_MeV > 0.0))
				return "DM_mediator_mass must be positive for the General form factor";
			P.m_dark_photon = d.mediator_mass_MeV * DMS_MeV;
		}
		const double Q = P.q_reference * P.q_reference + P.m_dark_photon * P.m_dark_photon;
		const double sigma_e_in = d.sigma_e_cm2 * DMS_cm * DMS_cm;
		const double epsilon = Q / 4.0 / P.mu_e * dm_sqrt(sigma_e_in / DM_PI / DMS_aEM / alpha_dark);
		P.sigma_p = 16.0 * DM_PI * DMS_aEM * alpha_dark * epsilon * epsilon * P.mu_p * P.mu_p / (Q * Q);
		P.sigma_e = 16.0 * DM_PI * DMS_aEM * alpha_dark * epsilon * epsilon * P.mu_e * P.mu_e / (Q * Q);
		for(int i = 0; i < S.n_iso; i++)
		{
			const double mu_N = dms_reduced_mass(P.mass, S.iso_mass[i]);
			P.sigma_N[i] = P.sigma_p * mu_N * mu_N / P.mu_p / P.mu_p * S.iso_Z[i] * S.iso_Z[i];
		}
		return "";
	}
	if(d.interaction == DMS_SD)
		return "SD interactions need obscura's nuclear spin table, which this build does not carry";
	return "unknown DM interaction";
}

uint64_t normal_mode_sync_interval(double sigma_cm2)
{
	if(!std::isfinite(sigma_cm2) || sigma_cm2 <= 0.0)
		return 128;
	if(sigma_cm2 >= 1.0e-35)
		return 64;
	if(sigma_cm2 >= 1.0e-36)
		return 128;
	if(sigma_cm2 >= 1.0e-37)
		return 1024;
	if(sigma_cm2 >= 1.0e-38)
		return 8192;
	if(sigma_cm2 >= 1.0e-39)
		return 65536;
	return 1048576;
}

} // namespace dms
