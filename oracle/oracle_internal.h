/* CPU ORACLE -- TEST INFRASTRUCTURE ONLY (see dms_oracle.h). Internal declarations. */
#ifndef DMS_ORACLE_INTERNAL_H
#define DMS_ORACLE_INTERNAL_H

#include <stdint.h>
#include <stddef.h>

#include "detmath.h"
#include "philox.h"
#include "dms_oracle.h"

/* Natural units (GeV, hbar = c = k_B = 1), CODATA 2018 / IAU 2015 values.
 * These replace libphysica::natural_units (not available offline); they
 * reproduce the reference's known answers v_esc(0) = 1384.13 +- 0.1 km/s
 * (tests/test_Solar_Model.cpp:78), <v_rel>(1e7 K, m_p, 1e-3) = 0.0017928
 * (tests/test_Solar_Model.cpp:275) and the SHM entering rate
 * 1.06689e30 /s +- 0.1% (tests/test_Reflection_Spectrum.cpp:72). */
#define OR_cm (1.0 / 1.973269804e-14)
#define OR_meter (100.0 * OR_cm)
#define OR_km (1000.0 * OR_meter)
#define OR_sec (299792458.0 * OR_meter)
#define OR_gram (1.0 / 1.78266192e-24)
#define OR_kg (1000.0 * OR_gram)
#define OR_Kelvin 8.617333262e-14
#define OR_eV 1.0e-9
#define OR_MeV 1.0e-3
#define OR_GeV 1.0
#define OR_mProton 0.93827208816
#define OR_mElectron 0.51099895e-3
#define OR_mNucleon 0.93149410242
#define OR_aEM (1.0 / 137.035999084)
#define OR_G_Newton (6.67430e-11 * OR_meter * OR_meter * OR_meter / OR_kg / OR_sec / OR_sec)
#define OR_mSun (1.98847e30 * OR_kg)
#define OR_rSun (6.957e8 * OR_meter)
#define OR_AU (149597870700.0 * OR_meter)
#define OR_dyne (OR_gram * OR_cm / OR_sec / OR_sec)
#define OR_R_SUN_KM 6.957e5

#define OR_MAX_ISO 64
#define OR_NCOL 35

typedef struct
{
	int n;           /* nodes */
	const double* x; /* nodes (owned by the solar model) */
	double* a;       /* n-1 coefficients each */
	double* b;
	double* c;
	double* d;
	double pref;
} or_spline;

struct or_solar
{
	int n; /* padded node count */
	double* raw; /* n x 35, natural units */
	double* r;   /* nodes */
	or_spline mass, temperature, vesc2, rho, n_e;
	int n_iso;
	int Z[OR_MAX_ISO];
	int A[OR_MAX_ISO];
	double m[OR_MAX_ISO];
	double abundance[OR_MAX_ISO];
	or_spline n_iso_spline[OR_MAX_ISO];
};

typedef struct
{
	int interaction, ff;
	double mass;
	double sigma_p, sigma_e; /* effective proton / electron cross sections */
	double m_dark_photon, q_reference;
	double fnfp;
	double mu_p, mu_e;
	double sigma_N[OR_MAX_ISO]; /* velocity-independent part of Sigma_Total_Nucleus */
} or_dm;

struct or_point
{
	const or_solar* S;
	or_dm dm;
	int use_grid;
	int Nr, Nv;
	double* grid;
	double inv_dr, inv_dv, vmax;
};

/* splines / solar */
double or_spline_eval(const or_spline* s, double x);
double or_mass(const or_solar* S, double r);
double or_temperature(const or_solar* S, double r);
double or_local_escape_speed(const or_solar* S, double r);
double or_n_nucleus(const or_solar* S, double r, int i);
double or_n_electron(const or_solar* S, double r);

/* rates */
double or_rate_electron(const or_point* P, double r, double v);
double or_rate_nucleus(const or_point* P, double r, double v, int i);
double or_rate_total(const or_point* P, double r, double v);
double or_rate_computed(const or_point* P, double r, double v);
double or_sigma_nucleus(const or_dm* dm, const or_solar* S, int i, double v);
double or_sigma_electron(const or_dm* dm, double v);

/* vectors */
static inline double or_dot(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static inline double or_norm(const double* a) { return dm_sqrt(or_dot(a, a)); }
static inline void or_cross(const double* a, const double* b, double* o)
{
	const double x = a[1] * b[2] - a[2] * b[1];
	const double y = a[2] * b[0] - a[0] * b[2];
	const double z = a[0] * b[1] - a[1] * b[0];
	o[0] = x;
	o[1] = y;
	o[2] = z;
}
static inline void or_normalize(const double* a, double* o)
{
	const double n = or_norm(a);
	o[0] = a[0] / n;
	o[1] = a[1] / n;
	o[2] = a[2] / n;
}

/* Kepler */
int or_kepler_shift_impl(dms_event* e, double R_final);

/* engine */
typedef struct
{
	const or_point* P;
	uint64_t max_time_steps, max_scatterings;
	double max_distance;
	int terminate_on_capture;
	uint64_t rk_budget;
	int nbins;
	double bin_width_km;
} or_sim_cfg;

/* Attribution marks of the op-counting build (oracle/opcount/): which part of a
 * trajectory the floating-point operations belong to.  They do nothing in the
 * oracle itself. */
#ifndef OR_OPCOUNT
#define OR_OPS_PUSH(part) ((void) 0)
#define OR_OPS_POP() ((void) 0)
#define OR_OPS_STEP_BEGIN() ((void) 0)
#define OR_OPS_STEP_CLASS(part) ((void) 0)
#define OR_OPS_STEP_END() ((void) 0)
#endif

int or_simulate_one(const or_sim_cfg* cfg, const dms_event* ic, uint32_t seed, uint32_t point_id, uint64_t traj_id,
                    dms_traj_record* rec, double* dt_hist, double* v2dt_hist, uint32_t* counts);
int or_scatter_impl(const or_point* P, dms_event* e, dm_rng* rng, uint64_t* tv_iterations);
void or_make_initial_condition(const or_solar* S, const dms_halo* H, dm_rng* rng, dms_event* out);

#endif
