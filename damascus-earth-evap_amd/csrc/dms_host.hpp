/*
 * Host-side tables of the trajectory path (private to libdmsun).
 */
#ifndef DMS_HOST_HPP
#define DMS_HOST_HPP

#include <string>
#include <vector>

#include "dms_physics.h"
#include "../../include/dms.h"

namespace dms
{

constexpr int kColumns = 35;
constexpr int kMaxIso = DMS_MAX_TARGETS - 1;

/* Solar_Model tables (src/Solar_Model.cpp:43-166) in the device layout. */
struct SolarTables
{
	int n = 0;
	int n_iso = 0;
	std::vector<double> raw; /* n x 35, natural units, padded */
	std::vector<double> x;
	std::vector<DmsCubic> mass, temp, vesc2, rho, dens; /* dens: (n-1) x 64 */
	std::vector<double> iso_pref, iso_mass;
	std::vector<int> iso_Z, iso_A;
	double x1 = 0.0, inv_h = 0.0;

	DmsSolarView view() const;
	double mass_density(double r) const;
};

/* DM particle parameters of one point (SI low mass, dark photon). */
struct PointParams
{
	int interaction = 0, ff = 0;
	double mass = 0.0, sigma_p = 0.0, sigma_e = 0.0;
	double mu_p = 0.0, mu_e = 0.0;
	double m_dark_photon = 0.0, q_reference = 0.0;
	double sigma_N[DMS_MAX_TARGETS] = {0};
};

/* Builds the tables; returns an empty string or an error message. */
std::string build_solar_tables(const double* raw_rows, size_t n_rows, SolarTables& out);
std::string build_point_params(const SolarTables& S, const dms_point_desc& d, PointParams& out);

/* Sync interval of the reference's round loop by cross section
 * (src/Data_Generation.cpp:59-74). */
uint64_t normal_mode_sync_interval(double sigma_cm2);

} // namespace dms

#endif
