/*
 * Unit system of the trajectory path: natural units (GeV, hbar = c = k_B = 1)
 * as in libphysica::natural_units, with CODATA 2018 / IAU 2015 values.
 * Valid C and HIP C++ (constant expressions, evaluated with IEEE rounding).
 */
#ifndef DMS_UNITS_H
#define DMS_UNITS_H

#define DMS_cm (1.0 / 1.973269804e-14)
#define DMS_meter (100.0 * DMS_cm)
#define DMS_km (1000.0 * DMS_meter)
#define DMS_sec (299792458.0 * DMS_meter)
#define DMS_gram (1.0 / 1.78266192e-24)
#define DMS_kg (1000.0 * DMS_gram)
#define DMS_Kelvin 8.617333262e-14
#define DMS_eV 1.0e-9
#define DMS_MeV 1.0e-3
#define DMS_GeV 1.0
#define DMS_mProton 0.93827208816
#define DMS_mElectron 0.51099895e-3
#define DMS_mNucleon 0.93149410242
#define DMS_aEM (1.0 / 137.035999084)
#define DMS_G_Newton (6.67430e-11 * DMS_meter * DMS_meter * DMS_meter / DMS_kg / DMS_sec / DMS_sec)
#define DMS_mSun (1.98847e30 * DMS_kg)
#define DMS_rSun (6.957e8 * DMS_meter)
#define DMS_AU (149597870700.0 * DMS_meter)
#define DMS_dyne (DMS_gram * DMS_cm / DMS_sec / DMS_sec)
#define DMS_R_SUN_KM 6.957e5

/* Engine constants, src/Simulation_Trajectory.cpp:29-43 and
 * include/Simulation_Trajectory.hpp:27-32,121 of the reference. */
#define DMS_RK45_MIN_STEP_FACTOR 0.1
#define DMS_RK45_MAX_STEP_FACTOR 4.0
#define DMS_RK45_MAX_INNER_RETRIES 256
#define DMS_FREE_ENERGY_DRIFT_REL_E_SCALE_EV 1.0e-30
#define DMS_MAX_OPTICAL_DEPTH_STEP 0.05
#define DMS_CAPTURE_MODE_MAX_OPTICAL_DEPTH_STEP 0.10
#define DMS_OPTICAL_DEPTH_RELATIVE_TOLERANCE 1.0e-2
#define DMS_OPTICAL_DEPTH_ABSOLUTE_TOLERANCE (1.0e-12 * DMS_MAX_OPTICAL_DEPTH_STEP)
#define DMS_MAX_OPTICAL_DEPTH_RETRIES 100u
#define DMS_MAX_OPTICAL_DEPTH_PIECES 4
#define DMS_CAPTURE_MODE_OPTICAL_DEPTH_PIECES 2
#define DMS_TARGET_VELOCITY_MAX_REJECTION_ATTEMPTS 10000UL
#define DMS_V_MAX 0.75
#define DMS_MAX_TARGETS 64 /* nuclear isotopes + 1 electron slot */

#endif
