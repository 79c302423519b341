/*
 * CPU ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, scalar FP64 restatement of the DaMaSCUS-SUN-EVAP trajectory path
 * (reference files cited per function in dms_oracle.c).  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it, and only
 * to check the GPU results or as the timed CPU baseline; the product library
 * never links it.
 *
 * It shares two things with the product: the public struct layouts of
 * include/dms.h (the boundary format) and the deterministic elementary
 * functions + Philox stream (detmath.h, philox.h), which play the role libm and
 * std::mt19937 play in the reference.  Everything else is restated here.
 */
#ifndef DMS_ORACLE_H
#define DMS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/dms.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_solar or_solar;
typedef struct or_point or_point;
typedef struct or_prop or_prop;

double or_constant(const char* name);

or_solar* or_solar_create(const double* raw_rows, size_t n_rows);
void or_solar_free(or_solar* S);
double or_solar_eval(const or_solar* S, int which, double r);
int or_solar_n_isotopes(const or_solar* S);
void or_solar_isotope(const or_solar* S, int i, int32_t* Z, int32_t* A, double* mass, double* abundance);
double or_thermal_averaged_relative_speed(double temperature, double mass_target, double v);

/* n_threads used for the rate-grid build. */
or_point* or_point_create(const or_solar* S, const dms_point_desc* desc, int n_threads);
void or_point_free(or_point* P);
/* which: 0 total (grid when enabled), 1 computed, 2 electron, 3+i nucleus i */
double or_point_rate(const or_point* P, int which, double r, double v);
int or_point_grid(const or_point* P, double* out, size_t n);
void or_point_cross_sections(const or_point* P, double* sigma_p_cm2, double* sigma_e_cm2);

/* SHM helpers (obscura Standard_Halo_Model restated). */
double or_shm_pdf_speed(const dms_halo* H, double v_natural);
double or_shm_average_speed(const dms_halo* H);
double or_shm_inverse_average_speed(const dms_halo* H);
double or_dm_entering_rate(const or_solar* S, const dms_halo* H, double mDM_GeV);

int or_initial_conditions(const or_solar* S, const dms_halo* H, uint32_t seed, uint32_t point_id,
                          uint64_t first_id, size_t n, double r_final_rsun, dms_event* out, int32_t* shift_ok);
int or_hyperbolic_kepler_shift(dms_event* e, double R_final);
int or_bound_kepler_return(const dms_event* out_ev, dms_event* in_ev);

/* Free_Particle_Propagator */
or_prop* or_prop_create(const dms_event* e);
void or_prop_free(or_prop* p);
int or_prop_step_solar(or_prop* p, const or_solar* S);
int or_prop_step_mass(or_prop* p, double constant_mass);
void or_prop_state(const or_prop* p, double* time, double* radius, double* speed, double* time_step, dms_event* ev3d);
void or_prop_set_time_step(or_prop* p, double dt);

/* Trajectory_Simulator::Simulate over given initial events. */
int or_simulate_events(const or_point* P, const dms_job_desc* job, const dms_event* initial, size_t n,
                       dms_traj_record* records, double* hist_rows, uint32_t* count_rows, int n_threads);

/* Trajectory_Simulator::Scatter on one event with an explicit Philox stream. */
int or_scatter(const or_point* P, dms_event* e, uint32_t seed, uint32_t point_id, uint64_t traj_id);

/* Sampler hooks (closed-form checks): Philox block; n draws of the CM angle
 * sampler, of Sample_Target, of Sample_Target_Velocity; one New_DM_Velocity. */
void or_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_sample_angles(const or_point* P, uint32_t seed, int target, double v, size_t n, double* out);
void or_sample_targets(const or_point* P, uint32_t seed, double r, double v, size_t n, int32_t* out);
size_t or_sample_target_velocities(uint32_t seed, double temperature, double target_mass, const double* vel_DM, size_t n,
                                   double* out);
int or_new_dm_velocity(uint32_t seed, double cos_a, double mDM, double mT, double* vel, const double* vT);

/* Test hooks of the known-answer checks: Solve_Linear_Rate_Optical_Depth_Fraction,
 * Solar_Interior_Fraction_Interval (returns 0 when the step has no interior
 * part) and the total cross section of target i (< 0: electron) in cm^2. */
double or_solve_linear_rate(double rs, double re, double dt, double target_tau);
int or_interior_interval(const dms_event* b, const dms_event* a, double r_before, double r_after, double* start, double* end);
double or_point_sigma_target(const or_point* P, int i, double v);

/* Simulation_Data::Generate_Data. */
void or_probe_math(int fn, const double* x, const double* y, double* out, size_t n);
int or_run_job(const or_point* P, const dms_job_desc* job, int n_threads, dms_counters* counters,
               double* hist8, uint64_t* counts2, dms_evap_event* events, size_t event_capacity,
               size_t* n_events, dms_traj_record* records, size_t record_capacity, size_t* n_records);

#ifdef __cplusplus
}
#endif

#endif
