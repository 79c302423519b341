/*
 * dms.h -- C ABI of the MI355X-native solar dark-matter trajectory engine.
 *
 * This is the drop-in boundary for the per-trajectory Monte Carlo path of
 * DaMaSCUS-SUN-EVAP.  In the reference the path sits behind C++ classes
 * (no FFI): Solar_Model, Trajectory_Simulator and Simulation_Data.  Each entry
 * point below names the reference interface it replaces (file:line, relative
 * to the reference repository root).
 *
 * Conventions
 *   - Plain C types only; no exceptions cross this boundary.
 *   - Return value: DMS_OK (0) on success, a negative DMS_ERR_* code on an API
 *     or device error; dms_last_error() gives the message.
 *   - Physical inputs and outputs use the reference's natural units
 *     (GeV, hbar = c = k_B = 1) unless a field name says otherwise
 *     (_GeV, _cm2, _kms, _sec, _km, _MeV, _rsun).
 *   - Per-trajectory numerical trouble is an outcome (dms_reason), not an
 *     error, exactly as TrajectoryTerminationReason in the reference
 *     (include/Simulation_Trajectory.hpp:35-48).
 *   - Caller owns every array it passes; inputs are copied during the call.
 *     The library owns device memory.  A context is bound to one GPU and may be
 *     used from several host threads: calls on one context are serialised.
 */
#ifndef DMS_H
#define DMS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMS_ABI_VERSION 2

enum dms_status
{
	DMS_OK = 0,
	DMS_ERR_ARG = -1,         /* invalid argument (reference: std::invalid_argument) */
	DMS_ERR_DEVICE = -2,      /* HIP runtime error */
	DMS_ERR_NOMEM = -3,       /* device or host allocation failed */
	DMS_ERR_UNSUPPORTED = -4, /* option outside this build (e.g. SD couplings) */
	DMS_ERR_STATE = -5,       /* call order violated (e.g. read before wait) */
	DMS_ERR_CANCELLED = -6    /* the job was cancelled (dms_job_cancel) */
};

/* TrajectoryTerminationReason, include/Simulation_Trajectory.hpp:35-48 */
enum dms_reason
{
	DMS_UNKNOWN = 0,
	DMS_OUTWARD_ESCAPE = 1,
	DMS_SCATTER = 2,
	DMS_WALL_TIME_LIMIT = 3, /* here: per-trajectory RK-attempt budget exhausted */
	DMS_MAX_FREE_STEPS = 4,
	DMS_MAX_SCATTERINGS = 5,
	DMS_NON_FINITE_STATE = 6,
	DMS_SPEED_LIMIT = 7,
	DMS_NUMERICAL_FAILURE = 8,
	DMS_CAPTURE_MODE = 9,
	DMS_ENERGY_DRIFT_ESCAPE = 10
};

enum dms_interaction
{
	DMS_SI = 0,         /* obscura::DM_Particle_SI, low-mass mode */
	DMS_SD = 1,         /* not supported in this build (needs nuclear spin data) */
	DMS_DARK_PHOTON = 2 /* src/Dark_Photon.cpp */
};

enum dms_form_factor
{
	DMS_FF_CONTACT = 0,
	DMS_FF_GENERAL = 1
};

enum dms_mode
{
	DMS_MODE_CAPTURE = 0,        /* run_mode "Capture" / capture_mode = true */
	DMS_MODE_PARAMETER_POINT = 1 /* run_mode "Parameter point" */
};

enum dms_precision
{
	DMS_FP64 = 0, /* plain FP64, bit-exact against the CPU restatement */
	DMS_MIXED = 1 /* FP32 rates/target sampling/kinematics, FP64 orbit and energy */
};

/* Event: src/Simulation_Utilities.cpp:47-66 (time, position, velocity). */
typedef struct
{
	double t;
	double r[3];
	double v[3];
} dms_event;

/* obscura Standard_Halo_Model parameters (config keys DM_local_density,
 * SHM_v0, SHM_vObserver, SHM_vEscape; src/Parameter_Scan.cpp via obscura). */
typedef struct
{
	double rho_GeV_cm3;
	double v0_kms;
	double vobs_kms[3];
	double vesc_kms;
} dms_halo;

/* DM particle and rate-table request (src/Parameter_Scan.cpp:386-520,
 * src/Dark_Photon.cpp:123-213, src/Solar_Model.cpp:395-465). */
typedef struct
{
	double mass_GeV;
	double sigma_p_cm2;        /* DM_cross_section_nucleon (SI); ignored for DP */
	double sigma_e_cm2;        /* DM_cross_section_electron */
	int32_t interaction;       /* dms_interaction */
	int32_t form_factor;       /* dms_form_factor */
	double mediator_mass_MeV;  /* DP "General" only */
	double fn_over_fp;         /* SI neutron/proton coupling ratio (1 = isospin conserving) */
	int32_t low_mass;          /* DM_light; only 1 is supported */
	/* Rate grid of Interpolate_Total_DM_Scattering_Rate(DM, N_radius, N_speed)
	 * (src/Solar_Model.cpp:395-465): interpolation_points radii in [0, R_sun]
	 * by speed_points speeds in [0, 0.75 c], radius-major.  speed_points = 0
	 * takes N_speed = N_radius (the square grid main.cpp builds from the config
	 * key interpolation_points, src/main.cpp:80).  Either count below 2
	 * disables the grid (:397-406): every rate is then computed directly. */
	int32_t interpolation_points; /* N_radius */
	int32_t speed_points;         /* N_speed (0: N_radius) */
	const double* host_rate_grid; /* optional N_radius*N_speed grid to upload instead of building it */
} dms_point_desc;

/* One batch job: Simulation_Data::Generate_Data (include/Data_Generation.hpp:134). */
typedef struct
{
	int32_t mode;                 /* dms_mode */
	int32_t precision;            /* dms_precision */
	uint64_t max_trajectories;    /* 0 = unlimited (config key max_trajectories) */
	uint64_t capture_target;      /* sample_size; 0 = run exactly max_trajectories */
	uint32_t seed;                /* fixed_seed (Philox key word 0) */
	uint32_t point_id;            /* Philox key word 1 */
	uint64_t first_trajectory_id; /* id of the first trajectory (counter words 0-1) */
	uint64_t max_scatterings;     /* maximum_number_of_scatterings */
	uint64_t max_free_steps;      /* maximum_time_steps of Trajectory_Simulator */
	uint64_t rk_attempt_budget;   /* wall-time analogue per trajectory; 0 = none */
	double r_max_rsun;            /* initial_and_final_radius / rSun (2.0 in main.cpp:72) */
	int32_t n_bins;               /* radial bins over [0, 2 R_sun); 2000 in the reference */
	int32_t record_trajectories;  /* keep one dms_traj_record per trajectory */
	dms_halo halo;
} dms_job_desc;

/* Per-trajectory outcome (Trajectory_Result + TrajectoryBincount scalars,
 * include/Simulation_Trajectory.hpp:53-114). */
typedef struct
{
	dms_event initial;   /* after the hyperbolic Kepler shift to R_max */
	dms_event final;
	uint64_t trajectory_id;
	uint64_t n_scatterings;
	int32_t reason;      /* dms_reason */
	uint32_t flags;      /* DMS_F_* */
	double t_capture_sec;
	double t_last_bound_sec;
	double t_final_unbinding_scatter_sec; /* NaN when unset */
	double t_boundary_escape_sec;         /* NaN when unset */
	double t_termination_sec;
	double r_first_negative_km;
	double E_first_negative_eV;
	double dE_first_negative_from_prev_eV;
	double max_free_energy_drift_eV;
	double max_free_energy_drift_rel;
	/* work counters (algorithmic-op accounting) */
	uint64_t n_rk_attempts;     /* RKF45 stage evaluations (inner attempts) */
	uint64_t n_steps_interior;  /* accepted steps with an optical-depth evaluation */
	uint64_t n_steps_exterior;  /* accepted steps outside the Sun */
	uint64_t n_tau_rejections;  /* optical-depth rejections */
	uint64_t n_tv_iterations;   /* Romano-Walsh rejection iterations */
} dms_traj_record;

#define DMS_F_CAPTURED 0x1u
#define DMS_F_EVENT_OBSERVED 0x2u
#define DMS_F_BOUNDARY_ESCAPE_OBSERVED 0x4u
#define DMS_F_SURVIVAL_VALID 0x8u
#define DMS_F_NUMERICALLY_INVALID_ESCAPE 0x10u
#define DMS_F_TRUNCATED 0x20u
#define DMS_F_SHIFT_FAILED 0x40u     /* initial Hyperbolic_Kepler_Shift failed */
#define DMS_F_FREE 0x80u             /* Trajectory_Result::Particle_Free */
#define DMS_F_REFLECTED 0x100u       /* Trajectory_Result::Particle_Reflected */
#define DMS_F_REFLECTION_SHIFT_FAILED 0x200u /* Hyperbolic_Kepler_Shift to 1 AU failed */

/* Run counters (src/Data_Generation.cpp:437-448, 974-1024). */
typedef struct
{
	uint64_t trajectories;
	uint64_t captured;
	uint64_t completed_outward_escapes;
	uint64_t free_particles;
	uint64_t reflected_particles;
	uint64_t complete_evaporation;
	uint64_t invalid_survival_captured;
	uint64_t initial_shift_failures;
	uint64_t final_reflection_shift_failures;
	uint64_t numerical_failures;
	uint64_t computational_truncations;
	uint64_t total_scatterings;
	uint64_t sync_rounds;
	uint64_t final_round_trajectories;
	uint64_t capture_target_overshoot;
	int32_t early_stop_reason;    /* SimulationStopReason, include/Data_Generation.hpp:56-63 */
	int32_t sync_interval;        /* normal_mode_mpi_sync_interval analogue */
	double computing_time_sec;    /* device time of the generate loop */
	uint64_t reason_counts[11];   /* trajectories per dms_reason */
	uint64_t n_rk_attempts, n_steps_interior, n_steps_exterior, n_tau_rejections, n_tv_iterations;
	uint64_t simulated_trajectories; /* including speculative ones past the stop point */
	uint64_t max_rk_attempts;        /* RKF45 attempts of the job's longest trajectory */
} dms_counters;

/* CompactEvaporationEvent, include/Data_Generation.hpp:48-54 */
typedef struct
{
	int32_t rank;
	int32_t pad;
	uint64_t trajectory_id;
	double completion_wall_time_sec;
	double lifetime_unbinding_sec;
} dms_evap_event;

/* Live progress of a job (the heartbeat's view, include/Snapshot_IO.hpp:27-57):
 * while its launch runs, counts of the trajectories the device has finished;
 * once done, the job's counters. */
typedef struct
{
	uint64_t trajectories_done;
	uint64_t captured;
	uint64_t scatterings;
	int32_t done;
	int32_t pad;
	uint64_t completed_outward_escapes; /* uncaptured trajectories that left (capture-classified with the captured) */
	uint64_t numerical_failures;
} dms_job_status;

typedef struct dms_ctx dms_ctx;
typedef struct dms_point dms_point;
typedef struct dms_job dms_job;

/* ---- library / context ------------------------------------------------- */

int dms_abi_version(void);

/* Solar_Model::Solar_Model + Import_Raw_Data (src/Solar_Model.cpp:43-166).
 * raw_rows: n_rows x 35 data rows of model_agss09.dat in file units
 * (M/M_sun, r/R_sun, T[K], rho[g/cm^3], P, L, 29 mass fractions).
 * device < 0 creates a host-only context (tables and CPU queries only). */
int dms_create(dms_ctx** ctx, int device, const double* raw_rows, size_t n_rows);
void dms_destroy(dms_ctx* ctx);
const char* dms_last_error(const dms_ctx* ctx);

/* Device memory one job may use for its trajectory buffers (default 32 GiB):
 * a capture-target job sizes its chunks by it (budget / per-trajectory bytes),
 * and a fixed-count job or batch whose launch would need more (per-lane
 * histogram scratch, per-slot events and summaries) runs as chunked launches
 * within it instead, with the same results.  Up to 2 jobs run on the device
 * at once, each on its own stream.  No reference counterpart. */
int dms_set_job_memory_budget(dms_ctx* ctx, size_t bytes);

/* Parse the data rows of a model_agss09.dat-format file (20 header lines). */
int dms_load_solar_file(const char* path, double* rows_out, size_t capacity_rows, size_t* n_rows);

/* Solar_Model queries, host-side tables (src/Solar_Model.cpp:168-261).
 * which: 0 Mass, 1 Temperature, 2 Local_Escape_Speed, 3 Mass_Density,
 *        4 Number_Density_Electron, 100+i Number_Density_Nucleus(i). */
int dms_solar_eval(const dms_ctx* ctx, int which, const double* r, double* out, size_t n);
/* Target isotope list (src/Solar_Model.cpp:143-163). */
int dms_isotopes(const dms_ctx* ctx, int32_t* Z, int32_t* A, double* mass, double* abundance, size_t capacity, size_t* n);
/* Named constant of the unit system (e.g. "rSun", "km", "sec"). */
double dms_constant(const char* name);

/* ---- parameter points ---------------------------------------------------- */

/* DM particle + Solar_Model::Interpolate_Total_DM_Scattering_Rate
 * (src/Solar_Model.cpp:395-465): builds the radius-major N_radius x N_speed
 * grid on the GPU (or uploads host_rate_grid). */
int dms_point_create(dms_ctx* ctx, const dms_point_desc* desc, dms_point** point);
void dms_point_destroy(dms_point* point);
/* Copy the rate grid (N_radius*N_speed doubles, radius-major) to host. */
int dms_point_rate_grid(const dms_point* point, double* out, size_t n);
/* Grid shape in use (Scattering_Rate_Interpolation_Radius_Points /
 * _Speed_Points, include/Solar_Model.hpp:74-75): 0 x 0 without a grid. */
int dms_point_grid_shape(const dms_point* point, int32_t* n_radius, int32_t* n_speed);
/* Effective proton/electron cross sections in cm^2 (DP derives sigma_p). */
int dms_point_cross_sections(const dms_point* point, double* sigma_p_cm2, double* sigma_e_cm2);
/* Solar_Model::Total_DM_Scattering_Rate (+_Computed, per-target rates) evaluated
 * on the device for n (r, v) pairs.  which: 0 total (grid when enabled),
 * 1 computed, 2 electron, 3+i nucleus i. */
int dms_point_rates(const dms_point* point, int which, const double* r, const double* v, double* out, size_t n);

/* ---- trajectory batches --------------------------------------------------- */

/* Initial_Conditions + Hyperbolic_Kepler_Shift (src/Simulation_Utilities.cpp:176-321,
 * src/Data_Generation.cpp:747-748) on the device for ids first_id .. first_id+n-1. */
int dms_initial_conditions(dms_ctx* ctx, const dms_halo* halo, uint32_t seed, uint32_t point_id,
                           uint64_t first_id, size_t n, double r_final_rsun,
                           dms_event* out, int32_t* shift_ok);

/* Trajectory_Simulator::Simulate (src/Simulation_Trajectory.cpp:1276-1400) for a
 * batch of given initial events; trajectory i uses Philox counter id first_id+i.
 * hist_rows (optional): n x 2 x n_bins doubles (dt then v2dt per trajectory). */
int dms_simulate_events(dms_point* point, const dms_job_desc* job, const dms_event* initial, size_t n,
                        dms_traj_record* records, double* hist_rows);

/* dms_simulate_events for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  d_initial (n events),
 * d_records (n records) and d_hist_rows (optional, n x 2 x n_bins doubles,
 * cleared by the call) are DEVICE pointers on the point's GPU; the kernels'
 * other memory -- the slot counter and the one-job table -- lives in a
 * 256-byte array of the library that every call on that GPU shares, so calls
 * that overlap on one GPU race and must be ordered by the caller (one stream,
 * or synchronisation).  The call allocates nothing: with stream != NULL (a
 * hipStream_t) it only enqueues its kernels there; with NULL it runs on the
 * context's stream and returns when they have finished.  The events must pass
 * Simulate's argument check (finite, t >= 0, r > 0, v > 0); unlike
 * dms_simulate_events the host does not read them.  At most 2^31 - 1
 * trajectories per call. */
int dms_simulate_events_noalloc(dms_point* point, const dms_job_desc* job, const dms_event* d_initial, size_t n,
                                dms_traj_record* d_records, double* d_hist_rows, void* stream);

/* ---- jobs (Simulation_Data::Generate_Data, src/Data_Generation.cpp:488-972) -- */

int dms_job_submit(dms_point* point, const dms_job_desc* desc, dms_job** job);
/* Several fixed-count jobs of one point (one Generate_Data call per job, as
 * separate ranks or separate runs would make them) simulated in ONE trajectory
 * launch, so that the long trajectories of one job overlap the bulk of the
 * others.  Every job needs max_trajectories > 0, capture_target 0, and the
 * same descriptor as jobs[0] apart from max_trajectories, first_trajectory_id
 * and record_trajectories; the batch holds at most 2^27 trajectories and
 * 65536 bins.  Each job's results equal those of dms_job_submit on the same
 * descriptor (counters, counts and events exactly; the histogram sums to the
 * rounding of their summation order).  jobs[k] receives the handle of descs[k];
 * every handle is waited on, read and destroyed separately. */
int dms_job_submit_batch(dms_point* point, const dms_job_desc* descs, size_t n, dms_job** jobs);
/* Fixed-count jobs of several parameter points (and seeds, point ids, halos)
 * in ONE trajectory launch: the batched-uneven workload of many (mass, cross
 * section) points, whose trajectory lengths differ by orders of magnitude,
 * sharing one persistent launch with lane refill.  points[k] is the point of
 * descs[k] (all of one context).  The jobs must agree in mode, precision,
 * max_scatterings, max_free_steps, rk_attempt_budget, r_max_rsun and n_bins;
 * limits as dms_job_submit_batch.  Each job's results equal those of
 * dms_job_submit on its own point and descriptor.  No reference counterpart
 * (one Generate_Data call per point). */
int dms_job_submit_multi(dms_point* const* points, const dms_job_desc* descs, size_t n, dms_job** jobs);
/* Autotuner of a multi-point launch: runs calibration_trajectories
 * trajectories of every job in one launch (ids past each job's own range, at
 * most calibration_rk_attempts RKF45 attempts each), measures the lane
 * iterations each job's trajectories need, and returns in order[0..n) the job
 * order to pass to dms_job_submit_multi (cheapest first) and, optionally, the
 * measured cost per trajectory of each job.  The order changes the launch time,
 * never a job's results.  No reference counterpart. */
int dms_multi_autotune(dms_point* const* points, const dms_job_desc* descs, size_t n, uint64_t calibration_trajectories,
                       uint64_t calibration_rk_attempts, uint32_t* order, double* cost_per_trajectory);
/* Launch knobs of a context: results never depend on them, only speed.
 * refill_batch: free lanes a wave waits for before it takes new slots (1 = at
 * once, up to 64), so that trajectory starts run together instead of one lane
 * at a time; a wave with no lane left running, or a launch whose slots are all
 * handed out, refills at once.
 * lone_serial: 0 (default) lets a trajectory left alone in its wave (every
 * other lane finished, no slot left) share its steps with a finished wave of
 * its block, which prepares the next step while the trajectory's wave
 * post-processes the current one; 1 runs it on its own wave only.  No
 * reference counterpart. */
typedef struct
{
	uint32_t refill_batch;
	uint32_t lone_serial;
} dms_launch_knobs;

int dms_set_launch_knobs(dms_ctx* ctx, const dms_launch_knobs* knobs);
int dms_get_launch_knobs(const dms_ctx* ctx, dms_launch_knobs* knobs);
/* Autotuner of the launch knobs: the calibration launch of dms_multi_autotune
 * (same jobs, same caps) run once per candidate refill_batch, `repeats` times
 * each in turn; seconds[i] receives candidate i's fastest kernel time.  The
 * first candidate is the reference: another is chosen (*best) only when it is
 * more than 10% faster.  With apply != 0 the context takes *best.  Submit no
 * other jobs on the context meanwhile (the knobs are context-wide).  No
 * reference counterpart. */
int dms_autotune_knobs(dms_point* const* points, const dms_job_desc* descs, size_t n, uint64_t calibration_trajectories,
                       uint64_t calibration_rk_attempts, const uint32_t* refill_candidates, size_t n_candidates,
                       int repeats, double* seconds, dms_launch_knobs* best, int apply);
int dms_job_poll(dms_job* job, dms_job_status* status);
/* Stop a running job: every lane of the launch it runs in ends within about
 * 1024 loop iterations (a few ms for a lone trajectory), so that a runaway
 * trajectory (no RKF45 budget, unbounded scatterings) cannot hold the device;
 * jobs sharing that launch (a batch, a multi-point launch) end with it.  Each
 * of them then returns DMS_ERR_CANCELLED from dms_job_wait and dms_job_read.
 * A request of a dms_server stops at its current point instead: its slots not
 * yet started are handed back and the trajectories already running finish.
 * Cancelling a finished job does nothing.  No reference counterpart (the
 * reference's runs end by their wall-time guard). */
int dms_job_cancel(dms_job* job);
int dms_job_wait(dms_job* job, dms_job_status* status);
/* hist: 8 x n_bins doubles in the order of bincount.txt columns
 *   cap_dt, cap_v2dt, cap_dt_sq, cap_v2dt_sq, nc_dt, nc_v2dt, nc_dt_sq, nc_v2dt_sq
 * (src/Data_Generation.cpp:789-795,816-822); counts: 2 x n_bins visit counts
 * (captured, not captured).  records (optional) need capacity >= trajectories. */
int dms_job_read(dms_job* job, dms_counters* counters, double* hist, uint64_t* counts, size_t n_bins,
                 dms_evap_event* events, size_t event_capacity, size_t* n_events,
                 dms_traj_record* records, size_t record_capacity, size_t* n_records);
/* Device timing of the last job: total kernel time and per-kernel average. */
int dms_job_timing(const dms_job* job, double* trajectory_kernel_ms, double* launches, double* other_ms);
void dms_job_destroy(dms_job* job);

/* ---- serving: requests join one resident launch ---------------------------- */

/* A server keeps one persistent trajectory launch on the device and appends
 * the trajectories of each submitted request to its work queue as slots while
 * it runs (continuous batching at slot level): a request starts on the lanes
 * that free up, without waiting for the others and without a launch of its
 * own.  Every request's results equal those of dms_job_submit on the same
 * point and descriptor (counters, counts and events exactly; the histogram
 * sums to the rounding of their summation order): its trajectories are
 * replayed in id order with the reference's round logic, as the ranks of
 * Simulation_Data::Generate_Data take trajectories inside their sync rounds
 * (src/Data_Generation.cpp:697-728).  With no request left the launch ends
 * (after about 2 ms) and the next request starts it again.  While it runs the
 * launch occupies the device's compute units, so jobs submitted outside the
 * server meanwhile run on what is left.  No reference counterpart (one
 * Generate_Data call per parameter point). */
typedef struct dms_server dms_server;

typedef struct
{
	/* mode, precision, max_scatterings, max_free_steps, rk_attempt_budget,
	 * r_max_rsun and n_bins of every request (the other fields are ignored) */
	dms_job_desc launch;
	uint32_t slot_ring;    /* slots in flight at most: power of two in [1024, 2^26] (0: 2^20) */
	uint32_t max_requests; /* requests admitted at once, others queue (0: 1024; at most 65536) */
	uint32_t max_rounds;   /* per-round histogram buffers (0: about 256 MiB of them; at most 65536) */
	uint32_t pad;
} dms_server_desc;

typedef struct
{
	uint64_t launches;        /* resident launches started: 1 + restarts after idle periods */
	uint64_t slots_published; /* trajectory slots appended to the launch */
	uint64_t slots_cancelled; /* slots handed back unstarted after their request had stopped */
	uint64_t requests_done;
	uint32_t requests_active; /* admitted and not finished */
	uint32_t running;         /* 1 while the launch is on the device */
} dms_server_stats;

int dms_server_create(dms_ctx* ctx, const dms_server_desc* desc, dms_server** server);
/* Queue a request; the returned job is waited on, polled, read and destroyed
 * like any other (dms_job_wait / dms_job_poll / dms_job_read /
 * dms_job_destroy) and stays valid after the server is destroyed.  The
 * descriptor must agree with the server's launch descriptor in mode,
 * precision, limits, r_max and n_bins; record_trajectories is not supported
 * (DMS_ERR_UNSUPPORTED). */
int dms_server_submit(dms_server* server, dms_point* point, const dms_job_desc* desc, dms_job** job);
int dms_server_get_stats(const dms_server* server, dms_server_stats* stats);
/* Waits for every submitted request to finish, ends the launch and frees the
 * server. */
void dms_server_destroy(dms_server* server);

/* ---- diagnostics ----------------------------------------------------------- */

/* Evaluate the deterministic elementary functions on the device for n inputs.
 * fn: 0 sqrt, 1 div (x/y), 2 exp, 3 log, 4 sin, 5 cos, 6 erf, 7 acos,
 * 8 x/y by the shared-reciprocal division of the RKF45 stages, 9 x/y by the
 * constant-divisor division (reciprocal 1/y, product, two FMAs). */
int dms_probe_math(dms_ctx* ctx, int fn, const double* x, const double* y, double* out, size_t n);

#ifdef __cplusplus
}
#endif

#endif
