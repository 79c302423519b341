/*
 * C++ host API of the MI355X solar dark-matter engine.
 *
 * The classes keep the shape of the reference's public API -- Solar_Model
 * (include/Solar_Model.hpp:26-78), Trajectory_Simulator and Trajectory_Result
 * (include/Simulation_Trajectory.hpp:53-180), Simulation_Data
 * (include/Data_Generation.hpp:65-155) and Configuration
 * (include/Parameter_Scan.hpp:19-39) -- and run every computation through the
 * C ABI of include/dms.h (HIP kernels on one GPU).  Quantities are in the
 * reference's natural units (GeV, hbar = c = k_B = 1); dmsun::units holds the
 * conversion factors.  Errors surface as exceptions, as in the reference:
 * std::invalid_argument for bad arguments, std::runtime_error otherwise.
 */
#ifndef DMSUN_HOST_HPP
#define DMSUN_HOST_HPP

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dms.h"

namespace dmsun
{

namespace units
{
double cm();
double km();
double sec();
double gram();
double Kelvin();
double GeV();
double MeV();
double eV();
double rSun();
double mSun();
double AU();
double G_Newton();
} // namespace units

/* src/Simulation_Utilities.hpp Event (time, position, velocity). */
struct Event
{
	double time = 0.0;
	std::array<double, 3> position{0.0, 0.0, 0.0};
	std::array<double, 3> velocity{0.0, 0.0, 0.0};

	Event() = default;
	Event(double t, const std::array<double, 3>& r, const std::array<double, 3>& v) : time(t), position(r), velocity(v) {}
	double Radius() const;
	double Speed() const;
	double Angular_Momentum() const;
};

/* TrajectoryTerminationReason, include/Simulation_Trajectory.hpp:35-48 */
enum class TrajectoryTerminationReason
{
	Unknown = DMS_UNKNOWN,
	OutwardEscape = DMS_OUTWARD_ESCAPE,
	Scatter = DMS_SCATTER,
	WallTimeLimit = DMS_WALL_TIME_LIMIT,
	MaxFreeSteps = DMS_MAX_FREE_STEPS,
	MaxScatterings = DMS_MAX_SCATTERINGS,
	NonFiniteState = DMS_NON_FINITE_STATE,
	SpeedLimit = DMS_SPEED_LIMIT,
	NumericalFailure = DMS_NUMERICAL_FAILURE,
	CaptureMode = DMS_CAPTURE_MODE,
	EnergyDriftEscape = DMS_ENERGY_DRIFT_ESCAPE
};
const char* Termination_Reason_Name(TrajectoryTerminationReason r);

/* SimulationStopReason, include/Data_Generation.hpp:56-63 */
enum class SimulationStopReason
{
	None = 0,
	MaxTrajectoriesReached = 1,
	CaptureTargetNotReached = 2,
	InitialShiftFailureFractionExceeded = 3,
	InvalidTrajectoryFractionExceeded = 4
};
const char* Stop_Reason_Key(SimulationStopReason reason);
const char* Stop_Reason_Display(SimulationStopReason reason);

/* TrajectoryBincount scalars and per-bin sums (include/Simulation_Trajectory.hpp:53-86). */
struct TrajectoryBincount
{
	std::vector<double> dt_hist, v2dt_hist; /* filled by Simulate when rows are requested */
	bool is_captured = false;
	double t_capture = -1.0, t_last_bound = -1.0, t_final_unbinding_scatter = 0.0, t_boundary_escape = 0.0;
	double t_termination = -1.0, r_first_negative_km = -1.0, E_first_negative_eV = 0.0, dE_first_negative_from_prev_eV = 0.0;
	bool event_observed = false, boundary_escape_observed = false, survival_valid = true, numerically_invalid_escape = false;
	double max_free_energy_drift_eV = 0.0, max_free_energy_drift_rel = 0.0;
	unsigned long int number_of_scatterings = 0;
	bool truncated = false;
	TrajectoryTerminationReason termination_reason = TrajectoryTerminationReason::Unknown;
};

/* Trajectory_Result, include/Simulation_Trajectory.hpp:95-114 */
struct Trajectory_Result
{
	Event initial_event, final_event;
	unsigned long int number_of_scatterings = 0;
	TrajectoryBincount bincount;
	uint32_t flags = 0; /* DMS_F_* */
	uint64_t trajectory_id = 0;

	bool Particle_Reflected() const { return (flags & DMS_F_REFLECTED) != 0; }
	bool Particle_Free() const { return (flags & DMS_F_FREE) != 0; }
	bool Particle_Captured() const { return (flags & DMS_F_CAPTURED) != 0; }
	void Print_Summary() const;
};

/* obscura::DM_Particle_SI (low mass) and DM_Particle_Dark_Photon, reduced to
 * the parameters the path uses. */
class DM_Particle
{
  public:
	double mass = 1.0;          /* GeV (natural units) */
	double sigma_proton = 0.0;  /* natural units (GeV^-2) */
	double sigma_electron = 0.0;
	int interaction = DMS_SI;
	int form_factor = DMS_FF_CONTACT;
	double mediator_mass = 0.0; /* natural units */
	double fn_over_fp = 1.0;
	bool low_mass = true;

	void Set_Mass(double m) { mass = m; }
	void Set_Low_Mass_Mode(bool on) { low_mass = on; }
	void Set_Sigma_Proton(double s) { sigma_proton = s; }
	void Set_Sigma_Electron(double s) { sigma_electron = s; }
	void Set_FormFactor_DM(const std::string& ff, double mediator);
	double Sigma_Proton() const { return sigma_proton; }
	double Sigma_Electron() const { return sigma_electron; }
	/* the point with an N_radius x N_speed rate grid (N_speed 0: N_radius) */
	dms_point_desc Descriptor(int interpolation_points, int speed_points = 0) const;
};
DM_Particle DM_Particle_SI(double mass);
DM_Particle DM_Particle_Dark_Photon(double mass, double sigma_electron);

/* obscura::Standard_Halo_Model (rho in GeV/cm^3, speeds in km/s). */
struct Standard_Halo_Model
{
	double rho_GeV_cm3 = 0.4, v0_kms = 220.0, vesc_kms = 544.0;
	std::array<double, 3> vobs_kms{11.1, 232.24, 7.25};
	void Set_Observer_Velocity(const std::array<double, 3>& v_kms) { vobs_kms = v_kms; }
	dms_halo Descriptor() const;
};

/* Solar_Model (include/Solar_Model.hpp:26-78) on one GPU. */
class Solar_Model
{
  public:
	explicit Solar_Model(int device = 0, const std::string& model_file = "");
	~Solar_Model();
	Solar_Model(const Solar_Model&) = delete;
	Solar_Model& operator=(const Solar_Model&) = delete;

	std::string name = "Standard Solar Model AGSS09";
	double Mass(double r) const;
	double Temperature(double r) const;
	double Local_Escape_Speed(double r) const;
	double Mass_Density(double r) const;
	double Number_Density_Nucleus(double r, unsigned int isotope) const;
	double Number_Density_Electron(double r) const;
	struct Isotope
	{
		int Z, A;
		double mass, abundance;
	};
	const std::vector<Isotope>& Target_Isotopes() const { return isotopes; }

	/* src/Solar_Model.cpp:395-465: builds the N_r x N_v table for DM on the GPU;
	 * N_r < 2 or N_v < 2 turns the table off (rates are then computed). */
	void Interpolate_Total_DM_Scattering_Rate(const DM_Particle& DM, unsigned int N_r, unsigned int N_v);
	unsigned int Scattering_Rate_Interpolation_Radius_Points() const { return radius_points; }
	unsigned int Scattering_Rate_Interpolation_Speed_Points() const { return speed_points; }
	double Total_DM_Scattering_Rate(const DM_Particle& DM, double r, double v);
	double Total_DM_Scattering_Rate_Computed(const DM_Particle& DM, double r, double v);
	void Print_Summary() const;

	dms_ctx* Context() const { return ctx; }
	/* the parameter point for DM: the interpolated one when it matches, else a
	 * direct-rate point (no table), as Total_DM_Scattering_Rate would use */
	dms_point* Point_For(const DM_Particle& DM);

  private:
	dms_ctx* ctx = nullptr;
	std::vector<Isotope> isotopes;
	dms_point* point = nullptr;
	dms_point_desc point_desc{};
	unsigned int radius_points = 0, speed_points = 0;
	double eval(int which, double r) const;
	double rate(const DM_Particle& DM, int which, double r, double v);
};

/* Trajectory_Simulator (include/Simulation_Trajectory.hpp:117-180).  Trajectories
 * draw from the Philox stream (fixed seed, rank) at consecutive ids. */
class Trajectory_Simulator
{
  public:
	Trajectory_Simulator(Solar_Model& model, unsigned long int max_time_steps = 1000000000000ul,
	                     unsigned long int max_scatterings = 100000000000000ul, double max_distance = -1.0);
	unsigned long int maximum_time_steps;
	unsigned long int maximum_scatterings;
	double maximum_distance;
	uint64_t rk_attempt_budget = 0; /* max_trajectory_wall_time_sec analogue (0 = none) */
	unsigned int current_mpi_rank = 0;
	unsigned long int current_trajectory_id = 0;

	void Fix_PRNG_Seed(unsigned int fixed_seed) { seed = fixed_seed; }
	void Enable_Capture_Mode(bool enabled) { capture_mode = enabled; }
	Trajectory_Result Simulate(const Event& initial_condition, const DM_Particle& DM, unsigned int mpi_rank = 0,
	                           bool histogram_rows = false);
	/* many trajectories in one launch (ids current_trajectory_id, +1, ...) */
	std::vector<Trajectory_Result> Simulate(const std::vector<Event>& initial_conditions, const DM_Particle& DM,
	                                        unsigned int mpi_rank = 0);

  private:
	Solar_Model& solar_model;
	unsigned int seed = 0;
	bool capture_mode = false;
	dms_job_desc job(unsigned int rank, size_t n) const;
};

/* Wilson score interval at z = 1.96 (src/Data_Generation.cpp:186-202). */
struct BinomialRateEstimate
{
	double rate = 0.0, standard_error = 0.0, ci_lower = 0.0, ci_upper = 0.0;
};
BinomialRateEstimate Estimate_Binomial_Rate(uint64_t trials, uint64_t successes);
/* sqrt(N Var) of per-trajectory bin contents (src/Data_Generation.cpp:240-251). */
double Snapshot_Bin_Error(double sum, double sum_sq, double count);

/* Simulation_Data (include/Data_Generation.hpp:65-155) */
class Simulation_Data
{
  public:
	Simulation_Data(unsigned int sample_size, unsigned int max_trajectories, double u_min = 0.0, unsigned int iso_rings = 1);
	void Configure(double initial_radius, unsigned int min_scattering, unsigned long int max_scattering,
	               unsigned long int max_free_steps = 1000000000000ul);
	void Set_Bins(int n_bins) { number_of_bins = n_bins; }
	void Set_RK_Attempt_Budget(uint64_t budget) { rk_attempt_budget = budget; }
	/* Progress snapshots (heartbeat analogue, include/Snapshot_IO.hpp:27-57,
	 * src/Snapshot_IO.cpp:508-646): every interval seconds while Generate_Data
	 * runs, <dir>/snapshot_<t>s.txt with the job's live device counters; at the
	 * first checkpoint after the job finished, the report with its final
	 * counters and histograms. */
	void Set_Snapshots(const std::string& dir, double interval_seconds)
	{
		snapshot_dir = dir;
		snapshot_interval = interval_seconds;
	}

	void Generate_Data(const DM_Particle& DM, Solar_Model& solar_model, const Standard_Halo_Model& halo_model,
	                   unsigned int fixed_seed = 0, bool capture_mode = false, unsigned int point_id = 0);

	void Write_Output_Files(const std::string& output_dir, const DM_Particle& DM) const;

	double Free_Ratio() const;
	double Capture_Ratio() const;
	double Reflection_Ratio() const;
	unsigned long int Valid_Trajectories() const;
	double Free_Ratio_Valid() const;
	double Capture_Ratio_Valid() const;
	double Reflection_Ratio_Valid() const;
	double Numerical_Failure_Ratio() const;
	void Print_Summary() const;
	void Print_Capture_Mode_Summary() const;

	/* results (the reference's private members, read-only here) */
	const dms_counters& Counters() const { return counters; }
	SimulationStopReason Early_Stop_Reason() const { return (SimulationStopReason) counters.early_stop_reason; }
	double Computing_Time() const { return counters.computing_time_sec; }
	int Bins() const { return number_of_bins; }
	/* hist[q][b]: q = cap_dt, cap_v2dt, cap_dt_sq, cap_v2dt_sq, nc_dt, nc_v2dt, nc_dt_sq, nc_v2dt_sq */
	const std::vector<double>& Histograms() const { return hist; }
	const std::vector<uint64_t>& Visit_Counts() const { return counts; }
	const std::vector<dms_evap_event>& Evaporation_Events() const { return events; }

	/* Build the result state directly (file writers and tests without a GPU). */
	void Set_Results(const dms_counters& c, const std::vector<double>& hist8, const std::vector<uint64_t>& counts2,
	                 const std::vector<dms_evap_event>& ev);

  private:
	unsigned int requested_captured_particles;
	unsigned long int max_trajectories;
	double initial_and_final_radius;
	unsigned long int maximum_number_of_scatterings = 100000000000000ul;
	unsigned long int maximum_free_time_steps = 1000000000000ul;
	int number_of_bins = 2000;
	uint64_t rk_attempt_budget = 0;
	std::string snapshot_dir;
	double snapshot_interval = 0.0;
	dms_counters counters{};
	std::vector<double> hist;
	std::vector<uint64_t> counts;
	std::vector<dms_evap_event> events;
};

/* Configuration (include/Parameter_Scan.hpp:19-39) read from a libconfig-style
 * file: the keys of src/Parameter_Scan.cpp:90-520 that the path uses. */
class Configuration
{
  public:
	explicit Configuration(const std::string& cfg_filename);
	std::string ID, run_mode, output_dir = "./";
	bool capture_mode = false;
	unsigned int sample_size = 0, max_trajectories = 0, fixed_seed = 0, interpolation_points = 0;
	unsigned long int maximum_number_of_scatterings = 100000000000000ul;
	double max_trajectory_wall_time_sec = 0.0;
	bool snapshot_enabled = false;
	double snapshot_interval_seconds = 60.0;
	DM_Particle DM;
	Standard_Halo_Model halo;
	void Print_Summary() const;

	/* raw settings (numbers as text, lists flattened) */
	std::map<std::string, std::vector<std::string>> settings;
};

/* Serving loop with continuous batching: requests (parameter points with a
 * captured-particle target) are admitted while the device-memory budget and
 * the concurrent-job limit allow; each admitted request runs as its own job
 * on its own streams, so requests that arrive while others are on the GPU
 * overlap with them.  Parameter points stay resident (rate tables are built
 * once per distinct point).  Latency = completion - arrival. */
struct Request
{
	DM_Particle DM;
	Standard_Halo_Model halo;
	uint64_t sample_size = 100;      /* captured-particle target */
	uint64_t max_trajectories = 0;   /* cap (0 = none) */
	uint64_t max_scatterings = 100000000000000ull; /* per trajectory (maximum_number_of_scatterings) */
	uint64_t rk_attempt_budget = 0;  /* per-trajectory wall-time guard analogue (0 = none) */
	unsigned int seed = 20260710, point_id = 0;
	int mode = DMS_MODE_PARAMETER_POINT;
	int precision = DMS_FP64;
	int n_bins = 2000;
	unsigned int interpolation_points = 1000;
};

struct Request_Result
{
	uint64_t id = 0;
	double arrival_sec = 0.0, start_sec = 0.0, finish_sec = 0.0; /* since the server started */
	int status = DMS_OK;
	std::string error;
	dms_counters counters{};
	std::vector<double> hist;
	std::vector<uint64_t> counts;
	std::vector<dms_evap_event> events;
	double Latency() const { return finish_sec - arrival_sec; }
};

class Server
{
  public:
	/* Requests join one resident trajectory launch per launch configuration
	 * (dms_server): max_concurrent requests at once; memory_budget: device
	 * bytes the resident parameter points' rate tables may hold together
	 * (unused points are evicted); per_job_budget: the context's memory budget
	 * for jobs run outside the server (dms_set_job_memory_budget). */
	Server(Solar_Model& model, size_t memory_budget = (size_t) 128 << 30, size_t per_job_budget = (size_t) 8 << 30,
	       int max_concurrent = 64);
	~Server();
	Server(const Server&) = delete;
	Server& operator=(const Server&) = delete;

	uint64_t Submit(const Request& r); /* thread-safe; returns the request id */
	void Wait_All();
	std::vector<Request_Result> Results() const; /* finished requests, by id */
	double Elapsed() const;
	/* counters of the resident launches, summed over launch configurations */
	dms_server_stats Stats() const;

  private:
	struct Impl;
	std::unique_ptr<Impl> impl;
};

/* p-th percentile (0..100, linear interpolation) */
double Percentile(std::vector<double> v, double p);

/* libconfig subset: `name = value;` with strings, numbers, booleans and
 * (possibly nested) lists `( ... )` / arrays `[ ... ]`; comments //, #, / * * /. */
std::map<std::string, std::vector<std::string>> Parse_Config_Text(const std::string& text);

/* Output files of Write_Output_Files (src/Data_Generation.cpp:253-348,1212-1300). */
void Write_Bincount_File(const std::string& path, const Simulation_Data& data, double mass_gev, double sigma_cm2);
void Write_Evaporation_Times_File(const std::string& path, const Simulation_Data& data, double mass_gev, double sigma_cm2);
/* One snapshot report (src/Snapshot_IO.cpp:508-646): the reference's
 * snapshot_<t>s.txt header, its device-status section (one "rank": the GPU
 * job) and, when hist is given (8 x n_bins in bincount.txt column order), the
 * histogram section; written through a temporary file. */
struct Snapshot_Report
{
	long long time_label_s = 0;
	double interval_s = 0.0;
	uint64_t run_id = 0;
	double mass_GeV = 0.0, sigma_cm2 = 0.0;
	uint64_t total_trajectories = 0, captured_particles = 0, valid_trajectories = 0, numerical_failures = 0;
	uint64_t scatterings = 0;
	bool done = false;
	double elapsed_s = 0.0;
	const std::vector<double>* hist = nullptr;
	int n_bins = 0;
	double bin_width_km = 0.0;
};
void Write_Snapshot_Report(const std::string& path, const Snapshot_Report& report);

} // namespace dmsun

#endif
