/*
 * Host-only parts of the C++ API: run statistics, the output files of
 * Simulation_Data::Write_Output_Files, run summaries and the configuration
 * reader.  Nothing here touches the GPU.
 */
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <sys/stat.h>
#include <unistd.h>

#include "dmsun.hpp"

namespace dmsun
{

const char* Termination_Reason_Name(TrajectoryTerminationReason r)
{
	static const char* names[] = {"Unknown",       "OutwardEscape",  "Scatter",    "WallTimeLimit",
	                              "MaxFreeSteps",  "MaxScatterings", "NonFiniteState", "SpeedLimit",
	                              "NumericalFailure", "CaptureMode", "EnergyDriftEscape"};
	const int i = (int) r;
	return (i >= 0 && i < 11) ? names[i] : "Unknown";
}

/* src/Data_Generation.cpp:204-238 */
const char* Stop_Reason_Key(SimulationStopReason reason)
{
	switch(reason)
	{
		case SimulationStopReason::MaxTrajectoriesReached: return "max_trajectories_reached";
		case SimulationStopReason::CaptureTargetNotReached: return "capture_target_not_reached";
		case SimulationStopReason::InitialShiftFailureFractionExceeded: return "initial_shift_failure_fraction_exceeded";
		case SimulationStopReason::InvalidTrajectoryFractionExceeded: return "invalid_trajectory_fraction_exceeded";
		default: return "none";
	}
}

const char* Stop_Reason_Display(SimulationStopReason reason)
{
	switch(reason)
	{
		case SimulationStopReason::MaxTrajectoriesReached: return "max_trajectories reached";
		case SimulationStopReason::CaptureTargetNotReached: return "capture target not reached";
		case SimulationStopReason::InitialShiftFailureFractionExceeded: return "initial shift failure fraction exceeded";
		case SimulationStopReason::InvalidTrajectoryFractionExceeded:
			return "numerical/computational invalid-trajectory fraction exceeded";
		default: return "none";
	}
}

BinomialRateEstimate Estimate_Binomial_Rate(uint64_t trials, uint64_t successes)
{
	BinomialRateEstimate e;
	if(trials == 0)
		return e;
	const double N = (double) trials;
	e.rate = (double) std::min(successes, trials) / N;
	e.standard_error = std::sqrt(e.rate * (1.0 - e.rate) / N);
	const double z = 1.96;
	const double denominator = 1.0 + z * z / N;
	const double center = e.rate + z * z / (2.0 * N);
	const double spread = z * std::sqrt(e.rate * (1.0 - e.rate) / N + z * z / (4.0 * N * N));
	e.ci_lower = std::max(0.0, (center - spread) / denominator);
	e.ci_upper = std::min(1.0, (center + spread) / denominator);
	return e;
}

double Snapshot_Bin_Error(double sum, double sum_sq, double count)
{
	if(count <= 1.0)
		return 0.0;
	const double mean = sum / count;
	double variance = sum_sq / count - mean * mean;
	if(variance < 0.0)
		variance = 0.0;
	return std::sqrt(count * variance);
}

/* ---- Simulation_Data: results and statistics ------------------------------ */

Simulation_Data::Simulation_Data(unsigned int sample_size, unsigned int max_traj, double, unsigned int)
    : requested_captured_particles(sample_size), max_trajectories(max_traj), initial_and_final_radius(-1.0)
{
}

void Simulation_Data::Configure(double initial_radius, unsigned int, unsigned long int max_scattering, unsigned long int max_free_steps)
{
	initial_and_final_radius = initial_radius;
	maximum_number_of_scatterings = max_scattering;
	maximum_free_time_steps = max_free_steps;
}

void Simulation_Data::Set_Results(const dms_counters& c, const std::vector<double>& hist8, const std::vector<uint64_t>& counts2,
                                  const std::vector<dms_evap_event>& ev)
{
	if(hist8.size() % 8 != 0 || counts2.size() * 4 != hist8.size())
		throw std::invalid_argument("Simulation_Data::Set_Results: histogram sizes do not match");
	counters = c;
	hist = hist8;
	counts = counts2;
	events = ev;
	number_of_bins = (int) (hist8.size() / 8);
}

unsigned long int Simulation_Data::Valid_Trajectories() const { return counters.captured + counters.completed_outward_escapes; }
static double ratio(uint64_t a, uint64_t b) { return b > 0 ? 1.0 * (double) a / (double) b : 0.0; }
double Simulation_Data::Free_Ratio() const { return ratio(counters.free_particles, counters.trajectories); }
double Simulation_Data::Capture_Ratio() const { return ratio(counters.captured, counters.trajectories); }
double Simulation_Data::Reflection_Ratio() const { return ratio(counters.reflected_particles, counters.trajectories); }
double Simulation_Data::Free_Ratio_Valid() const { return ratio(counters.free_particles, Valid_Trajectories()); }
double Simulation_Data::Capture_Ratio_Valid() const { return ratio(counters.captured, Valid_Trajectories()); }
double Simulation_Data::Reflection_Ratio_Valid() const { return ratio(counters.reflected_particles, Valid_Trajectories()); }
double Simulation_Data::Numerical_Failure_Ratio() const { return ratio(counters.numerical_failures, counters.trajectories); }

static const char* SEPARATOR = "----------------------------------------\n";

/* src/Data_Generation.cpp:1366-1404 */
void Simulation_Data::Print_Capture_Mode_Summary() const
{
	const dms_counters& c = counters;
	const BinomialRateEstimate raw = Estimate_Binomial_Rate(c.trajectories, c.captured);
	const BinomialRateEstimate valid = Estimate_Binomial_Rate(Valid_Trajectories(), c.captured);
	std::cout << SEPARATOR << "CAPTURE MODE summary\n\n"
	          << "Termination condition:\t\tpost-scatter E < 0\n"
	          << "File output:\t\t\tdisabled\n"
	          << "Simulated trajectories:\t\t" << c.trajectories << "\n"
	          << "Capture-classified trajectories:\t" << Valid_Trajectories() << "\n"
	          << "Unresolved non-captures:\t\t" << (c.trajectories - Valid_Trajectories()) << "\n"
	          << "Captured count:\t\t\t" << c.captured << "\n"
	          << std::fixed << std::setprecision(8) << "Capture rate raw:\t\t" << raw.rate << "\n"
	          << "Capture rate valid:\t\t" << valid.rate << "\n"
	          << "Capture rate raw 95% CI:\t[" << raw.ci_lower << ", " << raw.ci_upper << "]\n"
	          << "Capture rate valid 95% CI:\t[" << valid.ci_lower << ", " << valid.ci_upper << "]\n"
	          << "Numerical failure count:\t" << c.numerical_failures << "\n"
	          << "Computational truncations:\t" << c.computational_truncations << "\n"
	          << "Initial shift failures:\t\t" << c.initial_shift_failures << "\n"
	          << "Final reflection shift failures:\t" << c.final_reflection_shift_failures << "\n"
	          << "Numerical failure rate:\t\t" << Numerical_Failure_Ratio() << "\n";
	if(Early_Stop_Reason() != SimulationStopReason::None)
		std::cout << "*** EARLY STOP: " << Stop_Reason_Display(Early_Stop_Reason()) << " ***\n";
	if(Numerical_Failure_Ratio() > 1.0e-4)
		std::cout << "*** WARNING: numerical failure rate exceeded 0.0001 ***\n";
	const double rate = c.computing_time_sec > 0.0 ? (double) c.captured / c.computing_time_sec : 0.0;
	std::cout << "Captured particle rate [1/s]:\t" << std::setprecision(3) << rate << "\n"
	          << "Simulation time [s]:\t\t" << std::setprecision(6) << c.computing_time_sec << "\n"
	          << SEPARATOR << std::endl;
	std::cout.unsetf(std::ios::floatfield);
}

/* src/Data_Generation.cpp:1406-1460 */
void Simulation_Data::Print_Summary() const
{
	const dms_counters& c = counters;
	const double avg = c.trajectories > 0 ? (double) c.total_scatterings / (double) c.trajectories : 0.0;
	std::cout << SEPARATOR << "Simulation data summary\n\nResults:\n"
	          << "Simulated trajectories:\t\t" << c.trajectories << "\n"
	          << "Capture-classified trajectories:\t" << Valid_Trajectories() << "\n"
	          << "Unresolved non-captures:\t\t" << (c.trajectories - Valid_Trajectories()) << "\n"
	          << std::setprecision(4) << "Average # of scatterings:\t" << avg << "\n"
	          << "Free particles raw [%]:\t\t" << 100.0 * Free_Ratio() << "\n"
	          << "Reflected particles raw [%]:\t" << 100.0 * Reflection_Ratio() << "\n"
	          << "Captured particles raw [%]:\t" << 100.0 * Capture_Ratio() << "\n"
	          << "Free particles valid [%]:\t" << 100.0 * Free_Ratio_Valid() << "\n"
	          << "Reflected particles valid [%]:\t" << 100.0 * Reflection_Ratio_Valid() << "\n"
	          << "Captured particles valid [%]:\t" << 100.0 * Capture_Ratio_Valid() << "\n"
	          << "Captured count:\t\t\t" << c.captured << "\n"
	          << "Normal-mode MPI sync interval:\t" << c.sync_interval << "\n"
	          << "MPI sync rounds:\t\t" << c.sync_rounds << "\n"
	          << "Final MPI round trajectories:\t" << c.final_round_trajectories << "\n"
	          << "MPI tail trajectories:\t\t" << 0 << "\n"
	          << "Capture target overshoot:\t" << c.capture_target_overshoot << "\n"
	          << "Total # of scatterings:\t" << c.total_scatterings << "\n"
	          << "Numerical failure count:\t" << c.numerical_failures << "\n"
	          << "Computational truncations:\t" << c.computational_truncations << "\n"
	          << "Initial shift failures:\t\t" << c.initial_shift_failures << "\n"
	          << "Final reflection shift failures:\t" << c.final_reflection_shift_failures << "\n"
	          << std::fixed << std::setprecision(6) << "Numerical failure rate:\t\t" << Numerical_Failure_Ratio() << "\n"
	          << "Complete evaporation count:\t" << c.complete_evaporation << "\n"
	          << "Censored captured count:\t" << 0 << "\n"
	          << "Invalid survival count:\t\t" << c.invalid_survival_captured << "\n";
	const BinomialRateEstimate raw = Estimate_Binomial_Rate(c.trajectories, c.captured);
	const BinomialRateEstimate valid = Estimate_Binomial_Rate(Valid_Trajectories(), c.captured);
	std::cout << "Capture rate raw error (1σ):\t" << raw.standard_error << "\n"
	          << "Capture rate raw 95% CI:\t[" << raw.ci_lower << ", " << raw.ci_upper << "]\n"
	          << "Capture rate valid error (1σ):\t" << valid.standard_error << "\n"
	          << "Capture rate valid 95% CI:\t[" << valid.ci_lower << ", " << valid.ci_upper << "]\n";
	if(Early_Stop_Reason() != SimulationStopReason::None)
		std::cout << "*** EARLY STOP: " << Stop_Reason_Display(Early_Stop_Reason()) << " ***\n";
	if(Numerical_Failure_Ratio() > 1.0e-4)
		std::cout << "*** WARNING: numerical failure rate exceeded 0.0001 ***\n";
	std::cout << "Simulation time [s]:\t\t" << c.computing_time_sec << "\n" << SEPARATOR << std::endl;
	std::cout.unsetf(std::ios::floatfield);
}

/* ---- output files ------------------------------------------------------- */

/* Write_Report_Header + the header lambda of Write_Output_Files
 * (src/Data_Generation.cpp:253-278,1229-1256) */
static void write_header(std::ofstream& f, const Simulation_Data& d, double mass_gev, double sigma_cm2)
{
	const dms_counters& c = d.Counters();
	f << "# DM_mass_GeV = " << std::scientific << std::setprecision(6) << mass_gev << "\n";
	f << "# DM_sigma_cm2 = " << std::scientific << std::setprecision(6) << sigma_cm2 << "\n";
	f << "# total_trajectories = " << c.trajectories << "\n";
	f << "# captured_particles = " << c.captured << "\n";
	const BinomialRateEstimate raw = Estimate_Binomial_Rate(c.trajectories, c.captured);
	f << std::fixed << std::setprecision(8);
	f << "# capture_rate = " << raw.rate << "\n";
	f << "# capture_rate_raw = " << raw.rate << "\n";
	f << "# capture_rate_err = " << raw.standard_error << "\n";
	f << "# capture_rate_raw_err = " << raw.standard_error << "\n";
	f << "# capture_rate_CI_95_lower = " << raw.ci_lower << "\n";
	f << "# capture_rate_CI_95_upper = " << raw.ci_upper << "\n";
	f << "# capture_rate_raw_CI_95_lower = " << raw.ci_lower << "\n";
	f << "# capture_rate_raw_CI_95_upper = " << raw.ci_upper << "\n";
	if(d.Early_Stop_Reason() != SimulationStopReason::None)
		f << "# EARLY_STOP: " << Stop_Reason_Key(d.Early_Stop_Reason()) << "\n";
	const uint64_t valid = d.Valid_Trajectories();
	f << "# valid_trajectories = " << valid << "\n";
	f << "# completed_outward_escapes = " << c.completed_outward_escapes << "\n";
	f << "# unresolved_not_captured_trajectories = " << (c.trajectories > valid ? c.trajectories - valid : 0) << "\n";
	f << "# numerical_failures = " << c.numerical_failures << "\n";
	f << "# computational_truncations = " << c.computational_truncations << "\n";
	f << "# initial_shift_failures = " << c.initial_shift_failures << "\n";
	f << "# final_reflection_shift_failures = " << c.final_reflection_shift_failures << "\n";
	f << "# normal_mode_mpi_sync_interval = " << c.sync_interval << "\n";
	f << "# mpi_sync_rounds = " << c.sync_rounds << "\n";
	f << "# final_mpi_round_trajectories = " << c.final_round_trajectories << "\n";
	f << "# mpi_tail_trajectories = " << 0 << "\n";
	f << "# capture_target_overshoot = " << c.capture_target_overshoot << "\n";
	f << "# total_scatterings = " << c.total_scatterings << "\n";
	const double avg = c.trajectories > 0 ? (double) c.total_scatterings / (double) c.trajectories : 0.0;
	f.unsetf(std::ios::floatfield);
	f << "# average_scatterings = " << std::setprecision(12) << avg << "\n";
	f << "# simulation_time_seconds = " << std::setprecision(12) << c.computing_time_sec << "\n";
	const BinomialRateEstimate v = Estimate_Binomial_Rate(valid, c.captured);
	f << std::fixed << std::setprecision(8);
	f << "# capture_rate_valid = " << v.rate << "\n";
	f << "# capture_rate_valid_err = " << v.standard_error << "\n";
	f << "# capture_rate_valid_CI_95_lower = " << v.ci_lower << "\n";
	f << "# capture_rate_valid_CI_95_upper = " << v.ci_upper << "\n";
	f << "# numerical_failure_rate = " << d.Numerical_Failure_Ratio() << "\n";
}

void Write_Bincount_File(const std::string& path, const Simulation_Data& d, double mass_gev, double sigma_cm2)
{
	std::ofstream f(path);
	if(!f.is_open())
		throw std::runtime_error("cannot write " + path);
	write_header(f, d, mass_gev, sigma_cm2);
	const dms_counters& c = d.Counters();
	const uint64_t valid = d.Valid_Trajectories();
	const uint64_t excluded = c.trajectories > valid ? c.trajectories - valid : 0;
	f << "# not_captured_bincount_samples = " << c.completed_outward_escapes << "\n";
	if(excluded > 0)
		f << "# excluded_incomplete_not_captured = " << excluded << "\n";
	f << "# bin_index  cap_dt[s]  cap_v2dt[km2/s]  cap_err_dt[s]  cap_err_v2dt[km2/s]  not_cap_dt[s]  not_cap_v2dt[km2/s]  "
	     "not_cap_err_dt[s]  not_cap_err_v2dt[km2/s]\n";
	const int nb = d.Bins();
	const std::vector<double>& h = d.Histograms();
	const double N_cap = (double) c.captured;
	const double N_nc = (double) c.completed_outward_escapes;
	auto H = [&](int q, int b) { return h.empty() ? 0.0 : h[(size_t) q * nb + b]; };
	f << std::scientific << std::setprecision(10);
	for(int b = 0; b < nb; b++)
	{
		f << b << "\t" << H(0, b) << "\t" << H(1, b) << "\t" << Snapshot_Bin_Error(H(0, b), H(2, b), N_cap) << "\t"
		  << Snapshot_Bin_Error(H(1, b), H(3, b), N_cap) << "\t" << H(4, b) << "\t" << H(5, b) << "\t"
		  << Snapshot_Bin_Error(H(4, b), H(6, b), N_nc) << "\t" << Snapshot_Bin_Error(H(5, b), H(7, b), N_nc) << "\n";
	}
	if(!f.good())
		throw std::runtime_error("failed writing " + path);
}

/* Write_Final_Evaporation_Time_File (src/Data_Generation.cpp:281-348): events
 * sorted by (lifetime, rank, trajectory id), written through a temporary file. */
void Write_Evaporation_Times_File(const std::string& path, const Simulation_Data& d, double mass_gev, double sigma_cm2)
{
	const std::string tmp = path + ".final.tmp." + std::to_string(getpid());
	{
		std::ofstream f(tmp, std::ios::out | std::ios::trunc);
		if(!f.is_open())
			throw std::runtime_error("cannot write " + path);
		f << "# DaMaSCUS-SUN evaporation times\n";
		f << "# format_version = 3\n";
		f << "# DM_mass_GeV = " << std::scientific << std::setprecision(6) << mass_gev << "\n";
		f << "# DM_sigma_cm2 = " << std::scientific << std::setprecision(6) << sigma_cm2 << "\n";
		f << "# sorted_by = lifetime_unbinding_sec rank trajectory_id\n";
		f << "# rank trajectory_id lifetime_unbinding_sec\n";
		std::vector<dms_evap_event> ev = d.Evaporation_Events();
		std::sort(ev.begin(), ev.end(), [](const dms_evap_event& a, const dms_evap_event& b) {
			if(a.lifetime_unbinding_sec != b.lifetime_unbinding_sec)
				return a.lifetime_unbinding_sec < b.lifetime_unbinding_sec;
			if(a.rank != b.rank)
				return a.rank < b.rank;
			return a.trajectory_id < b.trajectory_id;
		});
		for(const dms_evap_event& e : ev)
			if(std::isfinite(e.lifetime_unbinding_sec) && e.lifetime_unbinding_sec >= 0.0)
				f << e.rank << "\t" << e.trajectory_id << "\t" << std::scientific << std::setprecision(10)
				  << e.lifetime_unbinding_sec << "\n";
		if(!f.good())
		{
			std::remove(tmp.c_str());
			throw std::runtime_error("failed writing " + path);
		}
	}
	if(std::rename(tmp.c_str(), path.c_str()) != 0)
	{
		std::remove(tmp.c_str());
		throw std::runtime_error("cannot move " + tmp + " to " + path);
	}
}

void Write_Snapshot_Report(const std::string& path, const Snapshot_Report& r)
{
	const std::string tmp = path + ".tmp." + std::to_string(getpid());
	{
		std::ofstream f(tmp, std::ios::out | std::ios::trunc);
		if(!f.is_open())
			return;
		f << "# snapshot_status = merged\n";
		f << "# snapshot_run_id = " << r.run_id << "\n";
		f << "# Cumulative snapshot report\n";
		f << "# ready_ranks = 0\n";
		f << "# missing_ranks = \n";
		f << "# snapshot_target_wall_time_s = " << r.time_label_s << "\n";
		f << "# snapshot_interval_s = " << std::fixed << std::setprecision(3) << r.interval_s << "\n";
		f << "# DM_mass_GeV = " << std::scientific << std::setprecision(6) << r.mass_GeV << "\n";
		f << "# DM_sigma_cm2 = " << std::scientific << std::setprecision(6) << r.sigma_cm2 << "\n";
		f << "# total_trajectories = " << r.total_trajectories << "\n";
		f << "# captured_particles = " << r.captured_particles << "\n";
		f << "# valid_trajectories = " << r.valid_trajectories << "\n";
		f << "# numerical_failures = " << r.numerical_failures << "\n";
		f << "# unresolved_not_captured_trajectories = "
		  << (r.total_trajectories > r.valid_trajectories ? r.total_trajectories - r.valid_trajectories : 0) << "\n";
		const BinomialRateEstimate raw = Estimate_Binomial_Rate(r.total_trajectories, r.captured_particles);
		const BinomialRateEstimate valid = Estimate_Binomial_Rate(r.valid_trajectories, r.captured_particles);
		f << std::fixed << std::setprecision(8);
		f << "# capture_rate = " << raw.rate << "\n";
		f << "# capture_rate_raw = " << raw.rate << "\n";
		f << "# capture_rate_err = " << raw.standard_error << "\n";
		f << "# capture_rate_raw_err = " << raw.standard_error << "\n";
		f << "# capture_rate_CI_95_lower = " << raw.ci_lower << "\n";
		f << "# capture_rate_CI_95_upper = " << raw.ci_upper << "\n";
		f << "# capture_rate_raw_CI_95_lower = " << raw.ci_lower << "\n";
		f << "# capture_rate_raw_CI_95_upper = " << raw.ci_upper << "\n";
		f << "# capture_rate_valid = " << valid.rate << "\n";
		f << "# capture_rate_valid_err = " << valid.standard_error << "\n";
		f << "# capture_rate_valid_CI_95_lower = " << valid.ci_lower << "\n";
		f << "# capture_rate_valid_CI_95_upper = " << valid.ci_upper << "\n";
		/* the device runs the job's trajectories in one launch: one status line */
		f << "#\n# [MPI rank status]\n";
		f << "# rank  state  trajectory_id  trajectory_wall_s  simulated_elapsed_s  scatterings  observed_at_wall_s\n";
		f << "# 0\t" << (r.done ? "done" : "running_uncaptured") << "\t" << r.total_trajectories << "\t" << std::scientific
		  << std::setprecision(10) << r.elapsed_s << "\t" << 0.0 << "\t" << r.scatterings << "\t" << r.elapsed_s << "\n";
		if(r.hist && r.n_bins > 0 && r.hist->size() >= 8 * (size_t) r.n_bins)
		{
			const std::vector<double>& h = *r.hist;
			const size_t nb = (size_t) r.n_bins;
			auto H = [&](int q, size_t b) { return h[(size_t) q * nb + b]; };
			const double N_cap = (double) r.captured_particles;
			const double N_nc = (double) (r.valid_trajectories - std::min(r.valid_trajectories, r.captured_particles));
			f << "#\n# [Bincount histogram]\n";
			f << "# bin_index  cap_dt[s]  cap_v2dt[km2/s]  cap_err_dt[s]  cap_err_v2dt[km2/s]  not_cap_dt[s]  not_cap_v2dt[km2/s]  "
			     "not_cap_err_dt[s]  not_cap_err_v2dt[km2/s]\n";
			for(size_t b = 0; b < nb; b++)
				f << b << "\t" << std::scientific << std::setprecision(10) << H(0, b) << "\t" << H(1, b) << "\t"
				  << Snapshot_Bin_Error(H(0, b), H(2, b), N_cap) << "\t" << Snapshot_Bin_Error(H(1, b), H(3, b), N_cap) << "\t"
				  << H(4, b) << "\t" << H(5, b) << "\t" << Snapshot_Bin_Error(H(4, b), H(6, b), N_nc) << "\t"
				  << Snapshot_Bin_Error(H(5, b), H(7, b), N_nc) << "\n";
		}
	}
	std::rename(tmp.c_str(), path.c_str());
}

static bool ensure_directory(const std::string& dir)
{
	if(dir.empty())
		return true;
	std::string partial;
	std::stringstream ss(dir);
	std::string part;
	if(dir[0] == '/')
		partial = "/";
	while(std::getline(ss, part, '/'))
	{
		if(part.empty())
			continue;
		partial += part + "/";
		struct stat st;
		if(stat(partial.c_str(), &st) != 0 && mkdir(partial.c_str(), 0755) != 0)
			return false;
	}
	struct stat st;
	return stat(dir.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

bool Ensure_Directory(const std::string& dir) { return ensure_directory(dir); }

/* src/Data_Generation.cpp:1212-1300 */
void Simulation_Data::Write_Output_Files(const std::string& output_dir, const DM_Particle& DM) const
{
	if(!ensure_directory(output_dir))
		throw std::runtime_error("Write_Output_Files: cannot create " + output_dir);
	const double mass_gev = DM.mass / units::GeV();
	const double sigma_cm2 = DM.Sigma_Proton() / (units::cm() * units::cm());
	const std::string dir = (output_dir.empty() || output_dir.back() == '/') ? output_dir : output_dir + "/";
	Write_Bincount_File(dir + "bincount.txt", *this, mass_gev, sigma_cm2);
	Write_Evaporation_Times_File(dir + "evaporation_times.txt", *this, mass_gev, sigma_cm2);
}

/* ---- configuration -------------------------------------------------------- */

namespace
{
struct Lexer
{
	const std::string& s;
	size_t i = 0;
	explicit Lexer(const std::string& t) : s(t) {}
	void skip()
	{
		for(;;)
		{
			while(i < s.size() && std::isspace((unsigned char) s[i]))
				i++;
			if(i < s.size() && (s[i] == '#' || (s[i] == '/' && i + 1 < s.size() && s[i + 1] == '/')))
			{
				while(i < s.size() && s[i] != '\n')
					i++;
				continue;
			}
			if(i + 1 < s.size() && s[i] == '/' && s[i + 1] == '*')
			{
				const size_t e = s.find("*/", i + 2);
				i = (e == std::string::npos) ? s.size() : e + 2;
				continue;
			}
			return;
		}
	}
	[[noreturn]] void error(const std::string& what) const
	{
		throw std::invalid_argument("configuration parse error at offset " + std::to_string(i) + ": " + what);
	}
	std::string name()
	{
		skip();
		const size_t b = i;
		while(i < s.size() && (std::isalnum((unsigned char) s[i]) || s[i] == '_' || s[i] == '-' || s[i] == '*'))
			i++;
		if(b == i)
			error("expected a setting name");
		return s.substr(b, i - b);
	}
	void value(std::vector<std::string>& out)
	{
		skip();
		if(i >= s.size())
			error("expected a value");
		const char c = s[i];
		if(c == '(' || c == '[')
		{
			const char close = (c == '(') ? ')' : ']';
			i++;
			skip();
			if(i < s.size() && s[i] == close)
			{
				i++;
				return;
			}
			for(;;)
			{
				value(out);
				skip();
				if(i < s.size() && s[i] == ',')
				{
					i++;
					continue;
				}
				if(i < s.size() && s[i] == close)
				{
					i++;
					return;
				}
				error("expected ',' or a closing bracket");
			}
		}
		if(c == '"')
		{
			std::string v;
			i++;
			while(i < s.size() && s[i] != '"')
			{
				if(s[i] == '\\' && i + 1 < s.size())
					i++;
				v += s[i++];
			}
			if(i >= s.size())
				error("unterminated string");
			i++;
			/* adjacent string literals are concatenated */
			skip();
			if(i < s.size() && s[i] == '"')
			{
				std::vector<std::string> more;
				value(more);
				v += more.empty() ? "" : more[0];
			}
			out.push_back(v);
			return;
		}
		const size_t b = i;
		while(i < s.size() && !std::isspace((unsigned char) s[i]) && s[i] != ';' && s[i] != ',' && s[i] != ')' && s[i] != ']')
			i++;
		if(b == i)
			error("expected a value");
		out.push_back(s.substr(b, i - b));
	}
};

double to_double(const std::map<std::string, std::vector<std::string>>& m, const std::string& key, double fallback, bool required)
{
	auto it = m.find(key);
	if(it == m.end() || it->second.empty())
	{
		if(required)
			throw std::invalid_argument("No '" + key + "' setting in configuration file.");
		return fallback;
	}
	const std::string& t = it->second[0];
	if(t == "true")
		return 1.0;
	if(t == "false")
		return 0.0;
	char* end = nullptr;
	std::string u = t;
	if(!u.empty() && (u.back() == 'L' || u.back() == 'l'))
		u.pop_back();
	const double v = std::strtod(u.c_str(), &end);
	if(end == u.c_str() || *end != '\0')
		throw std::invalid_argument("setting '" + key + "' is not a number: " + t);
	return v;
}

std::string to_string(const std::map<std::string, std::vector<std::string>>& m, const std::string& key, const std::string& fallback, bool required)
{
	auto it = m.find(key);
	if(it == m.end() || it->second.empty())
	{
		if(required)
			throw std::invalid_argument("No '" + key + "' setting in configuration file.");
		return fallback;
	}
	return it->second[0];
}
} // namespace

std::map<std::string, std::vector<std::string>> Parse_Config_Text(const std::string& text)
{
	std::map<std::string, std::vector<std::string>> m;
	Lexer L(text);
	for(;;)
	{
		L.skip();
		if(L.i >= text.size())
			break;
		const std::string key = L.name();
		L.skip();
		if(L.i >= text.size() || (text[L.i] != '=' && text[L.i] != ':'))
			L.error("expected '=' after " + key);
		L.i++;
		std::vector<std::string> v;
		L.value(v);
		L.skip();
		if(L.i < text.size() && (text[L.i] == ';' || text[L.i] == ','))
			L.i++;
		m[key] = v;
	}
	return m;
}

/* src/Parameter_Scan.cpp:90-520 (the keys of the "Parameter point" and
 * "Capture" run modes; validation messages as there) */
Configuration::Configuration(const std::string& cfg_filename)
{
	std::ifstream f(cfg_filename);
	if(!f.is_open())
		throw std::invalid_argument("Error in Configuration: cannot read " + cfg_filename);
	std::stringstream ss;
	ss << f.rdbuf();
	settings = Parse_Config_Text(ss.str());
	const auto& m = settings;
	ID = to_string(m, "ID", "default", false);
	const double ss_val = to_double(m, "sample_size", 0, true);
	if(ss_val <= 0)
		throw std::invalid_argument("Error in Configuration::Import_Parameter_Scan_Parameter(): 'sample_size' must be positive.");
	sample_size = (unsigned int) ss_val;
	const double seed = to_double(m, "fixed_seed", 0, false);
	if(seed < 0)
		throw std::invalid_argument("Error in Configuration::Import_Parameter_Scan_Parameter(): 'fixed_seed' must be non-negative.");
	fixed_seed = (unsigned int) seed;
	run_mode = to_string(m, "run_mode", "", true);
	capture_mode = (run_mode == "Capture");
	if(m.count("capture_mode"))
		capture_mode = to_double(m, "capture_mode", 0, false) != 0.0;
	if(run_mode == "Capture")
		capture_mode = true;
	const double ip = to_double(m, "interpolation_points", 0, false);
	if(ip < 0)
		throw std::invalid_argument("Error in Configuration::Import_Parameter_Scan_Parameter(): 'interpolation_points' must be non-negative.");
	interpolation_points = (unsigned int) ip;
	output_dir = to_string(m, "output_dir", "./", false);
	const double mt = to_double(m, "max_trajectories", 0, false);
	if(mt < 0)
		throw std::invalid_argument("Error in Configuration::Import_Parameter_Scan_Parameter(): 'max_trajectories' must be non-negative.");
	max_trajectories = (unsigned int) mt;
	if(m.count("maximum_number_of_scatterings"))
	{
		const double v = to_double(m, "maximum_number_of_scatterings", 0, false);
		if(!std::isfinite(v) || v < 0.0 || std::floor(v) != v)
			throw std::invalid_argument(
			    "Error in Configuration::Import_Parameter_Scan_Parameter(): 'maximum_number_of_scatterings' must be a non-negative integer.");
		maximum_number_of_scatterings = (unsigned long int) v;
	}
	snapshot_enabled = to_double(m, "snapshot_enabled", 0, false) != 0.0;
	if(snapshot_enabled && m.count("snapshot_interval"))
	{
		/* IsValidSnapshotIntervalSeconds: a positive integer number of seconds */
		const double v = to_double(m, "snapshot_interval", 60.0, false);
		if(!std::isfinite(v) || v <= 0.0 || std::floor(v) != v)
			throw std::invalid_argument(
			    "Error in Configuration::Import_Parameter_Scan_Parameter(): 'snapshot_interval' must be a positive integer number of seconds.");
		snapshot_interval_seconds = v;
	}
	max_trajectory_wall_time_sec = to_double(m, "max_trajectory_wall_time_sec", 0.0, false);
	if(!std::isfinite(max_trajectory_wall_time_sec) || max_trajectory_wall_time_sec < 0.0)
		throw std::invalid_argument(
		    "Error in Configuration::Import_Parameter_Scan_Parameter(): 'max_trajectory_wall_time_sec' must be finite and non-negative.");
	if(run_mode != "Parameter point" && run_mode != "Parameter scan" && run_mode != "Custom" && run_mode != "Capture")
		throw std::invalid_argument("Error in Configuration::Import_Parameter_Scan_Parameter(): Run mode " + run_mode + " not recognized.");

	/* Construct_DM_Particle, src/Parameter_Scan.cpp:386-520 */
	const double mass = to_double(m, "DM_mass", 0, true);
	const double fraction = to_double(m, "DM_fraction", 1.0, false);
	if(!std::isfinite(mass) || mass <= 0.0 || !std::isfinite(fraction) || fraction <= 0.0 || fraction > 1.0)
		throw std::invalid_argument("Error in Configuration::Construct_DM_Particle(): invalid DM_mass or DM_fraction.");
	const std::string interaction = to_string(m, "DM_interaction", "", true);
	DM.mass = mass * units::GeV();
	DM.low_mass = to_double(m, "DM_light", 0, true) != 0.0;
	const std::string ff = to_string(m, "DM_form_factor", "Contact", false);
	if(interaction == "SI" || interaction == "SD")
	{
		DM.interaction = (interaction == "SI") ? DMS_SI : DMS_SD;
		DM.sigma_proton = to_double(m, "DM_cross_section_nucleon", 0, true) * units::cm() * units::cm();
		DM.sigma_electron = to_double(m, "DM_cross_section_electron", 0.0, false) * units::cm() * units::cm();
		auto it = m.find("DM_relative_couplings");
		if(to_double(m, "DM_isospin_conserved", 1, false) == 0.0 && it != m.end() && it->second.size() == 2)
		{
			const double fp = std::strtod(it->second[0].c_str(), nullptr), fn = std::strtod(it->second[1].c_str(), nullptr);
			if(fp == 0.0)
				throw std::invalid_argument("Error in Configuration::Construct_DM_Particle_Standard(): f_p = 0 is not supported.");
			DM.fn_over_fp = fn / fp;
		}
		DM.Set_FormFactor_DM(ff, to_double(m, "DM_mediator_mass", 0.0, false) * units::MeV());
	}
	else if(interaction == "Dark photon")
	{
		DM.interaction = DMS_DARK_PHOTON;
		double mediator = -1.0;
		if(ff == "General")
		{
			mediator = to_double(m, "DM_mediator_mass", 0, true) * units::MeV();
			if(!std::isfinite(mediator) || mediator <= 0.0)
				throw std::invalid_argument("Error in Configuration::Construct_DM_Particle_Dark_Photon(): DM_mediator_mass must be positive.");
		}
		DM.Set_FormFactor_DM(ff, mediator);
		DM.sigma_electron = to_double(m, "DM_cross_section_electron", 0, true) * units::cm() * units::cm();
	}
	else
		throw std::invalid_argument("Error in Configuration::Construct_DM_Particle(): DM_interaction " + interaction + " not recognized.");
	if(!std::isfinite(DM.sigma_proton) || DM.sigma_proton < 0.0 || !std::isfinite(DM.sigma_electron) || DM.sigma_electron < 0.0)
		throw std::invalid_argument("Error in Configuration::Construct_DM_Particle(): cross sections must be finite and non-negative.");

	/* DM distribution: obscura Standard_Halo_Model keys */
	const std::string distr = to_string(m, "DM_distribution", "SHM", false);
	if(distr != "SHM")
		throw std::invalid_argument("Error in Configuration: DM_distribution " + distr + " is not supported (only SHM).");
	halo.rho_GeV_cm3 = to_double(m, "DM_local_density", 0.4, false);
	halo.v0_kms = to_double(m, "SHM_v0", 220.0, false);
	halo.vesc_kms = to_double(m, "SHM_vEscape", 544.0, false);
	auto vo = m.find("SHM_vObserver");
	if(vo != m.end())
	{
		if(vo->second.size() != 3)
			throw std::invalid_argument("Error in Configuration: SHM_vObserver needs three components.");
		for(int k = 0; k < 3; k++)
			halo.vobs_kms[k] = std::strtod(vo->second[k].c_str(), nullptr);
	}
}

void Configuration::Print_Summary() const
{
	std::cout << SEPARATOR << "Configuration " << ID << "\n\n"
	          << "\tRun mode:\t\t\t" << run_mode << "\n"
	          << "\tCapture mode:\t\t\t" << (capture_mode ? "[x]" : "[ ]") << "\n"
	          << "\tSample size:\t\t\t" << sample_size << "\n"
	          << "\tMax trajectories:\t\t" << max_trajectories << "\n"
	          << "\tFixed seed:\t\t\t" << fixed_seed << "\n"
	          << "\tInterpolation points:\t\t" << interpolation_points << "\n"
	          << "\tDM mass [GeV]:\t\t\t" << DM.mass / units::GeV() << "\n"
	          << "\tsigma_p [cm2]:\t\t\t" << DM.sigma_proton / (units::cm() * units::cm()) << "\n"
	          << "\tsigma_e [cm2]:\t\t\t" << DM.sigma_electron / (units::cm() * units::cm()) << "\n"
	          << SEPARATOR << std::endl;
}

void DM_Particle::Set_FormFactor_DM(const std::string& ff, double mediator)
{
	if(ff == "Contact")
		form_factor = DMS_FF_CONTACT;
	else if(ff == "General")
	{
		form_factor = DMS_FF_GENERAL;
		mediator_mass = mediator;
	}
	else
		throw std::invalid_argument("DM form factor " + ff + " is not supported (Contact or General).");
}

DM_Particle DM_Particle_SI(double mass)
{
	DM_Particle d;
	d.mass = mass;
	d.interaction = DMS_SI;
	return d;
}

DM_Particle DM_Particle_Dark_Photon(double mass, double sigma_electron)
{
	DM_Particle d;
	d.mass = mass;
	d.interaction = DMS_DARK_PHOTON;
	d.sigma_electron = sigma_electron;
	return d;
}

dms_point_desc DM_Particle::Descriptor(int interpolation_points, int speed_points) const
{
	dms_point_desc p{};
	p.mass_GeV = mass / units::GeV();
	p.sigma_p_cm2 = sigma_proton / (units::cm() * units::cm());
	p.sigma_e_cm2 = sigma_electron / (units::cm() * units::cm());
	p.interaction = interaction;
	p.form_factor = form_factor;
	p.mediator_mass_MeV = mediator_mass / units::MeV();
	p.fn_over_fp = fn_over_fp;
	p.low_mass = low_mass ? 1 : 0;
	p.interpolation_points = interpolation_points;
	p.speed_points = speed_points;
	p.host_rate_grid = nullptr;
	return p;
}

dms_halo Standard_Halo_Model::Descriptor() const
{
	dms_halo h;
	h.rho_GeV_cm3 = rho_GeV_cm3;
	h.v0_kms = v0_kms;
	h.vesc_kms = vesc_kms;
	for(int k = 0; k < 3; k++)
		h.vobs_kms[k] = vobs_kms[k];
	return h;
}

double Event::Radius() const { return std::sqrt(position[0] * position[0] + position[1] * position[1] + position[2] * position[2]); }
double Event::Speed() const { return std::sqrt(velocity[0] * velocity[0] + velocity[1] * velocity[1] + velocity[2] * velocity[2]); }
double Event::Angular_Momentum() const
{
	const double x = position[1] * velocity[2] - position[2] * velocity[1];
	const double y = position[2] * velocity[0] - position[0] * velocity[2];
	const double z = position[0] * velocity[1] - position[1] * velocity[0];
	return std::sqrt(x * x + y * y + z * z);
}

void Trajectory_Result::Print_Summary() const
{
	std::cout << "Trajectory " << trajectory_id << ": " << Termination_Reason_Name(bincount.termination_reason) << ", "
	          << number_of_scatterings << " scatterings, " << (Particle_Captured() ? "captured" : "not captured")
	          << (Particle_Free() ? ", free" : "") << (Particle_Reflected() ? ", reflected" : "") << std::endl;
}

} // namespace dmsun
