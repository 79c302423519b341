/*
 * GPU-backed parts of the C++ API: every computation goes through the C ABI
 * (include/dms.h), i.e. the HIP kernels of libdmsun.so.
 */
#include <chrono>
#include <cmath>
#include <thread>
#include <cstring>
#include <iostream>
#include <limits>
#include <stdexcept>

#include "dmsun.hpp"

namespace dmsun
{

bool Ensure_Directory(const std::string& dir);

namespace units
{
static double c(const char* n) { return dms_constant(n); }
double cm() { return c("cm"); }
double km() { return c("km"); }
double sec() { return c("sec"); }
double gram() { return c("gram"); }
double Kelvin() { return c("Kelvin"); }
double GeV() { return c("GeV"); }
double MeV() { return c("MeV"); }
double eV() { return c("eV"); }
double rSun() { return c("rSun"); }
double mSun() { return c("mSun"); }
double AU() { return c("AU"); }
double G_Newton() { return c("G_Newton"); }
} // namespace units

namespace
{
[[noreturn]] void raise(int status, const std::string& what, const dms_ctx* ctx)
{
	const std::string msg = what + ": " + dms_last_error(ctx);
	if(status == DMS_ERR_ARG || status == DMS_ERR_UNSUPPORTED)
		throw std::invalid_argument(msg);
	throw std::runtime_error(msg);
}

void check(int status, const char* what, const dms_ctx* ctx)
{
	if(status != DMS_OK)
		raise(status, what, ctx);
}

bool same_point(const dms_point_desc& a, const dms_point_desc& b)
{
	return a.mass_GeV == b.mass_GeV && a.sigma_p_cm2 == b.sigma_p_cm2 && a.sigma_e_cm2 == b.sigma_e_cm2
	       && a.interaction == b.interaction && a.form_factor == b.form_factor && a.mediator_mass_MeV == b.mediator_mass_MeV
	       && a.fn_over_fp == b.fn_over_fp && a.low_mass == b.low_mass;
}

Event to_event(const dms_event& e) { return Event(e.t, {e.r[0], e.r[1], e.r[2]}, {e.v[0], e.v[1], e.v[2]}); }

dms_event from_event(const Event& e)
{
	dms_event o;
	o.t = e.time;
	for(int k = 0; k < 3; k++)
	{
		o.r[k] = e.position[k];
		o.v[k] = e.velocity[k];
	}
	return o;
}

Trajectory_Result to_result(const dms_traj_record& r)
{
	Trajectory_Result t;
	t.initial_event = to_event(r.initial);
	t.final_event = to_event(r.final);
	t.number_of_scatterings = r.n_scatterings;
	t.flags = r.flags;
	t.trajectory_id = r.trajectory_id;
	TrajectoryBincount& b = t.bincount;
	b.is_captured = (r.flags & DMS_F_CAPTURED) != 0;
	b.t_capture = r.t_capture_sec;
	b.t_last_bound = r.t_last_bound_sec;
	b.t_final_unbinding_scatter = r.t_final_unbinding_scatter_sec;
	b.t_boundary_escape = r.t_boundary_escape_sec;
	b.t_termination = r.t_termination_sec;
	b.r_first_negative_km = r.r_first_negative_km;
	b.E_first_negative_eV = r.E_first_negative_eV;
	b.dE_first_negative_from_prev_eV = r.dE_first_negative_from_prev_eV;
	b.event_observed = (r.flags & DMS_F_EVENT_OBSERVED) != 0;
	b.boundary_escape_observed = (r.flags & DMS_F_BOUNDARY_ESCAPE_OBSERVED) != 0;
	b.survival_valid = (r.flags & DMS_F_SURVIVAL_VALID) != 0;
	b.numerically_invalid_escape = (r.flags & DMS_F_NUMERICALLY_INVALID_ESCAPE) != 0;
	b.truncated = (r.flags & DMS_F_TRUNCATED) != 0;
	b.max_free_energy_drift_eV = r.max_free_energy_drift_eV;
	b.max_free_energy_drift_rel = r.max_free_energy_drift_rel;
	b.number_of_scatterings = r.n_scatterings;
	b.termination_reason = (TrajectoryTerminationReason) r.reason;
	return t;
}
} // namespace

/* ---- Solar_Model -------------------------------------------------------- */

Solar_Model::Solar_Model(int device, const std::string& model_file)
{
	std::string path = model_file;
	if(path.empty())
	{
		const char* env = std::getenv("DMSUN_SOLAR_MODEL");
		path = env ? env : "data/model_agss09.dat";
	}
	size_t n = 0;
	if(dms_load_solar_file(path.c_str(), nullptr, 0, &n) != DMS_OK || n == 0)
		throw std::invalid_argument("Solar_Model: cannot read the solar model " + path);
	std::vector<double> rows(n * 35);
	check(dms_load_solar_file(path.c_str(), rows.data(), n, &n), "Solar_Model: dms_load_solar_file", nullptr);
	check(dms_create(&ctx, device, rows.data(), n), "Solar_Model: dms_create", nullptr);
	std::vector<int32_t> Z(128), A(128);
	std::vector<double> m(128), ab(128);
	size_t k = 0;
	check(dms_isotopes(ctx, Z.data(), A.data(), m.data(), ab.data(), 128, &k), "Solar_Model: dms_isotopes", ctx);
	for(size_t i = 0; i < k; i++)
		isotopes.push_back({Z[i], A[i], m[i], ab[i]});
}

Solar_Model::~Solar_Model()
{
	if(point)
		dms_point_destroy(point);
	dms_destroy(ctx);
}

double Solar_Model::eval(int which, double r) const
{
	double out = 0.0;
	check(dms_solar_eval(ctx, which, &r, &out, 1), "Solar_Model: dms_solar_eval", ctx);
	return out;
}

double Solar_Model::Mass(double r) const { return eval(0, r); }
double Solar_Model::Temperature(double r) const { return eval(1, r); }
double Solar_Model::Local_Escape_Speed(double r) const { return eval(2, r); }
double Solar_Model::Mass_Density(double r) const { return eval(3, r); }
double Solar_Model::Number_Density_Electron(double r) const { return eval(4, r); }
double Solar_Model::Number_Density_Nucleus(double r, unsigned int isotope) const
{
	if(isotope >= isotopes.size())
		throw std::invalid_argument("Solar_Model::Number_Density_Nucleus: isotope index out of range");
	return eval(100 + (int) isotope, r);
}

void Solar_Model::Interpolate_Total_DM_Scattering_Rate(const DM_Particle& DM, unsigned int N_r, unsigned int N_v)
{
	const bool grid = N_r >= 2 && N_v >= 2;
	if(grid && ((unsigned long long) N_r * N_v > (unsigned long long) std::numeric_limits<int>::max()))
		throw std::invalid_argument("scattering-rate interpolation grid is too large");
	dms_point* p = nullptr;
	const dms_point_desc d = grid ? DM.Descriptor((int) N_r, (int) N_v) : DM.Descriptor(0);
	check(dms_point_create(ctx, &d, &p), "Interpolate_Total_DM_Scattering_Rate", ctx);
	if(point)
		dms_point_destroy(point);
	point = p;
	point_desc = d;
	radius_points = grid ? N_r : 0;
	speed_points = grid ? N_v : 0;
}

dms_point* Solar_Model::Point_For(const DM_Particle& DM)
{
	const dms_point_desc d = DM.Descriptor((int) radius_points, (int) speed_points);
	if(point && same_point(d, point_desc))
		return point;
	/* no table for this particle: direct rates, as the reference without a table */
	dms_point* p = nullptr;
	dms_point_desc direct = DM.Descriptor(0);
	check(dms_point_create(ctx, &direct, &p), "Solar_Model: dms_point_create", ctx);
	if(point)
		dms_point_destroy(point);
	point = p;
	point_desc = direct;
	radius_points = 0;
	speed_points = 0;
	return point;
}

double Solar_Model::rate(const DM_Particle& DM, int which, double r, double v)
{
	dms_point* p = Point_For(DM);
	double out = 0.0;
	check(dms_point_rates(p, which, &r, &v, &out, 1), "Solar_Model: dms_point_rates", ctx);
	return out;
}

double Solar_Model::Total_DM_Scattering_Rate(const DM_Particle& DM, double r, double v) { return rate(DM, 0, r, v); }
double Solar_Model::Total_DM_Scattering_Rate_Computed(const DM_Particle& DM, double r, double v) { return rate(DM, 1, r, v); }

void Solar_Model::Print_Summary() const
{
	std::cout << "----------------------------------------\n"
	          << name << "\n\tTarget isotopes:\t" << isotopes.size() << "\n\tRate interpolation points:\t" << radius_points << " x " << speed_points
	          << "\n----------------------------------------" << std::endl;
}

/* ---- Trajectory_Simulator --------------------------------------------------- */

Trajectory_Simulator::Trajectory_Simulator(Solar_Model& model, unsigned long int max_time_steps, unsigned long int max_scatterings,
                                           double max_distance)
    : maximum_time_steps(max_time_steps), maximum_scatterings(max_scatterings),
      maximum_distance(max_distance < 0.0 ? 2.0 * units::rSun() : max_distance), solar_model(model)
{
	/* src/Simulation_Trajectory.cpp:590-591 */
	if(!std::isfinite(maximum_distance) || maximum_distance <= units::rSun())
		throw std::invalid_argument("Trajectory_Simulator: maximum distance must exceed the solar radius");
}

dms_job_desc Trajectory_Simulator::job(unsigned int rank, size_t n) const
{
	dms_job_desc j{};
	j.mode = capture_mode ? DMS_MODE_CAPTURE : DMS_MODE_PARAMETER_POINT;
	j.precision = DMS_FP64;
	j.max_trajectories = n;
	j.capture_target = 0;
	j.seed = seed;
	j.point_id = rank;
	j.first_trajectory_id = current_trajectory_id;
	j.max_scatterings = maximum_scatterings;
	j.max_free_steps = maximum_time_steps;
	j.rk_attempt_budget = rk_attempt_budget;
	j.r_max_rsun = maximum_distance / units::rSun();
	j.n_bins = 2000;
	j.record_trajectories = 1;
	j.halo = Standard_Halo_Model().Descriptor();
	return j;
}

Trajectory_Result Trajectory_Simulator::Simulate(const Event& initial_condition, const DM_Particle& DM, unsigned int mpi_rank,
                                                 bool histogram_rows)
{
	current_mpi_rank = mpi_rank;
	dms_point* p = solar_model.Point_For(DM);
	const dms_job_desc j = job(mpi_rank, 1);
	const dms_event e = from_event(initial_condition);
	dms_traj_record rec;
	std::vector<double> rows(histogram_rows ? 2 * 2000 : 0);
	check(dms_simulate_events(p, &j, &e, 1, &rec, histogram_rows ? rows.data() : nullptr), "Trajectory_Simulator::Simulate",
	      solar_model.Context());
	current_trajectory_id++;
	Trajectory_Result t = to_result(rec);
	if(histogram_rows)
	{
		t.bincount.dt_hist.assign(rows.begin(), rows.begin() + 2000);
		t.bincount.v2dt_hist.assign(rows.begin() + 2000, rows.end());
	}
	return t;
}

std::vector<Trajectory_Result> Trajectory_Simulator::Simulate(const std::vector<Event>& initial_conditions, const DM_Particle& DM,
                                                              unsigned int mpi_rank)
{
	current_mpi_rank = mpi_rank;
	std::vector<Trajectory_Result> out;
	if(initial_conditions.empty())
		return out;
	dms_point* p = solar_model.Point_For(DM);
	const dms_job_desc j = job(mpi_rank, initial_conditions.size());
	std::vector<dms_event> ev;
	for(const Event& e : initial_conditions)
		ev.push_back(from_event(e));
	std::vector<dms_traj_record> recs(ev.size());
	check(dms_simulate_events(p, &j, ev.data(), ev.size(), recs.data(), nullptr), "Trajectory_Simulator::Simulate",
	      solar_model.Context());
	current_trajectory_id += ev.size();
	for(const dms_traj_record& r : recs)
		out.push_back(to_result(r));
	return out;
}

/* ---- Simulation_Data::Generate_Data (src/Data_Generation.cpp:488-972) ----- */

void Simulation_Data::Generate_Data(const DM_Particle& DM, Solar_Model& solar_model, const Standard_Halo_Model& halo_model,
                                    unsigned int fixed_seed, bool capture_mode, unsigned int point_id)
{
	dms_point* p = solar_model.Point_For(DM);
	dms_job_desc j{};
	j.mode = capture_mode ? DMS_MODE_CAPTURE : DMS_MODE_PARAMETER_POINT;
	j.precision = DMS_FP64;
	j.capture_target = requested_captured_particles;
	j.max_trajectories = max_trajectories;
	j.seed = fixed_seed;
	j.point_id = point_id;
	j.first_trajectory_id = 0;
	j.max_scatterings = maximum_number_of_scatterings;
	j.max_free_steps = maximum_free_time_steps;
	j.rk_attempt_budget = rk_attempt_budget;
	j.r_max_rsun = (initial_and_final_radius > 0.0 ? initial_and_final_radius : 2.0 * units::rSun()) / units::rSun();
	j.n_bins = number_of_bins;
	j.record_trajectories = 0;
	j.halo = halo_model.Descriptor();
	dms_job* job = nullptr;
	check(dms_job_submit(p, &j, &job), "Simulation_Data::Generate_Data", solar_model.Context());
	dms_job_status st;
	const bool snapshots = snapshot_interval > 0.0 && !snapshot_dir.empty() && Ensure_Directory(snapshot_dir);
	const std::string snap_root = snapshot_dir.empty() ? "" : (snapshot_dir.back() == '/' ? snapshot_dir : snapshot_dir + "/");
	const auto t0 = std::chrono::steady_clock::now();
	auto elapsed = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
	Snapshot_Report rep;
	rep.interval_s = snapshot_interval;
	rep.run_id = (uint64_t) std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::system_clock::now().time_since_epoch()).count();
	rep.mass_GeV = DM.mass / units::GeV();
	rep.sigma_cm2 = (DM.interaction == DMS_DARK_PHOTON ? DM.sigma_electron : DM.sigma_proton) / (units::cm() * units::cm());
	rep.n_bins = number_of_bins;
	auto snapshot_path = [&](long long label) { return snap_root + "snapshot_" + std::to_string(label) + "s.txt"; };
	int snapshot_index = 0;
	if(snapshots)
	{
		/* SnapshotTimeLabelSeconds: checkpoint k at k x interval seconds of wall time */
		while(dms_job_poll(job, &st) == DMS_OK && !st.done)
		{
			if(elapsed() >= (snapshot_index + 1) * snapshot_interval)
			{
				snapshot_index++;
				rep.time_label_s = (long long) (snapshot_index * snapshot_interval);
				rep.total_trajectories = st.trajectories_done;
				rep.captured_particles = st.captured;
				rep.valid_trajectories = st.captured + st.completed_outward_escapes;
				rep.numerical_failures = st.numerical_failures;
				rep.scatterings = st.scatterings;
				rep.elapsed_s = elapsed();
				Write_Snapshot_Report(snapshot_path(rep.time_label_s), rep);
			}
			std::this_thread::sleep_for(std::chrono::milliseconds(50));
		}
	}
	const int ws = dms_job_wait(job, &st);
	if(ws != DMS_OK)
	{
		const std::string msg = std::string("Simulation_Data::Generate_Data: ") + dms_last_error(solar_model.Context());
		dms_job_destroy(job);
		throw std::runtime_error(msg);
	}
	size_t n_ev = 0;
	dms_job_read(job, nullptr, nullptr, nullptr, 0, nullptr, 0, &n_ev, nullptr, 0, nullptr);
	std::vector<double> h(8 * (size_t) number_of_bins);
	std::vector<uint64_t> c(2 * (size_t) number_of_bins);
	std::vector<dms_evap_event> ev(n_ev);
	dms_counters cnt;
	const int rs = dms_job_read(job, &cnt, h.data(), c.data(), (size_t) number_of_bins, ev.data(), ev.size(), &n_ev, nullptr, 0,
	                            nullptr);
	dms_job_destroy(job);
	check(rs, "Simulation_Data::Generate_Data: dms_job_read", solar_model.Context());
	ev.resize(n_ev);
	Set_Results(cnt, h, c, ev);
	if(snapshots)
	{
		/* the first checkpoint after the job finished carries its final results */
		snapshot_index = std::max(snapshot_index + 1, (int) std::ceil(elapsed() / snapshot_interval));
		rep.time_label_s = (long long) (snapshot_index * snapshot_interval);
		rep.total_trajectories = cnt.trajectories;
		rep.captured_particles = cnt.captured;
		rep.valid_trajectories = cnt.captured + cnt.completed_outward_escapes;
		rep.numerical_failures = cnt.numerical_failures;
		rep.scatterings = cnt.total_scatterings;
		rep.done = true;
		rep.elapsed_s = elapsed();
		rep.hist = capture_mode ? nullptr : &h;
		Write_Snapshot_Report(snapshot_path(rep.time_label_s), rep);
	}
}

} // namespace dmsun
