/*
 * dmsun_serve: a request stream of parameter points served with continuous
 * batching on one MI355X: every request's trajectories join one resident
 * trajectory launch as slots (dmsun::Server over dms_server).
 *
 *   dmsun_serve [--requests N] [--rate R] [--sample-size S] [--sample-size-max S2]
 *               [--max-trajectories M] [--max-scatterings K] [--rk-budget B]
 *               [--precision fp64|mixed] [--concurrency C] [--memory-budget GiB]
 *               [--per-job-budget GiB] [--seed K] [--interp N] [--bins N]
 *               [--check-alone K] [--json PATH] [--results PATH] [--device D] [--quiet]
 *
 * Requests arrive as a Poisson process of rate R per second.  Each asks for a
 * spin-independent DM point with the mass log-uniform in [0.1, 5] GeV and the
 * proton cross section log-uniform in [1e-38, 1e-35] cm^2, in parameter-point
 * mode with a target of S captured particles (log-uniform in [S, S2] with
 * --sample-size-max), a cap of M trajectories and of K scatterings per
 * trajectory (heavy captured particles do not evaporate: K bounds them).
 * Latency is completion minus arrival.  --check-alone K reruns the first K
 * requests one at a time after the mix and compares their results with the
 * ones they got while sharing the GPU.  One JSON line summarises the run
 * (with the resident launch's counters: launches started, slots published,
 * slots handed back unstarted after their request had stopped);
 * --results writes one JSON line per request: its point and job descriptor,
 * counters, visit counts, histogram sums and evaporation events.
 */
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <thread>

#include "dmsun.hpp"

namespace
{
struct Options
{
	int requests = 64;
	double rate = 8.0;
	uint64_t sample_size = 20, sample_size_max = 0, max_trajectories = 200000;
	uint64_t max_scatterings = 100000, rk_budget = 0;
	int precision = DMS_FP64;
	int concurrency = 64;
	double memory_budget_gib = 128.0, per_job_budget_gib = 8.0;
	unsigned int seed = 20260710;
	unsigned int interp = 1000;
	int bins = 2000;
	int check_alone = 0;
	std::string json, results;
	int device = 0;
	bool quiet = false;
};

/* one request and its result as a JSON line (doubles round-trip: %.17g) */
std::string result_json(const dmsun::Request& q, const dmsun::Request_Result& r)
{
	const dms_point_desc d = q.DM.Descriptor((int) q.interpolation_points);
	const dms_counters& c = r.counters;
	std::ostringstream s;
	s.precision(17);
	s << "{\"id\": " << r.id << ", \"status\": " << r.status << ", \"mass_GeV\": " << d.mass_GeV
	  << ", \"sigma_p_cm2\": " << d.sigma_p_cm2 << ", \"sigma_e_cm2\": " << d.sigma_e_cm2 << ", \"interp\": " << q.interpolation_points
	  << ", \"sample_size\": " << q.sample_size << ", \"max_trajectories\": " << q.max_trajectories
	  << ", \"max_scatterings\": " << q.max_scatterings << ", \"rk_budget\": " << q.rk_attempt_budget << ", \"seed\": " << q.seed
	  << ", \"point_id\": " << q.point_id << ", \"precision\": " << q.precision << ", \"bins\": " << q.n_bins
	  << ", \"halo\": {\"rho\": " << q.halo.rho_GeV_cm3 << ", \"v0\": " << q.halo.v0_kms << ", \"vesc\": " << q.halo.vesc_kms
	  << ", \"vobs\": [" << q.halo.vobs_kms[0] << ", " << q.halo.vobs_kms[1] << ", " << q.halo.vobs_kms[2] << "]}"
	  << ", \"latency_sec\": " << r.Latency() << ", \"counters\": {\"trajectories\": " << c.trajectories
	  << ", \"captured\": " << c.captured << ", \"completed_outward_escapes\": " << c.completed_outward_escapes
	  << ", \"free_particles\": " << c.free_particles << ", \"reflected_particles\": " << c.reflected_particles
	  << ", \"complete_evaporation\": " << c.complete_evaporation << ", \"invalid_survival_captured\": " << c.invalid_survival_captured
	  << ", \"numerical_failures\": " << c.numerical_failures << ", \"computational_truncations\": " << c.computational_truncations
	  << ", \"total_scatterings\": " << c.total_scatterings << ", \"sync_rounds\": " << c.sync_rounds
	  << ", \"capture_target_overshoot\": " << c.capture_target_overshoot << ", \"early_stop_reason\": " << c.early_stop_reason
	  << ", \"reason_counts\": [";
	for(int k = 0; k < 11; k++)
		s << (k ? ", " : "") << c.reason_counts[k];
	s << "]}, \"events\": [";
	for(size_t k = 0; k < r.events.size(); k++)
		s << (k ? ", " : "") << "[" << r.events[k].trajectory_id << ", " << r.events[k].lifetime_unbinding_sec << "]";
	s << "], \"counts\": [";
	for(size_t k = 0; k < r.counts.size(); k++)
		s << (k ? ", " : "") << r.counts[k];
	s << "], \"hist\": [";
	for(size_t k = 0; k < r.hist.size(); k++)
		s << (k ? ", " : "") << r.hist[k];
	s << "]}";
	return s.str();
}

bool same_result(const dmsun::Request_Result& a, const dmsun::Request_Result& b, std::string& why)
{
	const dms_counters &x = a.counters, &y = b.counters;
	if(a.status != DMS_OK || b.status != DMS_OK)
	{
		why = "status";
		return false;
	}
	if(x.trajectories != y.trajectories || x.captured != y.captured || x.total_scatterings != y.total_scatterings
	   || x.complete_evaporation != y.complete_evaporation || x.sync_rounds != y.sync_rounds
	   || std::memcmp(x.reason_counts, y.reason_counts, sizeof(x.reason_counts)) != 0)
	{
		why = "counters";
		return false;
	}
	if(a.counts != b.counts)
	{
		why = "visit counts";
		return false;
	}
	if(a.events.size() != b.events.size())
	{
		why = "event count";
		return false;
	}
	for(size_t i = 0; i < a.events.size(); i++)
		if(a.events[i].trajectory_id != b.events[i].trajectory_id
		   || a.events[i].lifetime_unbinding_sec != b.events[i].lifetime_unbinding_sec)
		{
			why = "events";
			return false;
		}
	/* histogram sums are accumulated with device atomics: equal to rounding */
	for(size_t i = 0; i < a.hist.size(); i++)
	{
		const double d = std::fabs(a.hist[i] - b.hist[i]), s = std::fabs(a.hist[i]) + std::fabs(b.hist[i]);
		if(d > 1e-12 * s)
		{
			why = "histogram";
			return false;
		}
	}
	return true;
}
} // namespace

int main(int argc, char** argv)
{
	Options o;
	for(int i = 1; i < argc; i++)
	{
		const std::string a = argv[i];
		auto next = [&](void) -> std::string {
			if(i + 1 >= argc)
				throw std::invalid_argument("missing value after " + a);
			return argv[++i];
		};
		try
		{
			if(a == "--requests")
				o.requests = std::stoi(next());
			else if(a == "--rate")
				o.rate = std::stod(next());
			else if(a == "--sample-size")
				o.sample_size = std::stoull(next());
			else if(a == "--sample-size-max")
				o.sample_size_max = std::stoull(next());
			else if(a == "--max-scatterings")
				o.max_scatterings = std::stoull(next());
			else if(a == "--rk-budget")
				o.rk_budget = std::stoull(next());
			else if(a == "--max-trajectories")
				o.max_trajectories = std::stoull(next());
			else if(a == "--precision")
			{
				const std::string p = next();
				if(p != "fp64" && p != "mixed")
					throw std::invalid_argument("precision must be fp64 or mixed");
				o.precision = (p == "mixed") ? DMS_MIXED : DMS_FP64;
			}
			else if(a == "--concurrency")
				o.concurrency = std::stoi(next());
			else if(a == "--memory-budget")
				o.memory_budget_gib = std::stod(next());
			else if(a == "--per-job-budget")
				o.per_job_budget_gib = std::stod(next());
			else if(a == "--seed")
				o.seed = (unsigned int) std::stoul(next());
			else if(a == "--interp")
				o.interp = (unsigned int) std::stoul(next());
			else if(a == "--bins")
				o.bins = std::stoi(next());
			else if(a == "--check-alone")
				o.check_alone = std::stoi(next());
			else if(a == "--json")
				o.json = next();
			else if(a == "--results")
				o.results = next();
			else if(a == "--device")
				o.device = std::stoi(next());
			else if(a == "--quiet")
				o.quiet = true;
			else
				throw std::invalid_argument("unknown option " + a);
		}
		catch(const std::exception& e)
		{
			std::cerr << e.what() << std::endl;
			return 2;
		}
	}
	if(o.requests < 1 || o.rate <= 0.0 || o.concurrency < 1 || o.check_alone < 0 || o.check_alone > o.requests)
	{
		std::cerr << "invalid request count, rate, concurrency or --check-alone" << std::endl;
		return 2;
	}
	try
	{
		dmsun::Solar_Model SSM(o.device);
		const size_t gib = (size_t) 1 << 30;
		std::vector<dmsun::Request> reqs;
		std::mt19937_64 rng(o.seed);
		std::uniform_real_distribution<double> u(0.0, 1.0);
		std::exponential_distribution<double> gap(o.rate);
		std::vector<double> gaps;
		for(int k = 0; k < o.requests; k++)
		{
			dmsun::Request r;
			const double m = std::pow(10.0, std::log10(0.1) + u(rng) * (std::log10(5.0) - std::log10(0.1)));
			const double s = std::pow(10.0, -38.0 + 3.0 * u(rng));
			r.DM = dmsun::DM_Particle_SI(m);
			r.DM.Set_Sigma_Proton(s * dmsun::units::cm() * dmsun::units::cm());
			r.sample_size = o.sample_size;
			if(o.sample_size_max > o.sample_size) /* log-uniform */
				r.sample_size = (uint64_t) std::llround(std::exp(std::log((double) o.sample_size)
				                                                 + u(rng) * std::log((double) o.sample_size_max / (double) o.sample_size)));
			r.max_scatterings = o.max_scatterings;
			r.rk_attempt_budget = o.rk_budget;
			r.max_trajectories = o.max_trajectories;
			r.seed = o.seed;
			r.point_id = (unsigned int) k;
			r.precision = o.precision;
			r.n_bins = o.bins;
			r.interpolation_points = o.interp;
			reqs.push_back(r);
			gaps.push_back(gap(rng));
		}

		std::vector<dmsun::Request_Result> mix;
		double elapsed = 0.0;
		dms_server_stats sst{};
		{
			dmsun::Server server(SSM, (size_t) (o.memory_budget_gib * gib), (size_t) (o.per_job_budget_gib * gib), o.concurrency);
			const auto t0 = std::chrono::steady_clock::now();
			double t_next = 0.0;
			for(int k = 0; k < o.requests; k++)
			{
				t_next += gaps[k];
				std::this_thread::sleep_until(t0 + std::chrono::duration<double>(t_next));
				server.Submit(reqs[k]);
			}
			int reported = 0;
			while(true)
			{
				const std::vector<dmsun::Request_Result> now = server.Results();
				if(!o.quiet && (int) now.size() != reported)
				{
					reported = (int) now.size();
					std::cerr << "served " << reported << "/" << o.requests << " at " << server.Elapsed() << " s" << std::endl;
				}
				if((int) now.size() == o.requests)
					break;
				std::this_thread::sleep_for(std::chrono::milliseconds(50));
			}
			server.Wait_All();
			elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
			mix = server.Results();
			sst = server.Stats();
		}

		int failed = 0;
		uint64_t traj = 0, captured = 0, simulated = 0;
		std::vector<double> lat, wait;
		for(const auto& r : mix)
		{
			if(r.status != DMS_OK)
			{
				failed++;
				std::cerr << "request " << r.id << " failed: " << r.error << std::endl;
				continue;
			}
			lat.push_back(r.Latency());
			wait.push_back(r.start_sec - r.arrival_sec);
			traj += r.counters.trajectories;
			captured += r.counters.captured;
			simulated += r.counters.simulated_trajectories;
		}

		int mismatches = 0;
		if(o.check_alone > 0)
		{
			dmsun::Server alone(SSM, (size_t) (o.memory_budget_gib * gib), (size_t) (o.per_job_budget_gib * gib), 1);
			for(int k = 0; k < o.check_alone; k++)
				alone.Submit(reqs[k]);
			alone.Wait_All();
			const std::vector<dmsun::Request_Result> solo = alone.Results();
			for(int k = 0; k < o.check_alone; k++)
			{
				std::string why;
				if(!same_result(mix[k], solo[k], why))
				{
					mismatches++;
					std::cerr << "request " << k << " differs alone and in the mix: " << why << std::endl;
				}
			}
		}

		std::ostringstream js;
		js.precision(6);
		js << "{\"requests\": " << o.requests << ", \"failed\": " << failed << ", \"rate_per_s\": " << o.rate
		   << ", \"precision\": \"" << (o.precision == DMS_MIXED ? "mixed" : "fp64") << "\", \"concurrency\": " << o.concurrency
		   << ", \"sample_size\": " << o.sample_size << ", \"sample_size_max\": " << std::max(o.sample_size, o.sample_size_max)
		   << ", \"max_trajectories\": " << o.max_trajectories << ", \"max_scatterings\": " << o.max_scatterings
		   << ", \"elapsed_sec\": " << elapsed << ", \"requests_per_s\": " << (double) (o.requests - failed) / elapsed
		   << ", \"trajectories\": " << traj << ", \"captured\": " << captured
		   << ", \"trajectories_per_s\": " << (double) traj / elapsed
		   << ", \"latency_p50_sec\": " << dmsun::Percentile(lat, 50.0) << ", \"latency_p99_sec\": " << dmsun::Percentile(lat, 99.0)
		   << ", \"latency_max_sec\": " << dmsun::Percentile(lat, 100.0) << ", \"queue_wait_p50_sec\": " << dmsun::Percentile(wait, 50.0)
		   << ", \"queue_wait_p99_sec\": " << dmsun::Percentile(wait, 99.0) << ", \"checked_alone\": " << o.check_alone
		   << ", \"alone_mismatches\": " << mismatches << ", \"simulated_trajectories\": " << simulated
		   << ", \"launches\": " << sst.launches << ", \"slots_published\": " << sst.slots_published
		   << ", \"slots_cancelled\": " << sst.slots_cancelled << "}";
		std::cout << js.str() << std::endl;
		if(!o.json.empty())
		{
			std::ofstream f(o.json);
			f << js.str() << "\n";
		}
		if(!o.results.empty())
		{
			std::ofstream f(o.results);
			for(size_t k = 0; k < mix.size(); k++)
				f << result_json(reqs[mix[k].id], mix[k]) << "\n";
		}
		return (failed == 0 && mismatches == 0) ? 0 : 1;
	}
	catch(const std::exception& e)
	{
		std::cerr << "error: " << e.what() << std::endl;
		return 1;
	}
}
