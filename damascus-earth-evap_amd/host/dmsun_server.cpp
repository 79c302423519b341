/*
 * Serving loop: continuous batching of parameter-point requests over the
 * concurrent jobs of one context (each job on its own streams).
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <stdexcept>
#include <thread>

#include "dmsun.hpp"

namespace dmsun
{

double Percentile(std::vector<double> v, double p)
{
	if(v.empty())
		return 0.0;
	std::sort(v.begin(), v.end());
	const double x = (p / 100.0) * (double) (v.size() - 1);
	const size_t i = (size_t) x;
	const double f = x - (double) i;
	return (i + 1 < v.size()) ? v[i] * (1.0 - f) + v[i + 1] * f : v.back();
}

namespace
{
bool same_desc(const dms_point_desc& a, const dms_point_desc& b)
{
	return a.mass_GeV == b.mass_GeV && a.sigma_p_cm2 == b.sigma_p_cm2 && a.sigma_e_cm2 == b.sigma_e_cm2
	       && a.interaction == b.interaction && a.form_factor == b.form_factor && a.mediator_mass_MeV == b.mediator_mass_MeV
	       && a.fn_over_fp == b.fn_over_fp && a.low_mass == b.low_mass && a.interpolation_points == b.interpolation_points
	       && a.speed_points == b.speed_points;
}
} // namespace

struct Server::Impl
{
	Solar_Model& model;
	size_t memory_budget, per_job_budget;
	int max_concurrent;
	std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

	struct Pending
	{
		uint64_t id;
		Request req;
		double arrival;
	};
	struct Running
	{
		uint64_t id;
		Request req;
		double arrival, start;
		dms_job* job;
	};
	std::mutex mu;
	std::condition_variable cv;
	std::deque<Pending> queue;
	std::vector<Running> running;
	std::map<uint64_t, Request_Result> done;
	std::vector<std::pair<dms_point_desc, dms_point*>> points; /* resident parameter points */
	uint64_t next_id = 0;
	bool stop = false;
	std::thread loop;

	Impl(Solar_Model& m, size_t mb, size_t pj, int mc) : model(m), memory_budget(mb), per_job_budget(pj), max_concurrent(mc)
	{
		if(dms_set_job_memory_budget(model.Context(), per_job_budget) != DMS_OK)
			throw std::invalid_argument(dms_last_error(model.Context()));
		loop = std::thread([this] { run(); });
	}
	~Impl()
	{
		{
			std::lock_guard<std::mutex> lock(mu);
			stop = true;
		}
		cv.notify_all();
		loop.join();
		for(auto& r : running)
			dms_job_destroy(r.job);
		for(auto& p : points)
			dms_point_destroy(p.second);
	}
	double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

	dms_point* point_for(const Request& r)
	{
		const dms_point_desc d = r.DM.Descriptor((int) r.interpolation_points);
		for(auto& p : points)
			if(same_desc(p.first, d))
				return p.second;
		dms_point* pt = nullptr;
		if(dms_point_create(model.Context(), &d, &pt) != DMS_OK)
			return nullptr;
		points.push_back({d, pt});
		return pt;
	}

	/* admission: concurrent-job limit and the device-memory budget */
	bool can_admit() const
	{
		return (int) running.size() < max_concurrent && (running.size() + 1) * per_job_budget <= memory_budget;
	}

	void start(Pending& p)
	{
		Request_Result res;
		res.id = p.id;
		res.arrival_sec = p.arrival;
		res.start_sec = now();
		dms_point* pt = point_for(p.req);
		if(pt == nullptr)
		{
			res.status = DMS_ERR_ARG;
			res.error = dms_last_error(model.Context());
			res.finish_sec = now();
			done[p.id] = res;
			return;
		}
		dms_job_desc j{};
		j.mode = p.req.mode;
		j.precision = p.req.precision;
		j.capture_target = p.req.sample_size;
		j.max_trajectories = p.req.max_trajectories;
		j.seed = p.req.seed;
		j.point_id = p.req.point_id;
		j.first_trajectory_id = 0;
		j.max_scatterings = p.req.max_scatterings;
		j.max_free_steps = 1000000000000ull;
		j.rk_attempt_budget = p.req.rk_attempt_budget;
		j.r_max_rsun = 2.0;
		j.n_bins = p.req.n_bins;
		j.halo = p.req.halo.Descriptor();
		dms_job* job = nullptr;
		if(dms_job_submit(pt, &j, &job) != DMS_OK)
		{
			res.status = DMS_ERR_ARG;
			res.error = dms_last_error(model.Context());
			res.finish_sec = now();
			done[p.id] = res;
			return;
		}
		running.push_back({p.id, p.req, p.arrival, res.start_sec, job});
	}

	void collect(Running& r)
	{
		Request_Result res;
		res.id = r.id;
		res.arrival_sec = r.arrival;
		res.start_sec = r.start;
		dms_job_status st;
		res.status = dms_job_wait(r.job, &st);
		res.finish_sec = now();
		if(res.status == DMS_OK)
		{
			const size_t nb = (size_t) r.req.n_bins;
			size_t n_ev = 0;
			dms_job_read(r.job, nullptr, nullptr, nullptr, 0, nullptr, 0, &n_ev, nullptr, 0, nullptr);
			res.hist.resize(8 * nb);
			res.counts.resize(2 * nb);
			res.events.resize(n_ev);
			res.status = dms_job_read(r.job, &res.counters, res.hist.data(), res.counts.data(), nb, res.events.data(), n_ev, &n_ev,
			                          nullptr, 0, nullptr);
			res.events.resize(n_ev);
		}
		if(res.status != DMS_OK)
			res.error = dms_last_error(model.Context());
		dms_job_destroy(r.job);
		done[r.id] = std::move(res);
	}

	void run()
	{
		std::unique_lock<std::mutex> lock(mu);
		for(;;)
		{
			/* retire finished jobs */
			for(size_t k = 0; k < running.size();)
			{
				dms_job_status st;
				if(dms_job_poll(running[k].job, &st) == DMS_OK && st.done)
				{
					collect(running[k]);
					running.erase(running.begin() + (long) k);
					cv.notify_all();
				}
				else
					k++;
			}
			/* admit queued requests in arrival order */
			while(!queue.empty() && can_admit())
			{
				Pending p = std::move(queue.front());
				queue.pop_front();
				start(p);
			}
			if(queue.empty() && running.empty())
			{
				cv.notify_all(); /* Wait_All */
				if(stop)
					return;
			}
			cv.wait_for(lock, std::chrono::milliseconds(5));
		}
	}
};

Server::Server(Solar_Model& model, size_t memory_budget, size_t per_job_budget, int max_concurrent)
    : impl(new Impl(model, memory_budget, per_job_budget, max_concurrent))
{
}

Server::~Server() = default;

uint64_t Server::Submit(const Request& r)
{
	const double arrival = impl->now(); /* before the lock: admission may be building a rate table */
	std::lock_guard<std::mutex> lock(impl->mu);
	const uint64_t id = impl->next_id++;
	impl->queue.push_back({id, r, arrival});
	impl->cv.notify_all();
	return id;
}

void Server::Wait_All()
{
	std::unique_lock<std::mutex> lock(impl->mu);
	impl->cv.wait(lock, [this] { return impl->queue.empty() && impl->running.empty(); });
}

std::vector<Request_Result> Server::Results() const
{
	std::lock_guard<std::mutex> lock(impl->mu);
	std::vector<Request_Result> out;
	for(const auto& kv : impl->done)
		out.push_back(kv.second);
	return out;
}

double Server::Elapsed() const { return impl->now(); }

} // namespace dmsun
