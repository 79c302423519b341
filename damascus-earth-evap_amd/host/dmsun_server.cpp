/*
 * Serving loop: parameter-point requests admitted into one resident trajectory
 * launch per launch configuration (dms_server), where each request's
 * trajectories join the running launch as slots; the request queue admits by
 * the concurrent-request limit and the device memory of the resident points'
 * rate tables (unused points are evicted, least recently used first).
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
#include <tuple>

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

/* device bytes of a point: the FP64 and FP32 rate tables and the cross sections */
size_t point_bytes(const dms_point_desc& d)
{
	const size_t nr = d.interpolation_points >= 2 ? (size_t) d.interpolation_points : 0;
	const size_t nv = nr == 0 ? 0 : (d.speed_points >= 2 ? (size_t) d.speed_points : (d.speed_points == 0 ? nr : 0));
	return nr * nv * (sizeof(double) + sizeof(float)) + 64 * sizeof(double);
}

/* requests share a resident launch when these fields agree */
using Launch_Key = std::tuple<int, int, uint64_t, uint64_t, uint64_t, double, int>;
Launch_Key launch_key(const dms_job_desc& j)
{
	return Launch_Key(j.mode, j.precision, j.max_scatterings, j.max_free_steps, j.rk_attempt_budget, j.r_max_rsun, j.n_bins);
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
	struct Resident
	{
		dms_point_desc desc;
		dms_point* pt;
		size_t bytes;
		int users;
		uint64_t last_use;
	};
	struct Running
	{
		uint64_t id;
		Request req;
		double arrival, start;
		dms_job* job;
		dms_point* pt;
	};
	mutable std::mutex mu;
	std::condition_variable cv;
	std::deque<Pending> queue;
	std::vector<Running> running;
	std::map<uint64_t, Request_Result> done;
	std::vector<Resident> points; /* resident parameter points */
	size_t resident_bytes = 0;
	uint64_t use_clock = 0;
	std::map<Launch_Key, dms_server*> servers;
	uint64_t next_id = 0;
	bool stop = false;
	std::thread loop;

	Impl(Solar_Model& m, size_t mb, size_t pj, int mc) : model(m), memory_budget(mb), per_job_budget(pj), max_concurrent(mc)
	{
		if(max_concurrent < 1 || max_concurrent > 65536)
			throw std::invalid_argument("Server: max_concurrent must be in [1, 65536]");
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
		for(auto& kv : servers)
			dms_server_destroy(kv.second);
		for(auto& p : points)
			dms_point_destroy(p.pt);
	}
	double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

	static dms_job_desc job_desc(const Request& r)
	{
		dms_job_desc j{};
		j.mode = r.mode;
		j.precision = r.precision;
		j.capture_target = r.sample_size;
		j.max_trajectories = r.max_trajectories;
		j.seed = r.seed;
		j.point_id = r.point_id;
		j.first_trajectory_id = 0;
		j.max_scatterings = r.max_scatterings;
		j.max_free_steps = 1000000000000ull;
		j.rk_attempt_budget = r.rk_attempt_budget;
		j.r_max_rsun = 2.0;
		j.n_bins = r.n_bins;
		j.halo = r.halo.Descriptor();
		return j;
	}

	Resident* find_point(const dms_point_desc& d)
	{
		for(auto& p : points)
			if(same_desc(p.desc, d))
				return &p;
		return nullptr;
	}

	/* room for a new point's tables: evict unused points, least recently
	 * used first */
	bool make_room(size_t bytes)
	{
		while(resident_bytes + bytes > memory_budget)
		{
			size_t victim = points.size();
			for(size_t k = 0; k < points.size(); k++)
				if(points[k].users == 0 && (victim == points.size() || points[k].last_use < points[victim].last_use))
					victim = k;
			if(victim == points.size())
				return false;
			dms_point_destroy(points[victim].pt);
			resident_bytes -= points[victim].bytes;
			points.erase(points.begin() + (long) victim);
		}
		return true;
	}

	/* admission: the concurrent-request limit and the device-memory budget of
	 * the resident points */
	bool can_admit(const Request& r)
	{
		if((int) running.size() >= max_concurrent)
			return false;
		const dms_point_desc d = r.DM.Descriptor((int) r.interpolation_points);
		if(find_point(d) != nullptr)
			return true;
		const size_t bytes = point_bytes(d);
		/* a point larger than the whole budget runs once nothing else does */
		return make_room(bytes) || running.empty();
	}

	dms_server* server_for(const dms_job_desc& j)
	{
		const Launch_Key key = launch_key(j);
		auto it = servers.find(key);
		if(it != servers.end())
			return it->second;
		dms_server_desc sd{};
		sd.launch = j;
		sd.max_requests = (uint32_t) max_concurrent;
		dms_server* srv = nullptr;
		if(dms_server_create(model.Context(), &sd, &srv) != DMS_OK)
			return nullptr;
		servers[key] = srv;
		return srv;
	}

	void fail_request(const Pending& p, double start)
	{
		Request_Result res;
		res.id = p.id;
		res.arrival_sec = p.arrival;
		res.start_sec = start;
		res.status = DMS_ERR_ARG;
		res.error = dms_last_error(model.Context());
		res.finish_sec = now();
		done[p.id] = res;
	}

	void start(Pending& p)
	{
		const double start_sec = now();
		const dms_point_desc d = p.req.DM.Descriptor((int) p.req.interpolation_points);
		Resident* rp = find_point(d);
		if(rp == nullptr)
		{
			dms_point* pt = nullptr;
			if(dms_point_create(model.Context(), &d, &pt) != DMS_OK)
			{
				fail_request(p, start_sec);
				return;
			}
			points.push_back({d, pt, point_bytes(d), 0, 0});
			resident_bytes += point_bytes(d);
			rp = &points.back();
		}
		const dms_job_desc j = job_desc(p.req);
		dms_server* srv = server_for(j);
		dms_job* job = nullptr;
		if(srv == nullptr || dms_server_submit(srv, rp->pt, &j, &job) != DMS_OK)
		{
			fail_request(p, start_sec);
			return;
		}
		rp->users++;
		rp->last_use = ++use_clock;
		running.push_back({p.id, p.req, p.arrival, start_sec, job, rp->pt});
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
		for(auto& p : points)
			if(p.pt == r.pt)
				p.users--;
		done[r.id] = std::move(res);
	}

	void run()
	{
		std::unique_lock<std::mutex> lock(mu);
		for(;;)
		{
			/* retire finished requests */
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
			while(!queue.empty() && can_admit(queue.front().req))
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
			cv.wait_for(lock, std::chrono::milliseconds(1));
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

dms_server_stats Server::Stats() const
{
	std::lock_guard<std::mutex> lock(impl->mu);
	dms_server_stats total{};
	for(const auto& kv : impl->servers)
	{
		dms_server_stats s{};
		if(dms_server_get_stats(kv.second, &s) != DMS_OK)
			continue;
		total.launches += s.launches;
		total.slots_published += s.slots_published;
		total.slots_cancelled += s.slots_cancelled;
		total.requests_done += s.requests_done;
		total.requests_active += s.requests_active;
		total.running += s.running;
	}
	return total;
}

} // namespace dmsun
