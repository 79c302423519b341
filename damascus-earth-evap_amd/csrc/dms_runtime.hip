/*
 * libdmsun: C ABI (include/dms.h) over the gfx950 trajectory kernels.
 *
 * Job flow (Simulation_Data::Generate_Data, src/Data_Generation.cpp:488-972):
 * trajectories are simulated speculatively in large id-ordered chunks
 * (initial conditions on the device, then the persistent trajectory kernel),
 * the host folds the per-trajectory summaries into the run totals strictly in
 * id order with the reference's round logic (sync interval, capture target,
 * run-level fuses), and a deterministic reduction kernel sums the per-bin rows
 * of the trajectories that count.  Results therefore do not depend on launch
 * geometry or chunking.
 */
#include <algorithm>
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "dms_host.hpp"
#include "dms_kernels.hip"

/* jobs (or batches, or a resident serving launch) that may run on the device
 * at the same time: with the context stream and the shared progress stream,
 * the process uses four streams, one per hardware queue of the device */
#define DMS_CONCURRENT_JOBS 2
/* cancel words of the context: one per running job or batch launch */
#define DMS_CANCEL_WORDS 64

struct dms_ctx
{
	int device = -1;
	hipStream_t stream = nullptr;
	hipStream_t progress_stream = nullptr; /* progress reads while launches run */
	/* work streams of concurrently running jobs */
	std::vector<hipStream_t> pool;
	std::vector<int> pool_free;
	std::mutex pool_mu;
	std::condition_variable pool_cv;
	/* device memory a chunked job may use for its trajectory buffers */
	size_t job_memory_budget = (size_t) 32 << 30;
	std::mutex mu;
	std::string err;
	dms::SolarTables T;
	int n_cu = 0;
	int traj_blocks_per_cu = 0;
	size_t traj_lds = 0;
	/* device copies of the solar tables */
	double* d_x = nullptr;
	DmsCubic *d_mass = nullptr, *d_temp = nullptr, *d_vesc2 = nullptr, *d_dens = nullptr;
	double *d_pref = nullptr, *d_imass = nullptr;
	int* d_Z = nullptr;
	DmsSolarView dev;
	/* device address of this GPU's copy of the allocation-free entry point's
	 * scratch array (hipGetSymbolAddress at set-up) */
	unsigned char* scratch_ptr = nullptr;
	/* cancel words of running launches: host memory mapped for the device,
	 * allocated once (freeing pinned memory would wait for the device) */
	unsigned int* h_cancel = nullptr;
	unsigned int* d_cancel = nullptr;
	std::vector<int> cancel_free;
	/* launch knobs (dms_set_launch_knobs, chosen by dms_autotune_knobs) */
	dms_launch_knobs knobs = {1u, 0u};
};

struct dms_point
{
	dms_ctx* ctx = nullptr;
	dms_point_desc desc;
	dms::PointParams P;
	double* d_sigmaN = nullptr;
	double* d_grid = nullptr;
	float* d_grid32 = nullptr;
	int Nr = 0, Nv = 0; /* rate grid radii x speeds; 0 x 0 without a grid */
	DmsPointView dev;
};

struct dms_job;

struct dms_batch
{
	std::vector<dms_job*> jobs;
	std::thread worker;
	std::mutex mu;
	bool joined = false;
	/* dms_job_cancel: the request, and the launch's cancel word once it has one */
	std::atomic<int> cancelled{0};
	std::atomic<int> cancel_word{-1};
};

struct dms_job
{
	dms_point* pt = nullptr;
	dms_job_desc desc;
	std::shared_ptr<dms_batch> batch;
	bool needs_rerun = false;
	double batch_size = 1.0;
	std::atomic<uint64_t> done_traj{0}, done_captured{0}, done_scat{0}, done_escaped{0}, done_failed{0};
	std::atomic<int> finished{0};
	/* jobs of a dms_server: completion is signalled here (the server may be
	 * gone by the time the caller waits) */
	struct dms_server* server = nullptr;
	std::mutex fmu;
	std::condition_variable fcv;
	std::atomic<int> cancel_requested{0};
	int status = DMS_OK;
	std::string err;
	dms_counters C;
	std::vector<double> hist;
	std::vector<unsigned long long> counts;
	std::vector<dms_evap_event> events;
	std::vector<dms_traj_record> records;
	double kernel_ms = 0.0, other_ms = 0.0, launches = 0.0;
};

namespace
{

thread_local std::string g_last_error;

int fail(dms_ctx* ctx, int code, const std::string& msg)
{
	if(ctx)
		ctx->err = msg;
	g_last_error = msg;
	return code;
}

#define DMS_HIP(ctx, call)                                                                                         \
	do                                                                                                           \
	{                                                                                                            \
		hipError_t e_ = (call);                                                                                  \
		if(e_ != hipSuccess)                                                                                     \
			return fail((ctx), DMS_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));               \
	} while(0)

template <typename T>
hipError_t dev_alloc_copy(T** d, const T* h, size_t n)
{
	hipError_t e = hipMalloc((void**) d, sizeof(T) * (n ? n : 1));
	if(e != hipSuccess)
		return e;
	if(n)
		e = hipMemcpy(*d, h, sizeof(T) * n, hipMemcpyHostToDevice);
	return e;
}

DmsPointView host_point_view(const dms_point* pt)
{
	DmsPointView v;
	v.mass = pt->P.mass;
	v.mu_e = pt->P.mu_e;
	v.interaction = pt->P.interaction;
	v.ff = pt->P.ff;
	v.m_dark_photon = pt->P.m_dark_photon;
	v.q_reference = pt->P.q_reference;
	v.sigma_e = pt->P.sigma_e;
	v.sigma_N = pt->P.sigma_N;
	const bool grid = pt->Nr >= 2 && pt->Nv >= 2;
	v.use_grid = grid;
	v.Nr = pt->Nr;
	v.Nv = pt->Nv;
	v.inv_dr = grid ? (double) (pt->Nr - 1) / DMS_rSun : 0.0;
	v.inv_dv = grid ? (double) (pt->Nv - 1) / DMS_V_MAX : 0.0;
	v.vmax = grid ? DMS_V_MAX : 0.0;
	v.grid = nullptr;
	v.grid32 = nullptr;
	return v;
}

DmsJobView job_view(const dms_job_desc& d)
{
	DmsJobView J;
	J.max_time_steps = d.max_free_steps;
	J.max_scatterings = d.max_scatterings;
	J.rk_budget = d.rk_attempt_budget;
	J.max_distance = d.r_max_rsun * DMS_rSun;
	J.terminate_on_capture = (d.mode == DMS_MODE_CAPTURE);
	J.nbins = d.n_bins > 0 ? d.n_bins : 2000;
	J.bin_width_km = (2.0 * DMS_R_SUN_KM) / J.nbins;
	J.seed = d.seed;
	J.point_id = d.point_id;
	J.first_id = d.first_trajectory_id;
	J.mixed = (d.precision == DMS_MIXED) ? 1 : 0;
	return J;
}

std::string check_job(const dms_job_desc& d)
{
	if(d.mode != DMS_MODE_CAPTURE && d.mode != DMS_MODE_PARAMETER_POINT)
		return "mode must be capture or parameter point";
	if(d.precision != DMS_FP64 && d.precision != DMS_MIXED)
		return "precision must be fp64 or mixed";
	if(!(d.r_max_rsun >= 1.0) || !std::isfinite(d.r_max_rsun))
		return "r_max_rsun must be finite and >= 1 (maximum distance of Trajectory_Simulator)";
	if(d.n_bins < 0 || d.n_bins > (1 << 20))
		return "n_bins out of range";
	return "";
}

DmsHaloView halo_view(const dms_ctx* ctx, const dms_halo& h)
{
	DmsHaloView H;
	dms_halo_init(H, h);
	const DmsSolarView V = ctx->T.view();
	H.vescR = dms_local_escape_speed(V, DMS_rSun);
	H.vesc_asym = dms_local_escape_speed(V, 1000.0 * DMS_AU);
	H.norm = dms_shm_moment(H, 1) + H.vescR * H.vescR * dms_shm_moment(H, -1);
	return H;
}

std::string check_halo(const dms_halo& h)
{
	if(!(h.v0_kms > 0.0) || !(h.vesc_kms > 0.0) || !std::isfinite(h.v0_kms) || !std::isfinite(h.vesc_kms))
		return "halo v0 and v_esc must be finite and positive";
	for(int c = 0; c < 3; c++)
		if(!std::isfinite(h.vobs_kms[c]))
			return "halo observer velocity must be finite";
	const double vo = std::sqrt(h.vobs_kms[0] * h.vobs_kms[0] + h.vobs_kms[1] * h.vobs_kms[1] + h.vobs_kms[2] * h.vobs_kms[2]);
	if(!(vo < h.vesc_kms))
		return "observer speed must be below the galactic escape speed";
	return "";
}

/* Stream-ordered device allocations of one job (hipMallocAsync/hipFreeAsync on
 * the job's stream), so that finishing a job never waits for other jobs'
 * kernels. */
template <typename T>
hipError_t dalloc(T** p, size_t count, hipStream_t st)
{
	*p = nullptr;
	return hipMallocAsync((void**) p, sizeof(T) * (count ? count : 1), st);
}
template <typename T>
void dfree(T*& p, hipStream_t st)
{
	if(p)
		hipFreeAsync(p, st);
	p = nullptr;
}

/* Device buffers of one chunk, grown on demand. */
struct ChunkBuffers
{
	dms_event* ics = nullptr;
	int32_t* ok = nullptr;
	DmsSummary* summ = nullptr;
	dms_traj_record* recs = nullptr;
	double* rows = nullptr;
	uint32_t* counts = nullptr;
	uint8_t* cls = nullptr;
	unsigned int* next = nullptr;
	size_t cap = 0, rows_cap = 0, recs_cap = 0;

	void release(hipStream_t st)
	{
		dfree(ics, st);
		dfree(ok, st);
		dfree(summ, st);
		dfree(recs, st);
		dfree(rows, st);
		dfree(counts, st);
		dfree(cls, st);
		dfree(next, st);
		*this = ChunkBuffers();
	}
};

hipError_t ensure(ChunkBuffers& B, size_t n, int nbins, bool rows, bool recs, hipStream_t st)
{
	hipError_t e = hipSuccess;
	if(B.next == nullptr && (e = dalloc(&B.next, 4, st)) != hipSuccess) /* slot counter + fill control */
		return e;
	if(n > B.cap)
	{
		dfree(B.ics, st);
		dfree(B.ok, st);
		dfree(B.summ, st);
		dfree(B.cls, st);
		if((e = dalloc(&B.ics, n, st)) != hipSuccess || (e = dalloc(&B.ok, n, st)) != hipSuccess
		   || (e = dalloc(&B.summ, n, st)) != hipSuccess || (e = dalloc(&B.cls, n, st)) != hipSuccess)
			return e;
		B.cap = n;
	}
	if(rows && n > B.rows_cap)
	{
		dfree(B.rows, st);
		dfree(B.counts, st);
		if((e = dalloc(&B.rows, 2 * (size_t) nbins * n, st)) != hipSuccess
		   || (e = dalloc(&B.counts, (size_t) nbins * n, st)) != hipSuccess)
			return e;
		B.rows_cap = n;
	}
	if(recs && n > B.recs_cap)
	{
		dfree(B.recs, st);
		if((e = dalloc(&B.recs, n, st)) != hipSuccess)
			return e;
		B.recs_cap = n;
	}
	return hipSuccess;
}

int traj_grid(const dms_ctx* ctx, size_t n)
{
	const size_t need = (n + DMS_TRAJ_BLOCK - 1) / DMS_TRAJ_BLOCK;
	const size_t resident = (size_t) ctx->n_cu * (size_t) (ctx->traj_blocks_per_cu > 0 ? ctx->traj_blocks_per_cu : 1);
	const size_t g = need < resident ? need : resident;
	return (int) (g > 0 ? g : 1);
}

/* Jobs sharing one trajectory launch (slot ranges and first trajectory ids). */
struct LaunchJobs
{
	std::vector<uint32_t> start;
	std::vector<uint64_t> first_id;
	uint32_t* d_start = nullptr;
	uint64_t* d_first = nullptr;
	unsigned int* d_done = nullptr;
	/* multi-point launch */
	std::vector<DmsPointView> points;
	std::vector<uint32_t> keys;
	DmsPointView* d_points = nullptr;
	uint32_t* d_keys = nullptr;

	/* the host vectors must stay unchanged until the stream has run the copies */
	hipError_t upload(hipStream_t st)
	{
		hipError_t e;
		if((e = dalloc(&d_start, start.size(), st)) != hipSuccess || (e = dalloc(&d_first, first_id.size(), st)) != hipSuccess
		   || (e = dalloc(&d_done, DMS_PROGRESS_WORDS * start.size(), st)) != hipSuccess)
			return e;
		if((e = hipMemsetAsync(d_done, 0, sizeof(unsigned int) * DMS_PROGRESS_WORDS * start.size(), st)) != hipSuccess)
			return e;
		if(!points.empty())
		{
			if((e = dalloc(&d_points, points.size(), st)) != hipSuccess || (e = dalloc(&d_keys, keys.size(), st)) != hipSuccess)
				return e;
			if((e = hipMemcpyAsync(d_points, points.data(), sizeof(DmsPointView) * points.size(), hipMemcpyHostToDevice, st))
			       != hipSuccess
			   || (e = hipMemcpyAsync(d_keys, keys.data(), sizeof(uint32_t) * keys.size(), hipMemcpyHostToDevice, st)) != hipSuccess)
				return e;
		}
		if((e = hipMemcpyAsync(d_start, start.data(), sizeof(uint32_t) * start.size(), hipMemcpyHostToDevice, st)) != hipSuccess)
			return e;
		return hipMemcpyAsync(d_first, first_id.data(), sizeof(uint64_t) * first_id.size(), hipMemcpyHostToDevice, st);
	}
	void release(hipStream_t st)
	{
		dfree(d_start, st);
		dfree(d_first, st);
		dfree(d_done, st);
		dfree(d_points, st);
		dfree(d_keys, st);
	}
};

/* Launch the trajectory kernel over n slots.  rows: per-trajectory rows in B;
 * lane_rows: per-lane scratch mode (R prepared by the caller). */
/* Called about every 20 ms while a launch runs with the finished-trajectory
 * count of each job of the launch. */
using ProgressFn = std::function<void(const std::vector<unsigned int>&)>;

/* Capture-target chunk: stop handing out slots fill_extra slots after the
 * launch's cap_stop-th captured trajectory.  Slots never started keep an
 * all-ones summary (reason -1); the host counts only the started prefix. */
struct FillStop
{
	uint32_t cap_stop, fill_extra;
};

hipError_t run_trajectories(dms_ctx* ctx, hipStream_t st, hipStream_t pst, const dms_point* pt, const DmsJobView& J, ChunkBuffers& B,
                            size_t n, bool rows, bool recs, const LaunchJobs& LJ, const DmsRows* lane_rows, float* ms,
                            const ProgressFn& progress = nullptr, const FillStop* stop = nullptr,
                            const unsigned int* cancel = nullptr)
{
	DmsTrajArgs a;
	a.S = ctx->dev;
	a.P = pt->dev;
	a.J = J;
	a.ics = B.ics;
	a.shift_ok = B.ok;
	a.n = (uint32_t) n;
	a.records = recs ? B.recs : nullptr;
	a.summ = B.summ;
	if(lane_rows)
		a.R = *lane_rows;
	else
	{
		a.R.rows = rows ? B.rows : nullptr;
		a.R.counts = rows ? B.counts : nullptr;
		a.R.bins = nullptr;
		a.R.touched = nullptr;
		a.R.per_lane = 0;
		a.R.hist = nullptr;
		a.R.hcounts = nullptr;
	}
	a.next = B.next;
	a.refill_batch = ctx->knobs.refill_batch;
	a.lone_serial = ctx->knobs.lone_serial;
	a.n_jobs = (uint32_t) LJ.start.size();
	a.job_start = LJ.d_start;
	a.job_first_id = LJ.d_first;
	a.job_done = LJ.d_done;
	a.job_points = LJ.d_points;
	a.job_keys = LJ.d_keys;
	a.fill_ctl = nullptr;
	a.cap_stop = 0;
	a.fill_extra = 0;
	a.cancel = cancel;
	hipError_t e;
	if((e = hipMemsetAsync(B.next, 0, sizeof(unsigned int), st)) != hipSuccess)
		return e;
	if(stop)
	{
		a.fill_ctl = B.next + 1;
		a.cap_stop = stop->cap_stop;
		a.fill_extra = stop->fill_extra;
		if((e = hipMemsetD32Async((hipDeviceptr_t) (B.next + 1), 0, 1, st)) != hipSuccess
		   || (e = hipMemsetD32Async((hipDeviceptr_t) (B.next + 2), (int) n, 1, st)) != hipSuccess
		   || (e = hipMemsetAsync(B.summ, 0xFF, sizeof(DmsSummary) * n, st)) != hipSuccess)
			return e;
	}
	if(rows && !lane_rows)
	{
		if((e = hipMemsetAsync(B.rows, 0, sizeof(double) * 2 * (size_t) J.nbins * n, st)) != hipSuccess)
			return e;
		if((e = hipMemsetAsync(B.counts, 0, sizeof(uint32_t) * (size_t) J.nbins * n, st)) != hipSuccess)
			return e;
	}
	hipEvent_t e0, e1;
	hipEventCreate(&e0);
	hipEventCreate(&e1);
	(void) hipGetLastError();
	hipEventRecord(e0, st);
	void (*kernel)(DmsTrajArgs) = LJ.d_points ? (J.mixed ? dms_traj_kernel_mixed_multi : dms_traj_kernel_multi)
	                                          : (J.mixed ? dms_traj_kernel_mixed : dms_traj_kernel);
	hipLaunchKernelGGL(kernel, dim3(traj_grid(ctx, n)), dim3(DMS_TRAJ_BLOCK), ctx->traj_lds, st, a);
	e = hipGetLastError();
	hipEventRecord(e1, st);
	if(e == hipSuccess && progress && LJ.d_done)
	{
		/* heartbeat: read the per-job finished counters on a second stream */
		std::vector<unsigned int> done(DMS_PROGRESS_WORDS * LJ.start.size(), 0u);
		while(hipEventQuery(e1) == hipErrorNotReady)
		{
			if(hipMemcpyAsync(done.data(), LJ.d_done, sizeof(unsigned int) * done.size(), hipMemcpyDeviceToHost, pst) == hipSuccess
			   && hipStreamSynchronize(pst) == hipSuccess)
				progress(done);
			std::this_thread::sleep_for(std::chrono::milliseconds(20));
		}
	}
	if(e == hipSuccess)
		e = hipEventSynchronize(e1);
	float t = 0.0f;
	hipEventElapsedTime(&t, e0, e1);
	if(ms)
		*ms = t;
	hipEventDestroy(e0);
	hipEventDestroy(e1);
	return e;
}

/* single-job launch over slots [0, n) with ids first_id .. */
hipError_t run_trajectories_single(dms_ctx* ctx, hipStream_t st, hipStream_t pst, const dms_point* pt, const DmsJobView& J,
                                   ChunkBuffers& B, size_t n, bool rows, bool recs, uint64_t first_id, float* ms,
                                   const ProgressFn& progress = nullptr, const FillStop* stop = nullptr,
                                   const unsigned int* cancel = nullptr)
{
	LaunchJobs LJ;
	LJ.start.push_back(0);
	LJ.first_id.push_back(first_id);
	hipError_t e = LJ.upload(st);
	if(e == hipSuccess)
		e = run_trajectories(ctx, st, pst, pt, J, B, n, rows, recs, LJ, nullptr, ms, progress, stop, cancel);
	LJ.release(st);
	return e;
}

double now_ms()
{
	return 1e-6 * (double) std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool is_numerical(int r)
{
	return r == DMS_NUMERICAL_FAILURE || r == DMS_NON_FINITE_STATE || r == DMS_SPEED_LIMIT || r == DMS_ENERGY_DRIFT_ESCAPE || r == DMS_UNKNOWN;
}
bool is_truncation(int r) { return r == DMS_WALL_TIME_LIMIT || r == DMS_MAX_FREE_STEPS || r == DMS_MAX_SCATTERINGS; }

/* Round logic of Generate_Data for one rank (src/Data_Generation.cpp:696-938). */
struct RoundReplay
{
	uint64_t target, max_traj, interval;
	bool capture, unlimited;
	uint64_t local_total = 0, captured = 0, left_in_round = 0;
	bool stopped = false;
	int stop = 0;
	dms_counters* C;

	void init(const dms_job_desc& d, const dms_point* pt, dms_counters* counters)
	{
		C = counters;
		capture = (d.mode == DMS_MODE_CAPTURE);
		target = d.capture_target == 0 ? UINT64_MAX : d.capture_target;
		unlimited = d.max_trajectories == 0;
		max_traj = unlimited ? UINT64_MAX : d.max_trajectories;
		interval = capture ? 32 : dms::normal_mode_sync_interval(pt->P.sigma_p / (DMS_cm * DMS_cm));
		C->sync_interval = (int32_t) interval;
	}

	bool open_round()
	{
		if(stopped)
			return false;
		if(left_in_round > 0)
			return true;
		if(!(captured < target))
		{
			stopped = true;
			return false;
		}
		const uint64_t remaining = target - captured;
		uint64_t capacity = 0;
		if(local_total < max_traj)
			capacity = (max_traj - local_total < interval) ? max_traj - local_total : interval;
		if(capacity == 0)
		{
			stop = 1;
			stopped = true;
			return false;
		}
		const uint64_t batch = capture ? (remaining < capacity ? remaining : capacity) : capacity;
		C->sync_rounds++;
		C->final_round_trajectories = batch;
		left_in_round = batch;
		return true;
	}

	void close_one()
	{
		left_in_round--;
		if(left_in_round == 0 && C->trajectories > 0)
		{
			const double shift_fraction = (double) C->initial_shift_failures / (double) C->trajectories;
			if(shift_fraction > 1.0e-2)
			{
				stop = 3;
				stopped = true;
				return;
			}
			const double invalid_fraction = (double) (C->numerical_failures + C->computational_truncations) / (double) C->trajectories;
			if(unlimited && invalid_fraction > 1.0e-2)
			{
				stop = 4;
				stopped = true;
			}
		}
	}

	void finish()
	{
		C->capture_target_overshoot = (captured > target) ? captured - target : 0;
		int s = stop;
		if(captured < target && s == 0)
			s = 2;
		C->early_stop_reason = s;
	}
};

/* simulate_one_trajectory accounting (src/Data_Generation.cpp:764-850) of one
 * summary.  Returns the histogram class: 1 captured, 2 completed escape, 0 none. */
int account_trajectory(dms_job* job, RoundReplay& R, const DmsSummary& s, double elapsed_sec)
{
	dms_counters& C = job->C;
	const bool capture = R.capture;
	int cls = 0;
	const bool shift_ok = !(s.flags & DMS_F_SHIFT_FAILED);
	if(!shift_ok)
	{
		C.initial_shift_failures++;
		C.numerical_failures++;
	}
	R.local_total++;
	C.trajectories++;
	C.total_scatterings += s.n_scat;
	C.reason_counts[s.reason]++;
	C.n_rk_attempts += s.n_rk;
	C.n_steps_interior += s.n_int;
	C.n_steps_exterior += s.n_ext;
	C.n_tau_rejections += s.n_rej;
	C.n_tv_iterations += s.n_tv;
	if(s.n_rk > C.max_rk_attempts)
		C.max_rk_attempts = s.n_rk;
	if(shift_ok && is_numerical(s.reason))
		C.numerical_failures++;
	if(is_truncation(s.reason))
		C.computational_truncations++;
	if(s.flags & DMS_F_CAPTURED)
	{
		C.captured++;
		R.captured++;
		if(!capture)
		{
			if(s.flags & DMS_F_EVENT_OBSERVED)
				C.complete_evaporation++;
			else
				C.invalid_survival_captured++;
			cls = 1;
			/* Build_Evaporation_Record + Make_Compact_Evaporation_Event (:118-176) */
			if(std::isfinite(s.t_capture_sec))
			{
				const bool nie = (s.flags & DMS_F_NUMERICALLY_INVALID_ESCAPE) || s.reason == DMS_ENERGY_DRIFT_ESCAPE;
				const bool sv = (s.flags & DMS_F_SURVIVAL_VALID) && !nie;
				const bool eo = sv && (s.flags & DMS_F_EVENT_OBSERVED);
				const double lu = eo ? (s.t_unbind_sec - s.t_capture_sec) : -1.0;
				const double lifetime = (std::isfinite(lu) && lu >= 0.0) ? lu : -1.0;
				if(sv && eo && lifetime >= 0.0)
				{
					dms_evap_event E;
					E.rank = 0;
					E.pad = 0;
					E.trajectory_id = C.trajectories;
					E.completion_wall_time_sec = elapsed_sec;
					E.lifetime_unbinding_sec = lifetime;
					job->events.push_back(E);
				}
			}
		}
	}
	else if(s.reason == DMS_OUTWARD_ESCAPE)
	{
		C.completed_outward_escapes++;
		if(!capture)
		{
			cls = 2;
			if(s.flags & DMS_F_FREE)
				C.free_particles++;
			else if(s.flags & DMS_F_REFLECTED)
			{
				C.reflected_particles++;
				if(s.flags & DMS_F_REFLECTION_SHIFT_FAILED)
				{
					C.final_reflection_shift_failures++;
					C.numerical_failures++;
				}
			}
		}
	}
	return cls;
}

/* Chunked job: speculative id-ordered chunks with per-trajectory rows, so that
 * capture-target runs count exactly the trajectories the reference would. */
void run_job_chunked(dms_job* job, hipStream_t st, hipStream_t pst, const unsigned int* cancel = nullptr,
                     const std::atomic<int>* cancelled = nullptr)
{
	dms_point* pt = job->pt;
	dms_ctx* ctx = pt->ctx;
	const dms_job_desc& d = job->desc;
	const DmsJobView J = job_view(d);
	const bool capture = (d.mode == DMS_MODE_CAPTURE);
	const bool want_rows = !capture;
	const bool want_recs = d.record_trajectories != 0;
	const int nb = J.nbins;
	dms_counters& C = job->C;
	memset(&C, 0, sizeof(C));
	job->events.clear();
	job->records.clear();
	job->hist.assign((size_t) 8 * nb, 0.0);
	job->counts.assign((size_t) 2 * nb, 0ull);
	RoundReplay R;
	R.init(d, pt, &C);
	const DmsHaloView H = halo_view(ctx, d.halo);

	ChunkBuffers B;
	double* d_hist = nullptr;
	unsigned long long* d_counts = nullptr;
	double* d_partial = nullptr;
	unsigned long long* d_pcount = nullptr;
	size_t partial_cap = 0;
	const uint32_t TB = 512;
	auto set_err = [&](const std::string& m) {
		job->status = DMS_ERR_DEVICE;
		job->err = m;
	};
	if(want_rows)
	{
		if(dalloc(&d_hist, 8 * (size_t) nb, st) != hipSuccess || dalloc(&d_counts, 2 * (size_t) nb, st) != hipSuccess)
		{
			set_err("device allocation of histogram accumulators failed");
			dfree(d_hist, st);
			dfree(d_counts, st);
			return;
		}
		hipMemsetAsync(d_hist, 0, sizeof(double) * 8 * nb, st);
		hipMemsetAsync(d_counts, 0, sizeof(unsigned long long) * 2 * nb, st);
	}
	/* Chunk size: bounded by a device-memory budget for the per-trajectory rows. */
	const size_t row_bytes = want_rows ? (size_t) nb * (2 * sizeof(double) + sizeof(uint32_t)) : 0;
	const size_t rec_bytes = want_recs ? sizeof(dms_traj_record) : 0;
	const size_t per_traj = row_bytes + rec_bytes + sizeof(dms_event) + sizeof(DmsSummary) + 8;
	const size_t budget = ctx->job_memory_budget;
	size_t max_chunk = budget / per_traj;
	if(max_chunk > (size_t) 1 << 26)
		max_chunk = (size_t) 1 << 26;
	uint64_t next_id = d.first_trajectory_id;
	uint64_t simulated = 0;
	const double t_start = now_ms();
	std::vector<DmsSummary> hsum;
	std::vector<uint8_t> hcls;
	std::vector<dms_traj_record> hrec;
	while(R.open_round())
	{
		if(cancelled && cancelled->load())
			break;
		uint64_t want;
		if(R.target == UINT64_MAX)
			want = R.max_traj - R.local_total;
		else
		{
			const double rate = (C.trajectories > 0 && C.captured > 0) ? (double) C.captured / (double) C.trajectories : 0.0;
			const uint64_t need = R.target - R.captured;
			/* Every launch waits for its longest trajectory, while the bulk of a
			 * few hundred thousand more trajectories costs a small fraction of
			 * that: size chunks so that one more launch is rarely needed.  The
			 * trajectories past the stop point are discarded (results do not
			 * depend on the chunking). */
			want = rate > 0.0 ? (uint64_t) (1.5 * (double) need / rate) + 65536 : 262144;
			if(!R.unlimited && want > R.max_traj - R.local_total)
				want = R.max_traj - R.local_total;
			if(want < R.left_in_round)
				want = R.left_in_round;
		}
		size_t n = (size_t) (want < max_chunk ? want : max_chunk);
		if(n == 0)
			n = 1;
		if(ensure(B, n, nb, want_rows, want_recs, st) != hipSuccess)
		{
			set_err("device allocation for a trajectory chunk failed");
			break;
		}
		const double t0 = now_ms();
		(void) hipGetLastError();
		hipLaunchKernelGGL(dms_ic_kernel, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, H, d.seed, d.point_id,
		                   next_id, (uint32_t) n, d.r_max_rsun * DMS_rSun, B.ics, B.ok);
		float kms = 0.0f;
		const uint64_t base_done = C.trajectories;
		/* capture target: the launch stops starting trajectories once it has
		 * captured a quarter more than still needed, plus enough slots to
		 * close the current sync round twice over */
		FillStop fs{0u, 0u};
		const bool fill_stop = R.target != UINT64_MAX;
		if(fill_stop)
		{
			const uint64_t need = R.target > R.captured ? R.target - R.captured : 1;
			const uint64_t cs = need + need / 4 + 32;
			const uint64_t fx = 2 * R.interval + 1024;
			fs.cap_stop = (uint32_t) (cs < 0x7fffffffULL ? cs : 0x7fffffffULL);
			fs.fill_extra = (uint32_t) (fx < (uint64_t) n ? fx : (uint64_t) n);
		}
		hipError_t e = run_trajectories_single(ctx, st, pst, pt, J, B, n, want_rows, want_recs, next_id, &kms,
		                                       [job, base_done, base_cap = C.captured, base_esc = C.completed_outward_escapes,
		                                        base_fail = C.numerical_failures](const std::vector<unsigned int>& d) {
			                                       job->done_traj = base_done + d[0];
			                                       job->done_captured = base_cap + d[1];
			                                       job->done_escaped = base_esc + d[2];
			                                       job->done_failed = base_fail + d[3];
		                                       },
		                                       fill_stop ? &fs : nullptr, cancel);
		if(e != hipSuccess)
		{
			set_err(std::string("trajectory kernel: ") + hipGetErrorString(e));
			break;
		}
		job->kernel_ms += kms;
		job->launches += 1.0;
		hsum.resize(n);
		hipMemcpyAsync(hsum.data(), B.summ, sizeof(DmsSummary) * n, hipMemcpyDeviceToHost, st);
		if(want_recs)
		{
			hrec.resize(n);
			hipMemcpyAsync(hrec.data(), B.recs, sizeof(dms_traj_record) * n, hipMemcpyDeviceToHost, st);
		}
		if(hipStreamSynchronize(st) != hipSuccess)
		{
			set_err("reading trajectory summaries failed");
			break;
		}
		/* started prefix (slots past the fill limit were never started) */
		if(fill_stop)
			for(size_t i = 0; i < n; i++)
				if(hsum[i].reason == -1)
				{
					n = i;
					break;
				}
		simulated += n;
		next_id += n;
		hcls.assign(n, 0);
		size_t counted = 0;
		const double elapsed_sec = (now_ms() - t_start) * 1e-3;
		for(size_t i = 0; i < n; i++)
		{
			if(!R.open_round())
				break;
			hcls[i] = (uint8_t) account_trajectory(job, R, hsum[i], elapsed_sec);
			if(want_recs)
				job->records.push_back(hrec[i]);
			counted = i + 1;
			R.close_one();
		}
		if(want_rows && counted > 0)
		{
			hipMemcpyAsync(B.cls, hcls.data(), counted, hipMemcpyHostToDevice, st);
			const uint32_t nblk = (uint32_t) ((counted + TB - 1) / TB);
			const size_t need = (size_t) nblk * nb;
			if(need > partial_cap)
			{
				dfree(d_partial, st);
				dfree(d_pcount, st);
				if(dalloc(&d_partial, 8 * need, st) != hipSuccess || dalloc(&d_pcount, 2 * need, st) != hipSuccess)
				{
					set_err("device allocation of histogram partials failed");
					break;
				}
				partial_cap = need;
			}
			hipLaunchKernelGGL(dms_hist_partial_kernel, dim3((unsigned) ((nb + 127) / 128), nblk), dim3(128), 0, st,
			                   B.rows, B.counts, B.cls, (uint32_t) counted, nb, TB, d_partial, d_pcount);
			hipLaunchKernelGGL(dms_hist_final_kernel, dim3((unsigned) ((nb + 127) / 128)), dim3(128), 0, st, d_partial,
			                   d_pcount, nblk, nb, d_hist, d_counts);
			if(hipStreamSynchronize(st) != hipSuccess)
			{
				set_err("histogram reduction failed");
				break;
			}
		}
		job->other_ms += (now_ms() - t0) - kms;
		job->done_traj = C.trajectories;
		job->done_captured = C.captured;
		job->done_scat = C.total_scatterings;
		job->done_escaped = C.completed_outward_escapes;
		job->done_failed = C.numerical_failures;
	}
	R.finish();
	C.simulated_trajectories = simulated;
	C.computing_time_sec = (now_ms() - t_start) * 1e-3;
	if(want_rows && job->status == DMS_OK)
	{
		hipMemcpyAsync(job->hist.data(), d_hist, sizeof(double) * 8 * nb, hipMemcpyDeviceToHost, st);
		hipMemcpyAsync(job->counts.data(), d_counts, sizeof(unsigned long long) * 2 * nb, hipMemcpyDeviceToHost, st);
	}
	dfree(d_hist, st);
	dfree(d_counts, st);
	dfree(d_partial, st);
	dfree(d_pcount, st);
	B.release(st);
	hipStreamSynchronize(st);
}

/* Fixed-count jobs (max_trajectories > 0, no capture target) of one point
 * streamed through a single persistent launch: lanes refill across job
 * boundaries, so the long tail of one job overlaps the bulk of the others.
 * Histograms use per-lane scratch rows flushed into per-job sums. */
void run_batch_fixed(std::vector<dms_job*>& jobs, hipStream_t st, hipStream_t pst, const unsigned int* cancel = nullptr,
                     const std::atomic<int>* cancelled = nullptr)
{
	dms_point* pt = jobs[0]->pt;
	dms_ctx* ctx = pt->ctx;
	const dms_job_desc& d0 = jobs[0]->desc;
	const DmsJobView J = job_view(d0);
	const bool capture = (d0.mode == DMS_MODE_CAPTURE);
	const bool want_rows = !capture;
	bool want_recs = false;
	for(dms_job* j : jobs)
		want_recs = want_recs || j->desc.record_trajectories != 0;
	const int nb = J.nbins;
	const size_t nj = jobs.size();
	LaunchJobs LJ;
	size_t n = 0;
	bool multi = false;
	for(dms_job* j : jobs)
		multi = multi || j->pt != pt || j->desc.seed != d0.seed || j->desc.point_id != d0.point_id;
	for(dms_job* j : jobs)
	{
		LJ.start.push_back((uint32_t) n);
		LJ.first_id.push_back(j->desc.first_trajectory_id);
		if(multi)
		{
			LJ.points.push_back(j->pt->dev);
			LJ.keys.push_back(j->desc.seed);
			LJ.keys.push_back(j->desc.point_id);
		}
		n += (size_t) j->desc.max_trajectories;
		memset(&j->C, 0, sizeof(j->C));
		j->events.clear();
		j->records.clear();
		j->hist.assign((size_t) 8 * nb, 0.0);
		j->counts.assign((size_t) 2 * nb, 0ull);
	}
	auto set_err = [&](const std::string& m) {
		for(dms_job* j : jobs)
		{
			j->status = DMS_ERR_DEVICE;
			j->err = m;
		}
	};
	const double t_start = now_ms();
	ChunkBuffers B;
	DmsRows LR;
	memset(&LR, 0, sizeof(LR));
	const size_t lanes = (size_t) traj_grid(ctx, n) * DMS_TRAJ_BLOCK;
	bool ok = LJ.upload(st) == hipSuccess && ensure(B, n, nb, false, want_recs, st) == hipSuccess;
	if(ok && want_rows)
	{
		LR.per_lane = 1;
		ok = dalloc(&LR.bins, (size_t) nb * lanes, st) == hipSuccess && dalloc(&LR.touched, (size_t) nb * lanes, st) == hipSuccess
		     && dalloc(&LR.hist, 8 * (size_t) nb * nj, st) == hipSuccess && dalloc(&LR.hcounts, 2 * (size_t) nb * nj, st) == hipSuccess;
		if(ok)
			ok = hipMemsetAsync(LR.bins, 0, sizeof(DmsBinAcc) * (size_t) nb * lanes, st) == hipSuccess
			     && hipMemsetAsync(LR.hist, 0, sizeof(double) * 8 * (size_t) nb * nj, st) == hipSuccess
			     && hipMemsetAsync(LR.hcounts, 0, sizeof(unsigned long long) * 2 * (size_t) nb * nj, st) == hipSuccess;
	}
	if(!ok)
		set_err("device allocation for a batched launch failed");
	float kms = 0.0f;
	if(ok)
	{
		(void) hipGetLastError();
		for(size_t k = 0; k < nj; k++)
		{
			const dms_job_desc& dk = jobs[k]->desc;
			const DmsHaloView H = halo_view(ctx, dk.halo);
			const size_t nk = (size_t) dk.max_trajectories;
			hipLaunchKernelGGL(dms_ic_kernel, dim3((unsigned) ((nk + 255) / 256)), dim3(256), 0, st, H, dk.seed, dk.point_id,
			                   dk.first_trajectory_id, (uint32_t) nk, dk.r_max_rsun * DMS_rSun, B.ics + LJ.start[k],
			                   B.ok + LJ.start[k]);
		}
		hipError_t e = run_trajectories(ctx, st, pst, pt, J, B, n, want_rows, want_recs, LJ, want_rows ? &LR : nullptr, &kms,
		                                [&jobs](const std::vector<unsigned int>& d) {
			                                for(size_t k = 0; k < jobs.size() && DMS_PROGRESS_WORDS * k + 3 < d.size(); k++)
			                                {
				                                jobs[k]->done_traj = d[DMS_PROGRESS_WORDS * k];
				                                jobs[k]->done_captured = d[DMS_PROGRESS_WORDS * k + 1];
				                                jobs[k]->done_escaped = d[DMS_PROGRESS_WORDS * k + 2];
				                                jobs[k]->done_failed = d[DMS_PROGRESS_WORDS * k + 3];
			                                }
		                                },
		                                nullptr, cancel);
		if(e != hipSuccess)
		{
			set_err(std::string("trajectory kernel: ") + hipGetErrorString(e));
			ok = false;
		}
	}
	if(ok)
	{
		std::vector<DmsSummary> hsum(n);
		hipMemcpyAsync(hsum.data(), B.summ, sizeof(DmsSummary) * n, hipMemcpyDeviceToHost, st);
		std::vector<dms_traj_record> hrec;
		if(want_recs)
		{
			hrec.resize(n);
			hipMemcpyAsync(hrec.data(), B.recs, sizeof(dms_traj_record) * n, hipMemcpyDeviceToHost, st);
		}
		if(want_rows)
			for(size_t k = 0; k < nj; k++)
			{
				hipMemcpyAsync(jobs[k]->hist.data(), LR.hist + (size_t) k * 8 * nb, sizeof(double) * 8 * nb, hipMemcpyDeviceToHost, st);
				hipMemcpyAsync(jobs[k]->counts.data(), LR.hcounts + (size_t) k * 2 * nb, sizeof(unsigned long long) * 2 * nb,
				               hipMemcpyDeviceToHost, st);
			}
		hipStreamSynchronize(st);
		const double elapsed_sec = (now_ms() - t_start) * 1e-3;
		for(size_t k = 0; k < nj; k++)
		{
			dms_job* job = jobs[k];
			RoundReplay R;
			R.init(job->desc, job->pt, &job->C);
			const size_t nk = (size_t) job->desc.max_trajectories;
			size_t counted = 0;
			for(size_t i = 0; i < nk; i++)
			{
				if(!R.open_round())
					break;
				account_trajectory(job, R, hsum[LJ.start[k] + i], elapsed_sec);
				if(job->desc.record_trajectories)
					job->records.push_back(hrec[LJ.start[k] + i]);
				counted = i + 1;
				R.close_one();
			}
			R.open_round();
			R.finish();
			job->C.simulated_trajectories = nk;
			job->kernel_ms = kms;
			job->launches = 1.0;
			job->batch_size = (double) nj;
			job->done_traj = job->C.trajectories;
			job->done_captured = job->C.captured;
			job->done_scat = job->C.total_scatterings;
			job->done_escaped = job->C.completed_outward_escapes;
			job->done_failed = job->C.numerical_failures;
			job->needs_rerun = want_rows && counted < nk;
		}
		const double total = (now_ms() - t_start) * 1e-3;
		for(dms_job* job : jobs)
		{
			job->C.computing_time_sec = total;
			job->other_ms = total * 1e3 - kms;
		}
	}
	dfree(LR.bins, st);
	dfree(LR.touched, st);
	dfree(LR.hist, st);
	dfree(LR.hcounts, st);
	LJ.release(st);
	B.release(st);
	hipStreamSynchronize(st);
	/* A run-level fuse stopped a job before its last trajectory: its histogram
	 * sums hold trajectories past the stop, so recount it exactly (chunked). */
	for(dms_job* job : jobs)
		if(job->status == DMS_OK && job->needs_rerun && !(cancelled && cancelled->load()))
			run_job_chunked(job, st, pst, cancel, cancelled);
}

/* Device bytes of a batched launch: per-lane histogram scratch (rows, counts,
 * touched list), per-slot initial event, shift flag, summary and record, and
 * the per-job class sums. */
size_t batch_device_bytes(const dms_ctx* ctx, const std::vector<dms_job*>& jobs)
{
	const dms_job_desc& d0 = jobs[0]->desc;
	const size_t nb = (size_t) job_view(d0).nbins;
	const bool rows = d0.mode != DMS_MODE_CAPTURE;
	size_t n = 0;
	bool recs = false;
	for(const dms_job* j : jobs)
	{
		n += (size_t) j->desc.max_trajectories;
		recs = recs || j->desc.record_trajectories != 0;
	}
	const size_t lanes = (size_t) traj_grid(ctx, n) * DMS_TRAJ_BLOCK;
	size_t bytes = n * (sizeof(dms_event) + sizeof(int32_t) + sizeof(DmsSummary) + (recs ? sizeof(dms_traj_record) : 0));
	if(rows)
		bytes += lanes * nb * (sizeof(DmsBinAcc) + sizeof(uint16_t))
		         + jobs.size() * nb * (8 * sizeof(double) + 2 * sizeof(unsigned long long));
	return bytes;
}

/* Largest slot count of one batched launch (initial conditions and summaries
 * of every trajectory stay resident: about 150 B each). */
#define DMS_BATCH_MAX_SLOTS ((uint64_t) 1 << 27)

bool batchable(const dms_job_desc& d)
{
	return d.capture_target == 0 && d.max_trajectories > 0 && d.max_trajectories <= DMS_BATCH_MAX_SLOTS
	       && job_view(d).nbins <= 65536;
}

bool same_batch_config(const dms_job_desc& a, const dms_job_desc& b)
{
	return a.mode == b.mode && a.precision == b.precision && a.seed == b.seed && a.point_id == b.point_id
	       && a.max_scatterings == b.max_scatterings && a.max_free_steps == b.max_free_steps
	       && a.rk_attempt_budget == b.rk_attempt_budget && a.r_max_rsun == b.r_max_rsun && a.n_bins == b.n_bins
	       && memcmp(&a.halo, &b.halo, sizeof(dms_halo)) == 0;
}

/* Jobs of different points (and keys/halos) can share a launch when the
 * kernel-wide job view agrees. */
bool same_launch_config(const dms_job_desc& a, const dms_job_desc& b)
{
	return a.mode == b.mode && a.precision == b.precision && a.max_scatterings == b.max_scatterings
	       && a.max_free_steps == b.max_free_steps && a.rk_attempt_budget == b.rk_attempt_budget && a.r_max_rsun == b.r_max_rsun
	       && a.n_bins == b.n_bins;
}

/* A job (or batch) runs on a stream pair of its own, so jobs submitted while
 * others run overlap on the device (continuous batching: the blocks of a new
 * launch take the compute units that finished launches leave). */
void run_batch_worker(std::shared_ptr<dms_batch> batch)
{
	dms_ctx* ctx = batch->jobs[0]->pt->ctx;
	int slot;
	{
		std::unique_lock<std::mutex> lock(ctx->pool_mu);
		ctx->pool_cv.wait(lock, [ctx] { return !ctx->pool_free.empty(); });
		slot = ctx->pool_free.back();
		ctx->pool_free.pop_back();
	}
	const hipStream_t st = ctx->pool[slot], pst = ctx->progress_stream;
	hipSetDevice(ctx->device);
	/* the launch's cancel word (dms_job_cancel); none left: not cancellable */
	int word = -1;
	{
		std::lock_guard<std::mutex> lock(ctx->pool_mu);
		if(!ctx->cancel_free.empty())
		{
			word = ctx->cancel_free.back();
			ctx->cancel_free.pop_back();
		}
	}
	const unsigned int* cancel = nullptr;
	if(word >= 0)
	{
		__atomic_store_n(ctx->h_cancel + word, 0u, __ATOMIC_SEQ_CST);
		batch->cancel_word.store(word);
		if(batch->cancelled.load())
			__atomic_store_n(ctx->h_cancel + word, 1u, __ATOMIC_SEQ_CST);
		cancel = ctx->d_cancel + word;
	}
	if(!batch->cancelled.load())
	{
		if(batch->jobs.size() == 1 && !batchable(batch->jobs[0]->desc))
			run_job_chunked(batch->jobs[0], st, pst, cancel, &batch->cancelled);
		else if(batch_device_bytes(ctx, batch->jobs) > ctx->job_memory_budget)
		{
			/* a batch whose buffers exceed the job memory budget: each job as
			 * chunked launches within the budget (the same results) */
			for(dms_job* j : batch->jobs)
				if(!batch->cancelled.load())
					run_job_chunked(j, st, pst, cancel, &batch->cancelled);
		}
		else
			run_batch_fixed(batch->jobs, st, pst, cancel, &batch->cancelled);
	}
	if(batch->cancelled.load())
		for(dms_job* j : batch->jobs)
		{
			j->status = DMS_ERR_CANCELLED;
			j->err = "cancelled (dms_job_cancel)";
		}
	{
		std::lock_guard<std::mutex> lock(ctx->pool_mu);
		ctx->pool_free.push_back(slot);
		if(word >= 0)
		{
			batch->cancel_word.store(-1);
			ctx->cancel_free.push_back(word);
		}
	}
	ctx->pool_cv.notify_one();
	for(dms_job* j : batch->jobs)
		j->finished = 1;
}

} // namespace

/* ---- serving: requests join one resident launch as slots -------------------
 *
 * One persistent trajectory launch (dms_traj_kernel_serve) stays on the device
 * while requests arrive; a host thread (the driver) appends each admitted
 * request's trajectories to the launch's work queue as slots, exactly as the
 * reference's ranks take trajectories one by one inside Generate_Data's sync
 * rounds (src/Data_Generation.cpp:697-728), and folds every handed-back
 * summary into its request strictly in trajectory-id order with the
 * reference's round logic (RoundReplay, account_trajectory).  Histogram rows
 * are summed on the device per sync round of a request (the run-level stops
 * fall on round boundaries only, so a request's counted trajectories are
 * always whole rounds); the driver adds the counted rounds' sums in round
 * order.  The host and the running launch talk only through host memory the
 * device maps (slot descriptors, request table, summaries, completion queue,
 * round mailboxes): no copy or kernel has to run beside the launch, so it
 * never waits on a hardware queue or compute units it holds itself.  With no
 * request left the driver parks the launch (stop flag, launch ends) and
 * relaunches it for the next request. */
#define DMS_SERVE_COUNTERS 8

struct ServeRoundSums
{
	std::vector<double> h;
	std::vector<unsigned long long> c;
};

struct ServeRequest
{
	dms_job* job = nullptr;
	uint32_t slot = 0; /* request table index */
	RoundReplay R;
	uint64_t interval = 1, limit = UINT64_MAX;
	uint64_t n_pub = 0;    /* trajectories published (offsets 0 .. n_pub-1) */
	uint64_t n_acc = 0;    /* replay cursor: offsets accounted */
	uint64_t in_flight = 0; /* published, not yet handed back */
	uint64_t rounds_out = 0; /* round buffers not yet returned */
	uint64_t simulated = 0;
	uint32_t cur_round = DMS_SERVE_NO_ROUND; /* buffer of the round being published */
	uint64_t cur_round_end = 0;              /* offset past the round being published */
	std::deque<DmsSummary> sums; /* offsets n_acc.. */
	std::deque<uint8_t> have;
	uint64_t counted_rounds = 0; /* rounds [0, counted_rounds) count */
	uint64_t next_add = 0;       /* next counted round to add to the histogram */
	std::map<uint64_t, ServeRoundSums> mb;
	bool final_count = false; /* replay stopped: the counted set is final */
	bool cancelled = false;
	bool by_request = false; /* stopped by dms_job_cancel */
	double t_start = 0.0;
};

struct ServeRoundBuf
{
	ServeRequest* req = nullptr;
	uint64_t k = 0, gen = 0;
};

template <typename T>
hipError_t host_alloc(T** p, size_t count)
{
	*p = nullptr;
	hipError_t e = hipHostMalloc((void**) p, sizeof(T) * (count ? count : 1), hipHostMallocMapped | hipHostMallocCoherent);
	if(e == hipSuccess)
		memset((void*) *p, 0, sizeof(T) * (count ? count : 1));
	return e;
}

template <typename T>
T* device_view(T* host)
{
	void* d = nullptr;
	if(host == nullptr || hipHostGetDevicePointer(&d, (void*) host, 0) != hipSuccess)
		return nullptr;
	return (T*) d;
}

struct dms_server
{
	dms_ctx* ctx = nullptr;
	dms_job_desc cfg;
	DmsJobView J;
	bool want_rows = false;
	int nb = 0;
	uint32_t ring = 0, mask = 0, max_reqs = 0, max_rounds = 0;
	int grid = 0;
	size_t lanes = 0;
	/* host memory the device maps */
	unsigned long long* h_ctl = nullptr;
	DmsServeSlot* h_slots = nullptr;
	DmsServeReq* h_reqs = nullptr;
	DmsSummary* h_results = nullptr;
	uint32_t* h_cq = nullptr;
	uint32_t* h_round_size = nullptr;
	unsigned long long *h_round_gen = nullptr, *h_round_stamp = nullptr;
	double* h_mb_hist = nullptr;
	unsigned long long* h_mb_counts = nullptr;
	/* device memory */
	unsigned long long* d_counters = nullptr;
	uint32_t* d_round_done = nullptr;
	double* d_round_hist = nullptr;
	unsigned long long* d_round_counts = nullptr;
	DmsBinAcc* d_bins = nullptr;
	uint16_t* d_touched = nullptr;
	/* driver state (driver thread only) */
	std::vector<ServeRequest*> pos_req;
	std::vector<uint64_t> pos_off;
	uint64_t pub = 0, cq_cursor = 0, gen = 0;
	std::vector<ServeRoundBuf> rounds;
	std::vector<uint32_t> free_rounds;
	std::vector<uint32_t> free_reqs;
	std::vector<ServeRequest*> active;
	size_t rr = 0;
	bool running = false;
	int pool_slot = -1;
	hipStream_t st = nullptr;
	hipEvent_t ended = nullptr;
	double idle_since = -1.0, last_check = 0.0;
	/* shared with submitting threads */
	std::mutex mu;
	std::condition_variable cv;
	std::deque<dms_job*> pending;
	bool closing = false;
	std::thread driver;
	std::atomic<uint64_t> n_launches{0}, n_published{0}, n_cancelled{0}, n_done{0}, n_active{0};
	std::atomic<int> is_running{0};
};

namespace
{

void serve_release(dms_server* S)
{
	hipHostFree(S->h_ctl);
	hipHostFree(S->h_slots);
	hipHostFree(S->h_reqs);
	hipHostFree(S->h_results);
	hipHostFree(S->h_cq);
	hipHostFree(S->h_round_size);
	hipHostFree(S->h_round_gen);
	hipHostFree(S->h_round_stamp);
	hipHostFree(S->h_mb_hist);
	hipHostFree(S->h_mb_counts);
	const hipStream_t cs = S->ctx->stream;
	std::lock_guard<std::mutex> lock(S->ctx->mu);
	dfree(S->d_counters, cs);
	dfree(S->d_round_done, cs);
	dfree(S->d_round_hist, cs);
	dfree(S->d_round_counts, cs);
	dfree(S->d_bins, cs);
	dfree(S->d_touched, cs);
	hipStreamSynchronize(cs);
	if(S->ended)
		hipEventDestroy(S->ended);
}

void serve_finish_job(dms_server* S, dms_job* job, int status, const std::string& err)
{
	job->status = status;
	job->err = err;
	S->n_done++;
	std::lock_guard<std::mutex> lock(job->fmu);
	job->finished = 1;
	job->fcv.notify_all();
}

/* (Re)start the resident launch: a stream from the context's pool, cleared
 * counters, the stop flag down, slots handed out from the published count. */
bool serve_launch(dms_server* S, std::string& err)
{
	dms_ctx* ctx = S->ctx;
	{
		std::unique_lock<std::mutex> lock(ctx->pool_mu);
		ctx->pool_cv.wait(lock, [ctx] { return !ctx->pool_free.empty(); });
		S->pool_slot = ctx->pool_free.back();
		ctx->pool_free.pop_back();
	}
	S->st = ctx->pool[S->pool_slot];
	hipSetDevice(ctx->device);
	DmsTrajArgs a;
	memset(&a, 0, sizeof(a));
	a.S = ctx->dev;
	a.J = S->J;
	a.R.per_lane = S->want_rows ? 1 : 0;
	a.R.bins = S->d_bins;
	a.R.touched = S->d_touched;
	a.R.hist = S->d_round_hist;
	a.R.hcounts = S->d_round_counts;
	a.V.ctl = device_view(S->h_ctl);
	a.V.base = S->pub;
	a.V.ring_mask = S->mask;
	a.V.slots = device_view(S->h_slots);
	a.V.reqs = device_view(S->h_reqs);
	a.V.results = device_view(S->h_results);
	a.V.cq = device_view(S->h_cq);
	a.V.counters = S->d_counters;
	a.V.round_size = device_view(S->h_round_size);
	a.V.round_gen = device_view(S->h_round_gen);
	a.V.round_stamp = device_view(S->h_round_stamp);
	a.V.mb_hist = device_view(S->h_mb_hist);
	a.V.mb_counts = device_view(S->h_mb_counts);
	a.V.round_done = S->d_round_done;
	if(!a.V.ctl || !a.V.slots || !a.V.reqs || !a.V.results || !a.V.cq || !a.V.round_size || !a.V.round_gen || !a.V.round_stamp
	   || (S->want_rows && (!a.V.mb_hist || !a.V.mb_counts)))
	{
		err = "hipHostGetDevicePointer failed for the serving launch's host memory";
		{
			std::lock_guard<std::mutex> lock(ctx->pool_mu);
			ctx->pool_free.push_back(S->pool_slot);
		}
		ctx->pool_cv.notify_one();
		S->pool_slot = -1;
		return false;
	}
	S->cq_cursor = 0;
	__atomic_store_n(S->h_ctl + 8, 0ull, __ATOMIC_RELEASE);
	(void) hipGetLastError();
	hipError_t e = hipMemsetAsync(S->d_counters, 0, DMS_SERVE_COUNTERS * sizeof(unsigned long long), S->st);
	if(e == hipSuccess)
	{
		void (*kernel)(DmsTrajArgs) = S->J.mixed ? dms_traj_kernel_mixed_serve : dms_traj_kernel_serve;
		hipLaunchKernelGGL(kernel, dim3(S->grid), dim3(DMS_TRAJ_BLOCK), ctx->traj_lds, S->st, a);
		e = hipGetLastError();
	}
	if(e == hipSuccess)
		e = hipEventRecord(S->ended, S->st);
	if(e != hipSuccess)
	{
		err = std::string("serving launch: ") + hipGetErrorString(e);
		hipStreamSynchronize(S->st);
		{
			std::lock_guard<std::mutex> lock(ctx->pool_mu);
			ctx->pool_free.push_back(S->pool_slot);
		}
		ctx->pool_cv.notify_one();
		S->pool_slot = -1;
		return false;
	}
	S->running = true;
	S->is_running = 1;
	S->n_launches++;
	return true;
}

/* Park the launch: every published slot has been handed back, so the lanes
 * hold unpublished slots only and end at the stop flag. */
bool serve_park(dms_server* S, std::string& err)
{
	__atomic_store_n(S->h_ctl + 8, 1ull, __ATOMIC_RELEASE);
	const hipError_t e = hipEventSynchronize(S->ended);
	S->running = false;
	S->is_running = 0;
	{
		std::lock_guard<std::mutex> lock(S->ctx->pool_mu);
		S->ctx->pool_free.push_back(S->pool_slot);
	}
	S->ctx->pool_cv.notify_one();
	S->pool_slot = -1;
	if(e != hipSuccess)
	{
		err = std::string("serving launch: ") + hipGetErrorString(e);
		return false;
	}
	return true;
}

bool serve_admit(dms_server* S, dms_job* job)
{
	if(S->free_reqs.empty())
		return false;
	ServeRequest* q = new ServeRequest();
	q->job = job;
	q->slot = S->free_reqs.back();
	S->free_reqs.pop_back();
	const dms_job_desc& d = job->desc;
	memset(&job->C, 0, sizeof(job->C));
	job->events.clear();
	job->records.clear();
	job->hist.assign((size_t) 8 * S->nb, 0.0);
	job->counts.assign((size_t) 2 * S->nb, 0ull);
	q->R.init(d, job->pt, &job->C);
	q->interval = q->R.interval > 0 ? q->R.interval : 1;
	q->limit = q->R.unlimited ? UINT64_MAX : q->R.max_traj;
	q->t_start = now_ms();
	DmsServeReq r;
	memset(&r, 0, sizeof(r));
	r.P = job->pt->dev;
	r.H = halo_view(S->ctx, d.halo);
	r.seed = d.seed;
	r.point_id = d.point_id;
	r.cancel = 0;
	/* written before any slot of the request is published (release below) */
	S->h_reqs[q->slot] = r;
	S->active.push_back(q);
	S->n_active = S->active.size();
	return true;
}

/* Trajectories the request should have published by now: all of a fixed
 * count; for a capture target, the expected remainder with margin, as the
 * chunked path sizes its chunks. */
uint64_t serve_want(const ServeRequest* q)
{
	if(q->R.target == UINT64_MAX)
		return q->limit;
	const dms_counters& C = q->job->C;
	const double rate = (C.trajectories > 0 && C.captured > 0) ? (double) C.captured / (double) C.trajectories : 0.0;
	const uint64_t need = q->R.target > q->R.captured ? q->R.target - q->R.captured : 1;
	uint64_t ahead = rate > 0.0 ? (uint64_t) (1.5 * (double) need / rate) + 4096 : 16384;
	uint64_t want = q->n_acc + ahead;
	if(want > q->limit)
		want = q->limit;
	return want;
}

/* Append slots, round-robin over the requests, a quantum each per pass; a
 * cancelled request only completes the round it had started (as handed-back
 * slots).  Returns the number of slots published. */
uint64_t serve_publish(dms_server* S)
{
	const uint64_t first = S->pub;
	const size_t n = S->active.size();
	const uint64_t quantum = std::max<uint64_t>(1024, S->ring / 16);
	bool ring_full = false;
	for(size_t i = 0; i < n && !ring_full; i++)
	{
		ServeRequest* q = S->active[(S->rr + i) % n];
		uint64_t want;
		if(q->cancelled)
			want = (q->cur_round != DMS_SERVE_NO_ROUND) ? q->cur_round_end : q->n_pub;
		else
			want = serve_want(q);
		/* one request holds at most half the ring */
		const uint64_t cap = q->in_flight < S->ring / 2 ? S->ring / 2 - q->in_flight : 0;
		uint64_t budget = std::min<uint64_t>(quantum, cap);
		while(budget > 0 && q->n_pub < want)
		{
			if(S->want_rows && q->cur_round == DMS_SERVE_NO_ROUND)
			{
				if(S->free_rounds.empty())
					break;
				const uint32_t b = S->free_rounds.back();
				S->free_rounds.pop_back();
				const uint64_t k = q->n_pub / q->interval;
				const uint64_t end = std::min<uint64_t>((k + 1) * q->interval, q->limit);
				ServeRoundBuf& rb = S->rounds[b];
				rb.req = q;
				rb.k = k;
				rb.gen = ++S->gen;
				S->h_round_size[b] = (uint32_t) (end - q->n_pub);
				S->h_round_gen[b] = rb.gen;
				q->cur_round = b;
				q->cur_round_end = end;
				q->rounds_out++;
			}
			const uint64_t p = S->pub & S->mask;
			if(S->pos_req[p] != nullptr)
			{
				ring_full = true;
				break;
			}
			DmsServeSlot sl;
			sl.req = q->slot;
			sl.round = S->want_rows ? q->cur_round : DMS_SERVE_NO_ROUND;
			sl.id = q->job->desc.first_trajectory_id + q->n_pub;
			S->h_slots[p] = sl;
			S->pos_req[p] = q;
			S->pos_off[p] = q->n_pub;
			S->pub++;
			q->n_pub++;
			q->in_flight++;
			budget--;
			if(S->want_rows && q->n_pub == q->cur_round_end)
				q->cur_round = DMS_SERVE_NO_ROUND;
		}
	}
	if(n > 0)
		S->rr = (S->rr + 1) % n;
	if(S->pub != first)
	{
		__atomic_store_n(S->h_ctl, (unsigned long long) S->pub, __ATOMIC_RELEASE);
		S->n_published += S->pub - first;
	}
	return S->pub - first;
}

/* Hand-backs from the completion queue: the summary moves to its request,
 * the ring position is free again. */
uint64_t serve_drain_cq(dms_server* S)
{
	uint64_t k = 0;
	for(;;)
	{
		uint32_t* e = S->h_cq + (S->cq_cursor & S->mask);
		const uint32_t v = __atomic_load_n(e, __ATOMIC_ACQUIRE);
		if(v == 0)
			break;
		__atomic_store_n(e, 0u, __ATOMIC_RELAXED);
		S->cq_cursor++;
		k++;
		const uint32_t p = v - 1;
		ServeRequest* q = S->pos_req[p];
		const uint64_t off = S->pos_off[p];
		DmsSummary sm;
		memcpy(&sm, (const void*) (S->h_results + p), sizeof(sm));
		S->pos_req[p] = nullptr;
		if(q == nullptr)
			continue;
		q->in_flight--;
		if(sm.reason == DMS_SERVE_CANCELLED)
		{
			S->n_cancelled++;
			continue;
		}
		q->simulated++;
		if(q->final_count || off < q->n_acc)
			continue;
		const uint64_t i = off - q->n_acc;
		if(q->sums.size() <= i)
		{
			q->sums.resize(i + 1);
			q->have.resize(i + 1, 0);
		}
		q->sums[i] = sm;
		q->have[i] = 1;
	}
	return k;
}

/* Round mailboxes stamped by the device: kept while the round may count. */
uint64_t serve_drain_rounds(dms_server* S)
{
	uint64_t k = 0;
	const size_t nb = (size_t) S->nb;
	for(uint32_t b = 0; b < S->max_rounds; b++)
	{
		ServeRoundBuf& rb = S->rounds[b];
		if(rb.req == nullptr || __atomic_load_n(S->h_round_stamp + b, __ATOMIC_ACQUIRE) != rb.gen)
			continue;
		ServeRequest* q = rb.req;
		if(!q->final_count || rb.k < q->counted_rounds)
		{
			ServeRoundSums& m = q->mb[rb.k];
			m.h.assign(S->h_mb_hist + (size_t) b * 8 * nb, S->h_mb_hist + (size_t) (b + 1) * 8 * nb);
			m.c.assign(S->h_mb_counts + (size_t) b * 2 * nb, S->h_mb_counts + (size_t) (b + 1) * 2 * nb);
		}
		q->rounds_out--;
		rb.req = nullptr;
		S->free_rounds.push_back(b);
		k++;
	}
	return k;
}

/* Replay of the request's summaries in id order, the counted rounds' sums
 * added in round order. */
void serve_replay(dms_server* S, ServeRequest* q)
{
	dms_job* job = q->job;
	const double elapsed_sec = (now_ms() - q->t_start) * 1e-3;
	if(!q->final_count && job->cancel_requested.load())
	{
		q->final_count = true;
		q->by_request = true;
	}
	while(!q->final_count)
	{
		if(!q->R.open_round())
		{
			q->final_count = true;
			break;
		}
		if(S->want_rows)
			q->counted_rounds = std::max<uint64_t>(q->counted_rounds, q->n_acc / q->interval + 1);
		if(q->have.empty() || !q->have.front())
			break;
		account_trajectory(job, q->R, q->sums.front(), elapsed_sec);
		q->R.close_one();
		q->sums.pop_front();
		q->have.pop_front();
		q->n_acc++;
	}
	if(q->final_count && !q->cancelled)
	{
		q->cancelled = true;
		__atomic_store_n(&S->h_reqs[q->slot].cancel, 1u, __ATOMIC_RELAXED);
		q->sums.clear();
		q->have.clear();
		/* sums of rounds past the counted ones are not needed */
		for(auto it = q->mb.begin(); it != q->mb.end();)
			it = (it->first >= q->counted_rounds) ? q->mb.erase(it) : std::next(it);
	}
	const size_t nb = (size_t) S->nb;
	while(q->next_add < q->counted_rounds)
	{
		auto it = q->mb.find(q->next_add);
		if(it == q->mb.end())
			break;
		for(size_t i = 0; i < 8 * nb; i++)
			job->hist[i] += it->second.h[i];
		for(size_t i = 0; i < 2 * nb; i++)
			job->counts[i] += it->second.c[i];
		q->mb.erase(it);
		q->next_add++;
	}
	job->done_traj = job->C.trajectories;
	job->done_captured = job->C.captured;
	job->done_scat = job->C.total_scatterings;
	job->done_escaped = job->C.completed_outward_escapes;
	job->done_failed = job->C.numerical_failures;
}

bool serve_request_done(const dms_server* S, const ServeRequest* q)
{
	return q->final_count && q->in_flight == 0 && q->rounds_out == 0 && (!S->want_rows || q->next_add >= q->counted_rounds);
}

void serve_fail_all(dms_server* S, const std::string& err)
{
	for(ServeRequest* q : S->active)
	{
		serve_finish_job(S, q->job, DMS_ERR_DEVICE, err);
		delete q;
	}
	S->active.clear();
	S->n_active = 0;
	std::deque<dms_job*> rest;
	{
		std::lock_guard<std::mutex> lock(S->mu);
		rest.swap(S->pending);
	}
	for(dms_job* j : rest)
		serve_finish_job(S, j, DMS_ERR_DEVICE, err);
}

void serve_driver(dms_server* S)
{
	hipSetDevice(S->ctx->device);
	bool broken = false;
	std::string broken_err;
	for(;;)
	{
		/* new requests */
		std::deque<dms_job*> arrivals;
		bool closing;
		{
			std::unique_lock<std::mutex> lock(S->mu);
			if(S->pending.empty() && S->active.empty() && !S->running)
				S->cv.wait(lock, [S] { return S->closing || !S->pending.empty(); });
			closing = S->closing;
			while(!S->pending.empty() && !S->free_reqs.empty() && arrivals.size() < S->free_reqs.size())
			{
				arrivals.push_back(S->pending.front());
				S->pending.pop_front();
			}
		}
		if(broken)
		{
			for(dms_job* j : arrivals)
				serve_finish_job(S, j, DMS_ERR_DEVICE, broken_err);
			if(closing)
				break;
			continue;
		}
		for(dms_job* j : arrivals)
			serve_admit(S, j);
		if(closing && S->active.empty() && arrivals.empty())
		{
			bool more;
			{
				std::lock_guard<std::mutex> lock(S->mu);
				more = !S->pending.empty();
			}
			if(!more)
				break;
		}
		std::string err;
		if(!S->active.empty() && !S->running && !serve_launch(S, err))
		{
			broken = true;
			broken_err = err;
			serve_fail_all(S, err);
			continue;
		}
		uint64_t work = 0;
		work += serve_drain_cq(S);
		if(S->want_rows)
			work += serve_drain_rounds(S);
		for(size_t i = 0; i < S->active.size();)
		{
			ServeRequest* q = S->active[i];
			serve_replay(S, q);
			if(serve_request_done(S, q))
			{
				dms_job* job = q->job;
				q->R.finish();
				job->C.simulated_trajectories = q->simulated;
				job->C.computing_time_sec = (now_ms() - q->t_start) * 1e-3;
				job->kernel_ms = job->C.computing_time_sec * 1e3;
				job->launches = 0.0;
				job->other_ms = 0.0;
				S->free_reqs.push_back(q->slot);
				S->active.erase(S->active.begin() + (long) i);
				S->n_active = S->active.size();
				const bool by_request = q->by_request;
				delete q;
				if(by_request)
					serve_finish_job(S, job, DMS_ERR_CANCELLED, "cancelled (dms_job_cancel)");
				else
					serve_finish_job(S, job, DMS_OK, "");
				work++;
				continue;
			}
			i++;
		}
		work += serve_publish(S);
		const double t = now_ms();
		if(S->running)
		{
			if(!S->active.empty())
				S->idle_since = -1.0;
			else if(S->idle_since < 0.0)
				S->idle_since = t;
			bool pending_empty;
			{
				std::lock_guard<std::mutex> lock(S->mu);
				pending_empty = S->pending.empty();
			}
			/* park after 2 ms without requests (at once when closing) */
			if(S->active.empty() && pending_empty && (closing || t - S->idle_since > 2.0))
			{
				if(!serve_park(S, err))
				{
					broken = true;
					broken_err = err;
					serve_fail_all(S, err);
				}
				continue;
			}
			/* a launch that ended by itself, or faulted */
			if(t - S->last_check > 10.0)
			{
				S->last_check = t;
				const hipError_t e = hipEventQuery(S->ended);
				if(e != hipErrorNotReady)
				{
					broken = true;
					broken_err = e == hipSuccess ? std::string("serving launch ended unexpectedly")
					                             : std::string("serving launch: ") + hipGetErrorString(e);
					S->running = false;
					S->is_running = 0;
					serve_fail_all(S, broken_err);
					continue;
				}
			}
		}
		if(work == 0)
			std::this_thread::sleep_for(std::chrono::microseconds(100));
	}
	if(S->running)
	{
		std::string err;
		serve_park(S, err);
	}
}

} // namespace

/* Scratch memory of dms_simulate_events_noalloc: the launch's slot counter and
 * fill-control words, and its one-job table (first slot, first trajectory id).
 * Each GPU holds its own copy for the life of the process; host code reaches
 * it only through the address hipGetSymbolAddress returns for the current
 * device (the array's name in host code is not a device address). */
#define DMS_NOALLOC_SCRATCH_BYTES 256
#define DMS_NOALLOC_MAX_TRAJECTORIES 0x7fffffffu
static __device__ __align__(256) unsigned char scratch[DMS_NOALLOC_SCRATCH_BYTES];

namespace
{
/* byte offsets inside scratch */
constexpr size_t kScratchNext = 0;     /* unsigned int[4]: slot counter, fill control */
constexpr size_t kScratchStart = 64;   /* uint32_t job_start[1] */
constexpr size_t kScratchFirst = 128;  /* uint64_t job_first_id[1] */

std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[256] = {};

/* once per device, with that device current */
hipError_t scratch_address(int device, unsigned char** out)
{
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if(device < 0 || device >= 256)
		return hipErrorInvalidDevice;
	if(g_scratch_ptr[device] == nullptr)
	{
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if(e != hipSuccess)
			return e;
		g_scratch_ptr[device] = static_cast<unsigned char*>(p);
	}
	*out = g_scratch_ptr[device];
	return hipSuccess;
}
} // namespace

/* zero the slot counter and fill control, write the one-job table (a kernel,
 * not hipMemsetAsync, so that the launch sequence can be captured in a graph) */
__global__ void dms_noalloc_setup_kernel(unsigned char* scratch_dev, uint32_t n, uint64_t first_id)
{
	if(threadIdx.x < 4)
		reinterpret_cast<unsigned int*>(scratch_dev + kScratchNext)[threadIdx.x] = (threadIdx.x == 2) ? n : 0u;
	if(threadIdx.x == 0)
	{
		*reinterpret_cast<uint32_t*>(scratch_dev + kScratchStart) = 0u;
		*reinterpret_cast<uint64_t*>(scratch_dev + kScratchFirst) = first_id;
	}
}

__global__ void dms_zero_kernel(double* p, size_t n)
{
	for(size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x)
		p[i] = 0.0;
}

/* ======================= C ABI ============================================= */

extern "C" {

int dms_abi_version(void) { return DMS_ABI_VERSION; }

const char* dms_last_error(const dms_ctx* ctx)
{
	if(ctx && !ctx->err.empty())
		return ctx->err.c_str();
	return g_last_error.c_str();
}

double dms_constant(const char* name)
{
	static const struct
	{
		const char* name;
		double value;
	} table[] = {{"cm", DMS_cm}, {"meter", DMS_meter}, {"km", DMS_km}, {"sec", DMS_sec}, {"gram", DMS_gram},
	             {"kg", DMS_kg}, {"Kelvin", DMS_Kelvin}, {"eV", DMS_eV}, {"MeV", DMS_MeV}, {"GeV", DMS_GeV},
	             {"mProton", DMS_mProton}, {"mElectron", DMS_mElectron}, {"mNucleon", DMS_mNucleon},
	             {"aEM", DMS_aEM}, {"G_Newton", DMS_G_Newton}, {"mSun", DMS_mSun}, {"rSun", DMS_rSun},
	             {"AU", DMS_AU}, {"pb", 1.0e-36 * DMS_cm * DMS_cm}};
	if(name == nullptr)
		return dm_nan();
	for(const auto& e : table)
		if(strcmp(e.name, name) == 0)
			return e.value;
	return dm_nan();
}

int dms_load_solar_file(const char* path, double* rows_out, size_t capacity_rows, size_t* n_rows)
{
	if(path == nullptr || n_rows == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_load_solar_file: null argument");
	FILE* f = fopen(path, "r");
	if(f == nullptr)
		return fail(nullptr, DMS_ERR_ARG, std::string("cannot open solar model file ") + path);
	char line[4096];
	size_t rows = 0;
	int lineno = 0;
	while(fgets(line, sizeof(line), f))
	{
		if(lineno++ < 20)
			continue;
		double v[dms::kColumns];
		int k = 0;
		char* p = line;
		while(k < dms::kColumns)
		{
			char* end = nullptr;
			const double x = strtod(p, &end);
			if(end == p)
				break;
			v[k++] = x;
			p = end;
		}
		if(k != dms::kColumns)
			continue;
		if(rows_out != nullptr && rows < capacity_rows)
			memcpy(rows_out + rows * dms::kColumns, v, sizeof(v));
		rows++;
	}
	fclose(f);
	*n_rows = rows;
	return DMS_OK;
}

int dms_create(dms_ctx** out, int device, const double* raw_rows, size_t n_rows)
{
	if(out == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_create: null output handle");
	*out = nullptr;
	dms_ctx* ctx = new dms_ctx();
	const std::string msg = dms::build_solar_tables(raw_rows, n_rows, ctx->T);
	if(!msg.empty())
	{
		delete ctx;
		return fail(nullptr, DMS_ERR_ARG, "dms_create: " + msg);
	}
	ctx->device = device;
	if(device >= 0)
	{
		int count = 0;
		if(hipGetDeviceCount(&count) != hipSuccess || device >= count)
		{
			delete ctx;
			return fail(nullptr, DMS_ERR_DEVICE, "dms_create: no such HIP device");
		}
		auto cleanup = [&](const char* what) {
			delete ctx;
			return fail(nullptr, DMS_ERR_DEVICE, std::string("dms_create: ") + what);
		};
		if(hipSetDevice(device) != hipSuccess)
			return cleanup("hipSetDevice failed");
		if(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess
		   || hipStreamCreateWithFlags(&ctx->progress_stream, hipStreamNonBlocking) != hipSuccess)
			return cleanup("hipStreamCreate failed");
		for(int k = 0; k < DMS_CONCURRENT_JOBS; k++)
		{
			hipStream_t sk = nullptr;
			if(hipStreamCreateWithFlags(&sk, hipStreamNonBlocking) != hipSuccess)
				return cleanup("hipStreamCreate failed");
			ctx->pool.push_back(sk);
		}
		for(int k = 0; k < DMS_CONCURRENT_JOBS; k++)
			ctx->pool_free.push_back(k);
		hipDeviceProp_t prop;
		if(hipGetDeviceProperties(&prop, device) != hipSuccess)
			return cleanup("hipGetDeviceProperties failed");
		ctx->n_cu = prop.multiProcessorCount;
		const size_t lds = dms_lds_bytes(ctx->T.n);
		int max_lds = 0;
		if(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || (size_t) max_lds < lds + 16384)
			return cleanup("solar profile too large for the LDS-resident Mass/escape-speed tables");
		if(hipFuncSetAttribute((const void*) dms_traj_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds) != hipSuccess
		   || hipFuncSetAttribute((const void*) dms_traj_kernel_mixed, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds) != hipSuccess
		   || hipFuncSetAttribute((const void*) dms_traj_kernel_multi, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds) != hipSuccess
		   || hipFuncSetAttribute((const void*) dms_traj_kernel_mixed_multi, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds)
		          != hipSuccess
		   || hipFuncSetAttribute((const void*) dms_traj_kernel_serve, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds) != hipSuccess
		   || hipFuncSetAttribute((const void*) dms_traj_kernel_mixed_serve, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds)
		          != hipSuccess)
			return cleanup("hipFuncSetAttribute(MaxDynamicSharedMemorySize) failed");
		int blocks = 0;
		if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, dms_traj_kernel, DMS_TRAJ_BLOCK, lds) != hipSuccess)
			blocks = 1;
		ctx->traj_blocks_per_cu = blocks > 0 ? blocks : 1;
		ctx->traj_lds = lds;
		const dms::SolarTables& T = ctx->T;
		if(dev_alloc_copy(&ctx->d_x, T.x.data(), T.x.size()) != hipSuccess
		   || dev_alloc_copy(&ctx->d_mass, T.mass.data(), T.mass.size()) != hipSuccess
		   || dev_alloc_copy(&ctx->d_temp, T.temp.data(), T.temp.size()) != hipSuccess
		   || dev_alloc_copy(&ctx->d_vesc2, T.vesc2.data(), T.vesc2.size()) != hipSuccess
		   || dev_alloc_copy(&ctx->d_dens, T.dens.data(), T.dens.size()) != hipSuccess
		   || dev_alloc_copy(&ctx->d_pref, T.iso_pref.data(), T.iso_pref.size()) != hipSuccess
		   || dev_alloc_copy(&ctx->d_imass, T.iso_mass.data(), T.iso_mass.size()) != hipSuccess
		   || dev_alloc_copy(&ctx->d_Z, T.iso_Z.data(), T.iso_Z.size()) != hipSuccess)
			return cleanup("device table upload failed");
		ctx->dev = T.view();
		ctx->dev.x = ctx->d_x;
		ctx->dev.mass = ctx->d_mass;
		ctx->dev.temp = ctx->d_temp;
		ctx->dev.vesc2 = ctx->d_vesc2;
		ctx->dev.dens = ctx->d_dens;
		ctx->dev.iso_pref = ctx->d_pref;
		ctx->dev.iso_mass = ctx->d_imass;
		ctx->dev.iso_Z = ctx->d_Z;
		if(scratch_address(device, &ctx->scratch_ptr) != hipSuccess)
			return cleanup("hipGetSymbolAddress of the scratch array failed");
		if(hipHostMalloc((void**) &ctx->h_cancel, sizeof(unsigned int) * DMS_CANCEL_WORDS, hipHostMallocMapped | hipHostMallocCoherent)
		       != hipSuccess
		   || hipHostGetDevicePointer((void**) &ctx->d_cancel, ctx->h_cancel, 0) != hipSuccess)
			return cleanup("pinned allocation of the cancel words failed");
		memset(ctx->h_cancel, 0, sizeof(unsigned int) * DMS_CANCEL_WORDS);
		for(int k = DMS_CANCEL_WORDS - 1; k >= 0; k--)
			ctx->cancel_free.push_back(k);
	}
	*out = ctx;
	return DMS_OK;
}

void dms_destroy(dms_ctx* ctx)
{
	if(ctx == nullptr)
		return;
	if(ctx->device >= 0)
	{
		hipFree(ctx->d_x);
		hipFree(ctx->d_mass);
		hipFree(ctx->d_temp);
		hipFree(ctx->d_vesc2);
		hipFree(ctx->d_dens);
		hipFree(ctx->d_pref);
		hipFree(ctx->d_imass);
		hipFree(ctx->d_Z);
		if(ctx->stream)
			hipStreamDestroy(ctx->stream);
		if(ctx->progress_stream)
			hipStreamDestroy(ctx->progress_stream);
		for(hipStream_t sk : ctx->pool)
			hipStreamDestroy(sk);
		if(ctx->h_cancel)
			hipHostFree(ctx->h_cancel);
	}
	delete ctx;
}

int dms_set_job_memory_budget(dms_ctx* ctx, size_t bytes)
{
	if(ctx == nullptr || bytes < ((size_t) 64 << 20))
		return fail(ctx, DMS_ERR_ARG, "dms_set_job_memory_budget: null context or budget below 64 MiB");
	ctx->job_memory_budget = bytes;
	return DMS_OK;
}

int dms_solar_eval(const dms_ctx* ctx, int which, const double* r, double* out, size_t n)
{
	if(ctx == nullptr || (n > 0 && (r == nullptr || out == nullptr)))
		return fail(nullptr, DMS_ERR_ARG, "dms_solar_eval: null argument");
	const DmsSolarView V = ctx->T.view();
	for(size_t i = 0; i < n; i++)
	{
		const double x = r[i];
		double res;
		if(which == 0)
			res = dms_mass(V, x);
		else if(which == 1)
			res = dms_temperature(V, x);
		else if(which == 2)
			res = dms_local_escape_speed(V, x);
		else if(which == 3)
			res = ctx->T.mass_density(x);
		else if(which == 4 || (which >= 100 && which - 100 < V.n_iso))
		{
			const int k = (which == 4) ? DMS_MAX_TARGETS - 1 : which - 100;
			double xx = x;
			if(!dm_isfinite(xx))
				res = dm_nan();
			else
			{
				if(xx < 0.0)
					xx = 0.0;
				if(xx > DMS_rSun)
					res = 0.0;
				else
				{
					const int j = dms_interval(V, xx);
					res = dms_density_at(V, j, xx - V.x[j], k);
				}
			}
		}
		else
			return fail(nullptr, DMS_ERR_ARG, "dms_solar_eval: unknown quantity");
		out[i] = res;
	}
	return DMS_OK;
}

int dms_isotopes(const dms_ctx* ctx, int32_t* Z, int32_t* A, double* mass, double* abundance, size_t capacity, size_t* n)
{
	if(ctx == nullptr || n == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_isotopes: null argument");
	*n = (size_t) ctx->T.n_iso;
	for(size_t i = 0; i < *n && i < capacity; i++)
	{
		if(Z)
			Z[i] = ctx->T.iso_Z[i];
		if(A)
			A[i] = ctx->T.iso_A[i];
		if(mass)
			mass[i] = ctx->T.iso_mass[i];
		if(abundance)
			abundance[i] = ctx->T.iso_pref[i];
	}
	return DMS_OK;
}

int dms_point_create(dms_ctx* ctx, const dms_point_desc* desc, dms_point** out)
{
	if(ctx == nullptr || desc == nullptr || out == nullptr)
		return fail(ctx, DMS_ERR_ARG, "dms_point_create: null argument");
	*out = nullptr;
	if(ctx->device < 0)
		return fail(ctx, DMS_ERR_DEVICE, "dms_point_create: context has no GPU");
	dms_point* pt = new dms_point();
	pt->ctx = ctx;
	pt->desc = *desc;
	pt->desc.host_rate_grid = nullptr;
	const std::string msg = dms::build_point_params(ctx->T, *desc, pt->P);
	if(!msg.empty())
	{
		delete pt;
		const bool unsupported = desc->interaction == DMS_SD || desc->low_mass != 1;
		return fail(ctx, unsupported ? DMS_ERR_UNSUPPORTED : DMS_ERR_ARG, "dms_point_create: " + msg);
	}
	/* Interpolate_Total_DM_Scattering_Rate(DM, N_radius, N_speed): either count
	 * below 2 disables the grid (src/Solar_Model.cpp:397-406) */
	const int Nr = desc->interpolation_points;
	const int Nv = (desc->speed_points == 0) ? Nr : desc->speed_points;
	if(Nr > 20000 || Nv > 20000)
	{
		delete pt;
		return fail(ctx, DMS_ERR_ARG, "dms_point_create: interpolation_points or speed_points too large");
	}
	const bool grid = Nr >= 2 && Nv >= 2;
	pt->Nr = grid ? Nr : 0;
	pt->Nv = grid ? Nv : 0;
	pt->desc.interpolation_points = pt->Nr;
	pt->desc.speed_points = pt->Nv;
	std::lock_guard<std::mutex> lock(ctx->mu);
	hipSetDevice(ctx->device); /* callers may be on any host thread */
	/* Everything here is ordered on the context stream (stream-ordered
	 * allocations, asynchronous copies): no call waits for the null stream or
	 * the whole device, so points can be made while a resident serving launch
	 * runs on another stream. */
	const hipStream_t st = ctx->stream;
	auto bail = [&](const std::string& what) {
		dfree(pt->d_sigmaN, st);
		dfree(pt->d_grid, st);
		dfree(pt->d_grid32, st);
		hipStreamSynchronize(st);
		delete pt;
		return fail(ctx, DMS_ERR_DEVICE, "dms_point_create: " + what);
	};
	if(dalloc(&pt->d_sigmaN, DMS_MAX_TARGETS, st) != hipSuccess
	   || hipMemcpyAsync(pt->d_sigmaN, pt->P.sigma_N, sizeof(double) * DMS_MAX_TARGETS, hipMemcpyHostToDevice, st) != hipSuccess
	   || hipStreamSynchronize(st) != hipSuccess)
		return bail("sigma upload failed");
	pt->dev = host_point_view(pt);
	pt->dev.sigma_N = pt->d_sigmaN;
	if(grid)
	{
		const size_t cells = (size_t) pt->Nr * (size_t) pt->Nv;
		if(dalloc(&pt->d_grid, cells, st) != hipSuccess)
			return bail("grid allocation failed");
		if(desc->host_rate_grid != nullptr)
		{
			if(hipMemcpyAsync(pt->d_grid, desc->host_rate_grid, sizeof(double) * cells, hipMemcpyHostToDevice, st) != hipSuccess
			   || hipStreamSynchronize(st) != hipSuccess)
				return bail("grid upload failed");
		}
		else
		{
			(void) hipGetLastError();
			hipLaunchKernelGGL(dms_grid_kernel, dim3((unsigned) ((cells + 255) / 256)), dim3(256), 0, ctx->stream, pt->dev, ctx->dev,
			                   pt->Nr, pt->Nv, DMS_V_MAX, pt->d_grid);
			if(hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess)
				return bail("rate-grid kernel failed");
		}
		pt->dev.grid = pt->d_grid;
		/* FP32 copy (1/s) for the mixed-precision kernel */
		if(dalloc(&pt->d_grid32, cells, st) != hipSuccess)
			return bail("FP32 grid allocation failed");
		(void) hipGetLastError();
		hipLaunchKernelGGL(dms_grid_f32_kernel, dim3((unsigned) ((cells + 255) / 256)), dim3(256), 0, ctx->stream, pt->d_grid, cells,
		                   pt->d_grid32);
		if(hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess)
			return bail("FP32 grid conversion failed");
		pt->dev.grid32 = pt->d_grid32;
	}
	*out = pt;
	return DMS_OK;
}

void dms_point_destroy(dms_point* pt)
{
	if(pt == nullptr)
		return;
	/* stream-ordered frees: no wait for launches on other streams */
	hipSetDevice(pt->ctx->device);
	std::lock_guard<std::mutex> lock(pt->ctx->mu);
	dfree(pt->d_sigmaN, pt->ctx->stream);
	dfree(pt->d_grid, pt->ctx->stream);
	dfree(pt->d_grid32, pt->ctx->stream);
	hipStreamSynchronize(pt->ctx->stream);
	delete pt;
}

int dms_point_rate_grid(const dms_point* pt, double* out, size_t n)
{
	if(pt == nullptr || out == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_point_rate_grid: null argument");
	if(pt->Nr < 2)
		return fail(pt->ctx, DMS_ERR_STATE, "dms_point_rate_grid: point has no rate grid");
	const size_t cells = (size_t) pt->Nr * (size_t) pt->Nv;
	if(n < cells)
		return fail(pt->ctx, DMS_ERR_ARG, "dms_point_rate_grid: output too small");
	std::lock_guard<std::mutex> lock(pt->ctx->mu);
	hipSetDevice(pt->ctx->device);
	DMS_HIP(pt->ctx, hipMemcpyAsync(out, pt->d_grid, sizeof(double) * cells, hipMemcpyDeviceToHost, pt->ctx->stream));
	DMS_HIP(pt->ctx, hipStreamSynchronize(pt->ctx->stream));
	return DMS_OK;
}

int dms_point_grid_shape(const dms_point* pt, int32_t* n_radius, int32_t* n_speed)
{
	if(pt == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_point_grid_shape: null argument");
	if(n_radius)
		*n_radius = pt->Nr;
	if(n_speed)
		*n_speed = pt->Nv;
	return DMS_OK;
}

int dms_point_cross_sections(const dms_point* pt, double* sigma_p_cm2, double* sigma_e_cm2)
{
	if(pt == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_point_cross_sections: null argument");
	if(sigma_p_cm2)
		*sigma_p_cm2 = pt->P.sigma_p / (DMS_cm * DMS_cm);
	if(sigma_e_cm2)
		*sigma_e_cm2 = pt->P.sigma_e / (DMS_cm * DMS_cm);
	return DMS_OK;
}

int dms_point_rates(const dms_point* pt, int which, const double* r, const double* v, double* out, size_t n)
{
	if(pt == nullptr || (n > 0 && (r == nullptr || v == nullptr || out == nullptr)))
		return fail(nullptr, DMS_ERR_ARG, "dms_point_rates: null argument");
	if(which < 0 || which > 2 + pt->ctx->T.n_iso)
		return fail(pt->ctx, DMS_ERR_ARG, "dms_point_rates: unknown rate");
	if(n == 0)
		return DMS_OK;
	dms_ctx* ctx = pt->ctx;
	std::lock_guard<std::mutex> lock(ctx->mu);
	hipSetDevice(ctx->device); /* callers may be on any host thread */
	const hipStream_t st = ctx->stream;
	double *dr = nullptr, *dv = nullptr, *dout = nullptr;
	hipError_t e = dalloc(&dr, n, st);
	if(e == hipSuccess)
		e = dalloc(&dv, n, st);
	if(e == hipSuccess)
		e = dalloc(&dout, n, st);
	if(e == hipSuccess)
		e = hipMemcpyAsync(dr, r, sizeof(double) * n, hipMemcpyHostToDevice, st);
	if(e == hipSuccess)
		e = hipMemcpyAsync(dv, v, sizeof(double) * n, hipMemcpyHostToDevice, st);
	if(e == hipSuccess)
	{
		(void) hipGetLastError();
		hipLaunchKernelGGL(dms_rates_kernel, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, pt->dev, ctx->dev, which, dr, dv,
		                   dout, n);
		e = hipGetLastError();
	}
	if(e == hipSuccess)
		e = hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, st);
	dfree(dr, st);
	dfree(dv, st);
	dfree(dout, st);
	const hipError_t es = hipStreamSynchronize(st);
	if(e == hipSuccess)
		e = es;
	if(e != hipSuccess)
		return fail(ctx, DMS_ERR_DEVICE, std::string("dms_point_rates: ") + hipGetErrorString(e));
	return DMS_OK;
}

int dms_initial_conditions(dms_ctx* ctx, const dms_halo* halo, uint32_t seed, uint32_t point_id, uint64_t first_id, size_t n,
                           double r_final_rsun, dms_event* out, int32_t* shift_ok)
{
	if(ctx == nullptr || halo == nullptr || (n > 0 && out == nullptr))
		return fail(ctx, DMS_ERR_ARG, "dms_initial_conditions: null argument");
	if(ctx->device < 0)
		return fail(ctx, DMS_ERR_DEVICE, "dms_initial_conditions: context has no GPU");
	const std::string hm = check_halo(*halo);
	if(!hm.empty())
		return fail(ctx, DMS_ERR_ARG, "dms_initial_conditions: " + hm);
	if(n == 0)
		return DMS_OK;
	if(n > 0xffffffffu)
		return fail(ctx, DMS_ERR_ARG, "dms_initial_conditions: too many events");
	std::lock_guard<std::mutex> lock(ctx->mu);
	hipSetDevice(ctx->device); /* callers may be on any host thread */
	const DmsHaloView H = halo_view(ctx, *halo);
	const hipStream_t st = ctx->stream;
	dms_event* d_ev = nullptr;
	int32_t* d_ok = nullptr;
	hipError_t e = dalloc(&d_ev, n, st);
	if(e == hipSuccess)
		e = dalloc(&d_ok, n, st);
	const double R = (r_final_rsun > 0.0) ? r_final_rsun * DMS_rSun : 0.0;
	if(e == hipSuccess)
	{
		(void) hipGetLastError();
		hipLaunchKernelGGL(dms_ic_kernel, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, H, seed, point_id, first_id,
		                   (uint32_t) n, R, d_ev, d_ok);
		e = hipGetLastError();
	}
	if(e == hipSuccess)
		e = hipMemcpyAsync(out, d_ev, sizeof(dms_event) * n, hipMemcpyDeviceToHost, st);
	if(e == hipSuccess && shift_ok)
		e = hipMemcpyAsync(shift_ok, d_ok, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st);
	dfree(d_ev, st);
	dfree(d_ok, st);
	const hipError_t es = hipStreamSynchronize(st);
	if(e == hipSuccess)
		e = es;
	if(e != hipSuccess)
		return fail(ctx, DMS_ERR_DEVICE, std::string("dms_initial_conditions: ") + hipGetErrorString(e));
	return DMS_OK;
}

int dms_simulate_events(dms_point* pt, const dms_job_desc* job, const dms_event* initial, size_t n, dms_traj_record* records,
                        double* hist_rows)
{
	if(pt == nullptr || job == nullptr || (n > 0 && (initial == nullptr || records == nullptr)))
		return fail(pt ? pt->ctx : nullptr, DMS_ERR_ARG, "dms_simulate_events: null argument");
	dms_ctx* ctx = pt->ctx;
	const std::string jm = check_job(*job);
	if(!jm.empty())
		return fail(ctx, DMS_ERR_ARG, "dms_simulate_events: " + jm);
	if(n == 0)
		return DMS_OK;
	if(n > 0x7fffffffu)
		return fail(ctx, DMS_ERR_ARG, "dms_simulate_events: too many events");
	for(size_t i = 0; i < n; i++)
	{
		/* Simulate's argument check, src/Simulation_Trajectory.cpp:1278-1283 */
		const double r = std::sqrt(initial[i].r[0] * initial[i].r[0] + initial[i].r[1] * initial[i].r[1] + initial[i].r[2] * initial[i].r[2]);
		const double v = std::sqrt(initial[i].v[0] * initial[i].v[0] + initial[i].v[1] * initial[i].v[1] + initial[i].v[2] * initial[i].v[2]);
		if(!std::isfinite(initial[i].t) || initial[i].t < 0.0 || !std::isfinite(r) || r <= 0.0 || !std::isfinite(v) || v <= 0.0)
			return fail(ctx, DMS_ERR_ARG, "dms_simulate_events: initial state is invalid (Trajectory_Simulator::Simulate)");
	}
	std::lock_guard<std::mutex> lock(ctx->mu);
	hipSetDevice(ctx->device); /* callers may be on any host thread */
	const DmsJobView J = job_view(*job);
	const bool rows = hist_rows != nullptr && !J.terminate_on_capture;
	ChunkBuffers B;
	if(ensure(B, n, J.nbins, rows, true, ctx->stream) != hipSuccess)
	{
		B.release(ctx->stream);
		return fail(ctx, DMS_ERR_NOMEM, "dms_simulate_events: device allocation failed");
	}
	hipMemcpyAsync(B.ics, initial, sizeof(dms_event) * n, hipMemcpyHostToDevice, ctx->stream);
	hipMemsetAsync(B.ok, 0xff, sizeof(int32_t) * n, ctx->stream);
	float ms = 0.0f;
	hipError_t e = run_trajectories_single(ctx, ctx->stream, ctx->progress_stream, pt, J, B, n, rows, true, J.first_id, &ms);
	if(e == hipSuccess)
		e = hipMemcpyAsync(records, B.recs, sizeof(dms_traj_record) * n, hipMemcpyDeviceToHost, ctx->stream);
	if(e == hipSuccess && hist_rows)
	{
		if(rows)
			e = hipMemcpyAsync(hist_rows, B.rows, sizeof(double) * 2 * (size_t) J.nbins * n, hipMemcpyDeviceToHost, ctx->stream);
		else
			memset(hist_rows, 0, sizeof(double) * 2 * (size_t) J.nbins * n);
	}
	B.release(ctx->stream);
	hipStreamSynchronize(ctx->stream);
	if(e != hipSuccess)
		return fail(ctx, DMS_ERR_DEVICE, std::string("dms_simulate_events: ") + hipGetErrorString(e));
	return DMS_OK;
}

int dms_simulate_events_noalloc(dms_point* pt, const dms_job_desc* job, const dms_event* d_initial, size_t n,
                                dms_traj_record* d_records, double* d_hist_rows, void* stream)
{
	if(pt == nullptr || job == nullptr || (n > 0 && (d_initial == nullptr || d_records == nullptr)))
		return fail(pt ? pt->ctx : nullptr, DMS_ERR_ARG, "dms_simulate_events_noalloc: null argument");
	dms_ctx* ctx = pt->ctx;
	const std::string jm = check_job(*job);
	if(!jm.empty())
		return fail(ctx, DMS_ERR_ARG, "dms_simulate_events_noalloc: " + jm);
	if(n > DMS_NOALLOC_MAX_TRAJECTORIES)
		return fail(ctx, DMS_ERR_ARG,
		            "dms_simulate_events_noalloc: at most 2^31 - 1 trajectories per call (range of the scratch slot counter)");
	if(ctx->scratch_ptr == nullptr)
		return fail(ctx, DMS_ERR_DEVICE, "dms_simulate_events_noalloc: context has no GPU");
	if(n == 0)
		return DMS_OK;
	std::lock_guard<std::mutex> lock(ctx->mu);
	hipSetDevice(ctx->device); /* callers may be on any host thread */
	const hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
	const DmsJobView J = job_view(*job);
	const bool rows = d_hist_rows != nullptr && !J.terminate_on_capture;
	unsigned char* s = ctx->scratch_ptr;
	(void) hipGetLastError();
	hipLaunchKernelGGL(dms_noalloc_setup_kernel, dim3(1), dim3(64), 0, st, s, (uint32_t) n, J.first_id);
	if(d_hist_rows)
	{
		/* per-trajectory rows start at zero (capture mode leaves them zero) */
		const size_t cells = 2 * (size_t) J.nbins * n;
		const size_t blocks = std::min<size_t>((cells + 255) / 256, (size_t) 65536);
		hipLaunchKernelGGL(dms_zero_kernel, dim3((unsigned) blocks), dim3(256), 0, st, d_hist_rows, cells);
	}
	DmsTrajArgs a;
	a.S = ctx->dev;
	a.P = pt->dev;
	a.J = J;
	a.ics = d_initial;
	a.shift_ok = nullptr; /* every event is a valid start */
	a.n = (uint32_t) n;
	a.records = d_records;
	a.summ = nullptr;
	a.R.rows = rows ? d_hist_rows : nullptr;
	a.R.counts = nullptr;
	a.R.bins = nullptr;
	a.R.touched = nullptr;
	a.R.per_lane = 0;
	a.R.hist = nullptr;
	a.R.hcounts = nullptr;
	a.next = reinterpret_cast<unsigned int*>(s + kScratchNext);
	a.refill_batch = 1;
	a.lone_serial = ctx->knobs.lone_serial;
	a.n_jobs = 1;
	a.job_start = reinterpret_cast<const uint32_t*>(s + kScratchStart);
	a.job_first_id = reinterpret_cast<const uint64_t*>(s + kScratchFirst);
	a.job_done = nullptr;
	a.job_points = nullptr;
	a.job_keys = nullptr;
	a.fill_ctl = nullptr;
	a.cap_stop = 0;
	a.fill_extra = 0;
	a.cancel = nullptr;
	hipLaunchKernelGGL(J.mixed ? dms_traj_kernel_mixed : dms_traj_kernel, dim3(traj_grid(ctx, n)), dim3(DMS_TRAJ_BLOCK), ctx->traj_lds,
	                   st, a);
	hipError_t e = hipGetLastError();
	if(e == hipSuccess && stream == nullptr)
		e = hipStreamSynchronize(st);
	if(e != hipSuccess)
		return fail(ctx, DMS_ERR_DEVICE, std::string("dms_simulate_events_noalloc: ") + hipGetErrorString(e));
	return DMS_OK;
}

namespace
{
std::string check_submit(const dms_job_desc& d)
{
	std::string m = check_job(d);
	if(m.empty())
		m = check_halo(d.halo);
	if(m.empty() && d.max_trajectories == 0 && d.capture_target == 0)
		m = "a job needs max_trajectories or capture_target";
	return m;
}

void start_batch(dms_point* const* pts, const dms_job_desc* descs, size_t n, dms_job** out)
{
	auto batch = std::make_shared<dms_batch>();
	for(size_t k = 0; k < n; k++)
	{
		dms_job* job = new dms_job();
		job->pt = pts[k];
		job->desc = descs[k];
		job->batch = batch;
		memset(&job->C, 0, sizeof(job->C));
		batch->jobs.push_back(job);
		out[k] = job;
	}
	batch->worker = std::thread(run_batch_worker, batch);
}

void join_batch(dms_job* job)
{
	if(job->server != nullptr)
	{
		std::unique_lock<std::mutex> lock(job->fmu);
		job->fcv.wait(lock, [job] { return job->finished.load() != 0; });
		return;
	}
	dms_batch* b = job->batch.get();
	if(b == nullptr)
		return;
	std::lock_guard<std::mutex> lock(b->mu);
	if(!b->joined && b->worker.joinable())
		b->worker.join();
	b->joined = true;
}
} // namespace

int dms_job_submit(dms_point* pt, const dms_job_desc* desc, dms_job** out)
{
	if(pt == nullptr || desc == nullptr || out == nullptr)
		return fail(pt ? pt->ctx : nullptr, DMS_ERR_ARG, "dms_job_submit: null argument");
	*out = nullptr;
	const std::string m = check_submit(*desc);
	if(!m.empty())
		return fail(pt->ctx, DMS_ERR_ARG, "dms_job_submit: " + m);
	start_batch(&pt, desc, 1, out);
	return DMS_OK;
}

int dms_job_submit_batch(dms_point* pt, const dms_job_desc* descs, size_t n, dms_job** out)
{
	if(pt == nullptr || descs == nullptr || out == nullptr || n == 0)
		return fail(pt ? pt->ctx : nullptr, DMS_ERR_ARG, "dms_job_submit_batch: null argument or empty batch");
	for(size_t k = 0; k < n; k++)
		out[k] = nullptr;
	uint64_t total = 0;
	for(size_t k = 0; k < n; k++)
	{
		const std::string m = check_submit(descs[k]);
		if(!m.empty())
			return fail(pt->ctx, DMS_ERR_ARG, "dms_job_submit_batch: job " + std::to_string(k) + ": " + m);
		if(n > 1 && !batchable(descs[k]))
			return fail(pt->ctx, DMS_ERR_ARG,
			            "dms_job_submit_batch: job " + std::to_string(k)
			                + ": batched jobs need a fixed trajectory count (max_trajectories > 0, capture_target 0) and at most "
			                + std::to_string(DMS_BATCH_MAX_SLOTS) + " trajectories and 65536 bins");
		if(n > 1 && !same_batch_config(descs[0], descs[k]))
			return fail(pt->ctx, DMS_ERR_ARG,
			            "dms_job_submit_batch: job " + std::to_string(k)
			                + " differs from job 0 in a field other than max_trajectories, first_trajectory_id or record_trajectories");
		total += descs[k].max_trajectories;
	}
	if(n > 1 && total > DMS_BATCH_MAX_SLOTS)
		return fail(pt->ctx, DMS_ERR_ARG, "dms_job_submit_batch: the batch holds more than "
		                                      + std::to_string(DMS_BATCH_MAX_SLOTS) + " trajectories");
	std::vector<dms_point*> pts(n, pt);
	start_batch(pts.data(), descs, n, out);
	return DMS_OK;
}

int dms_job_submit_multi(dms_point* const* points, const dms_job_desc* descs, size_t n, dms_job** out)
{
	if(points == nullptr || descs == nullptr || out == nullptr || n == 0 || points[0] == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_job_submit_multi: null argument or empty batch");
	dms_ctx* ctx = points[0]->ctx;
	for(size_t k = 0; k < n; k++)
		out[k] = nullptr;
	uint64_t total = 0;
	for(size_t k = 0; k < n; k++)
	{
		if(points[k] == nullptr || points[k]->ctx != ctx)
			return fail(ctx, DMS_ERR_ARG, "dms_job_submit_multi: job " + std::to_string(k) + ": point is null or of another context");
		const std::string m = check_submit(descs[k]);
		if(!m.empty())
			return fail(ctx, DMS_ERR_ARG, "dms_job_submit_multi: job " + std::to_string(k) + ": " + m);
		if(!batchable(descs[k]))
			return fail(ctx, DMS_ERR_ARG,
			            "dms_job_submit_multi: job " + std::to_string(k)
			                + ": batched jobs need a fixed trajectory count (max_trajectories > 0, capture_target 0) and at most "
			                + std::to_string(DMS_BATCH_MAX_SLOTS) + " trajectories and 65536 bins");
		if(!same_launch_config(descs[0], descs[k]))
			return fail(ctx, DMS_ERR_ARG,
			            "dms_job_submit_multi: job " + std::to_string(k)
			                + " differs from job 0 in mode, precision, limits, r_max or n_bins");
		total += descs[k].max_trajectories;
	}
	if(total > DMS_BATCH_MAX_SLOTS)
		return fail(ctx, DMS_ERR_ARG, "dms_job_submit_multi: the batch holds more than " + std::to_string(DMS_BATCH_MAX_SLOTS) + " trajectories");
	start_batch(points, descs, n, out);
	return DMS_OK;
}

namespace
{
/* The autotuners' calibration launch: the same points, keys and limits,
 * trajectory ids past each job's own range, an RKF45-attempt cap that bounds
 * the launch.  Returns each job's lane iterations per trajectory and the
 * launch's kernel time. */
int calibration_launch(dms_point* const* points, const dms_job_desc* descs, size_t n, uint64_t calibration_trajectories,
                       uint64_t calibration_rk_attempts, const char* who, std::vector<double>* cost, double* kernel_ms)
{
	dms_ctx* ctx = points[0]->ctx;
	if(calibration_trajectories == 0 || calibration_trajectories * n > DMS_BATCH_MAX_SLOTS || calibration_rk_attempts == 0)
		return fail(ctx, DMS_ERR_ARG, std::string(who) + ": calibration size must be positive and fit one launch");
	std::vector<dms_job_desc> cal(descs, descs + n);
	for(size_t k = 0; k < n; k++)
	{
		cal[k].max_trajectories = calibration_trajectories;
		cal[k].capture_target = 0;
		cal[k].record_trajectories = 0;
		cal[k].first_trajectory_id = descs[k].first_trajectory_id + descs[k].max_trajectories;
		if(cal[k].rk_attempt_budget == 0 || cal[k].rk_attempt_budget > calibration_rk_attempts)
			cal[k].rk_attempt_budget = calibration_rk_attempts;
	}
	std::vector<dms_job*> jobs(n, nullptr);
	const int st = dms_job_submit_multi(points, cal.data(), n, jobs.data());
	if(st != DMS_OK)
		return st;
	if(cost)
		cost->assign(n, 0.0);
	int status = DMS_OK;
	for(size_t k = 0; k < n; k++)
	{
		if(dms_job_wait(jobs[k], nullptr) != DMS_OK)
			status = DMS_ERR_DEVICE;
		else
		{
			const dms_counters& c = jobs[k]->C;
			/* lane iterations: one per RKF45 attempt, about four per scattering */
			if(cost)
				(*cost)[k] =
				    c.trajectories ? ((double) c.n_rk_attempts + 4.0 * (double) c.total_scatterings) / (double) c.trajectories : 0.0;
			if(kernel_ms && k == 0)
				*kernel_ms = jobs[k]->kernel_ms;
		}
		dms_job_destroy(jobs[k]);
	}
	if(status != DMS_OK)
		return fail(ctx, status, std::string(who) + ": calibration launch failed");
	return DMS_OK;
}
} // namespace

int dms_multi_autotune(dms_point* const* points, const dms_job_desc* descs, size_t n, uint64_t calibration_trajectories,
                       uint64_t calibration_rk_attempts, uint32_t* order, double* cost_per_trajectory)
{
	if(points == nullptr || descs == nullptr || order == nullptr || n == 0 || points[0] == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_multi_autotune: null argument or empty batch");
	std::vector<double> cost;
	const int st = calibration_launch(points, descs, n, calibration_trajectories, calibration_rk_attempts, "dms_multi_autotune",
	                                  &cost, nullptr);
	if(st != DMS_OK)
		return st;
	/* cheapest first: the long trajectories of the costly points start last and
	 * finish in waves the short ones have left, instead of sharing full waves */
	std::vector<uint32_t> idx(n);
	for(size_t k = 0; k < n; k++)
		idx[k] = (uint32_t) k;
	std::stable_sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) { return cost[x] < cost[y]; });
	for(size_t k = 0; k < n; k++)
	{
		order[k] = idx[k];
		if(cost_per_trajectory)
			cost_per_trajectory[k] = cost[k];
	}
	return DMS_OK;
}

int dms_set_launch_knobs(dms_ctx* ctx, const dms_launch_knobs* knobs)
{
	if(ctx == nullptr || knobs == nullptr)
		return fail(ctx, DMS_ERR_ARG, "dms_set_launch_knobs: null argument");
	if(knobs->refill_batch < 1 || knobs->refill_batch > 64)
		return fail(ctx, DMS_ERR_ARG, "dms_set_launch_knobs: refill_batch must be in [1, 64]");
	if(knobs->lone_serial > 1)
		return fail(ctx, DMS_ERR_ARG, "dms_set_launch_knobs: lone_serial must be 0 or 1");
	ctx->knobs = *knobs;
	return DMS_OK;
}

int dms_get_launch_knobs(const dms_ctx* ctx, dms_launch_knobs* knobs)
{
	if(ctx == nullptr || knobs == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_get_launch_knobs: null argument");
	*knobs = ctx->knobs;
	return DMS_OK;
}

int dms_autotune_knobs(dms_point* const* points, const dms_job_desc* descs, size_t n, uint64_t calibration_trajectories,
                       uint64_t calibration_rk_attempts, const uint32_t* refill_candidates, size_t n_candidates, int repeats,
                       double* seconds, dms_launch_knobs* best, int apply)
{
	if(points == nullptr || descs == nullptr || n == 0 || points[0] == nullptr || refill_candidates == nullptr || n_candidates == 0
	   || best == nullptr || repeats < 1)
		return fail(nullptr, DMS_ERR_ARG, "dms_autotune_knobs: null argument, no candidates or repeats < 1");
	dms_ctx* ctx = points[0]->ctx;
	for(size_t i = 0; i < n_candidates; i++)
		if(refill_candidates[i] < 1 || refill_candidates[i] > 64)
			return fail(ctx, DMS_ERR_ARG, "dms_autotune_knobs: refill_batch candidates must be in [1, 64]");
	const dms_launch_knobs saved = ctx->knobs;
	std::vector<double> t(n_candidates, std::numeric_limits<double>::infinity());
	/* candidates in turn, repeated: drifts of the device hit them alike */
	for(int r = 0; r < repeats; r++)
		for(size_t i = 0; i < n_candidates; i++)
		{
			ctx->knobs.refill_batch = refill_candidates[i];
			double ms = 0.0;
			const int st = calibration_launch(points, descs, n, calibration_trajectories, calibration_rk_attempts,
			                                  "dms_autotune_knobs", nullptr, &ms);
			if(st != DMS_OK)
			{
				ctx->knobs = saved;
				return st;
			}
			t[i] = std::min(t[i], ms * 1e-3);
		}
	/* the first candidate is the reference; another replaces it only when it is
	 * more than 10% faster.  The calibration launch (short trajectories, each
	 * cut at the calibration budget) favours batched starts more than a full
	 * launch does: on the 65-point workload it ranked 16 up to 5% ahead of 1
	 * (mixed precision), while the full launch took 13% longer with 16 (FP64:
	 * 2.14, 2.16 and 2.34 s for 1, 4 and 16) */
	size_t b = 0;
	for(size_t i = 1; i < n_candidates; i++)
		if(t[i] < 0.90 * t[0] && t[i] < t[b])
			b = i;
	best->refill_batch = refill_candidates[b];
	best->lone_serial = saved.lone_serial;
	if(seconds)
		for(size_t i = 0; i < n_candidates; i++)
			seconds[i] = t[i];
	ctx->knobs = apply ? *best : saved;
	return DMS_OK;
}

int dms_server_create(dms_ctx* ctx, const dms_server_desc* desc, dms_server** out)
{
	if(ctx == nullptr || desc == nullptr || out == nullptr)
		return fail(ctx, DMS_ERR_ARG, "dms_server_create: null argument");
	*out = nullptr;
	if(ctx->device < 0)
		return fail(ctx, DMS_ERR_STATE, "dms_server_create: context has no device");
	const std::string m = check_job(desc->launch);
	if(!m.empty())
		return fail(ctx, DMS_ERR_ARG, "dms_server_create: " + m);
	const uint32_t ring = desc->slot_ring ? desc->slot_ring : (1u << 20);
	if(ring < 1024 || ring > (1u << 26) || (ring & (ring - 1)) != 0)
		return fail(ctx, DMS_ERR_ARG, "dms_server_create: slot_ring must be a power of two in [1024, 2^26]");
	const uint32_t max_reqs = desc->max_requests ? desc->max_requests : 1024;
	if(max_reqs > 65536)
		return fail(ctx, DMS_ERR_ARG, "dms_server_create: max_requests must be at most 65536");
	dms_server* S = new dms_server();
	S->ctx = ctx;
	S->cfg = desc->launch;
	S->J = job_view(desc->launch);
	S->want_rows = desc->launch.mode != DMS_MODE_CAPTURE;
	S->nb = S->J.nbins;
	if(S->want_rows && S->nb > 65536)
	{
		delete S;
		return fail(ctx, DMS_ERR_ARG, "dms_server_create: at most 65536 bins");
	}
	S->ring = ring;
	S->mask = ring - 1;
	S->max_reqs = max_reqs;
	uint32_t max_rounds = 0;
	if(S->want_rows)
	{
		const size_t per_round = (size_t) 80 * (size_t) S->nb;
		max_rounds = desc->max_rounds ? desc->max_rounds
		                              : (uint32_t) std::min<size_t>(4096, std::max<size_t>(64, ((size_t) 256 << 20) / per_round));
		if(max_rounds > 65536)
		{
			delete S;
			return fail(ctx, DMS_ERR_ARG, "dms_server_create: max_rounds must be at most 65536");
		}
	}
	S->max_rounds = max_rounds;
	S->grid = traj_grid(ctx, (size_t) 1 << 40);
	S->lanes = (size_t) S->grid * DMS_TRAJ_BLOCK;
	hipSetDevice(ctx->device);
	const size_t nb = (size_t) S->nb;
	const size_t R = max_rounds;
	bool ok = host_alloc(&S->h_ctl, 16) == hipSuccess && host_alloc(&S->h_slots, ring) == hipSuccess
	          && host_alloc(&S->h_reqs, max_reqs) == hipSuccess && host_alloc(&S->h_results, ring) == hipSuccess
	          && host_alloc(&S->h_cq, ring) == hipSuccess && host_alloc(&S->h_round_size, R) == hipSuccess
	          && host_alloc(&S->h_round_gen, R) == hipSuccess && host_alloc(&S->h_round_stamp, R) == hipSuccess;
	if(ok && S->want_rows)
		ok = host_alloc(&S->h_mb_hist, R * 8 * nb) == hipSuccess && host_alloc(&S->h_mb_counts, R * 2 * nb) == hipSuccess;
	/* device buffers: stream-ordered on the context stream, so that making a
	 * server never waits for launches on other streams */
	const hipStream_t cs = ctx->stream;
	std::unique_lock<std::mutex> ctx_lock(ctx->mu);
	if(ok)
		ok = dalloc(&S->d_counters, DMS_SERVE_COUNTERS, cs) == hipSuccess && dalloc(&S->d_round_done, R, cs) == hipSuccess
		     && hipMemsetAsync(S->d_round_done, 0, sizeof(uint32_t) * (R ? R : 1), cs) == hipSuccess;
	if(ok && S->want_rows)
		ok = dalloc(&S->d_round_hist, R * 8 * nb, cs) == hipSuccess && dalloc(&S->d_round_counts, R * 2 * nb, cs) == hipSuccess
		     && dalloc(&S->d_bins, S->lanes * nb, cs) == hipSuccess && dalloc(&S->d_touched, S->lanes * nb, cs) == hipSuccess
		     && hipMemsetAsync(S->d_round_hist, 0, sizeof(double) * R * 8 * nb, cs) == hipSuccess
		     && hipMemsetAsync(S->d_round_counts, 0, sizeof(unsigned long long) * R * 2 * nb, cs) == hipSuccess
		     && hipMemsetAsync(S->d_bins, 0, sizeof(DmsBinAcc) * S->lanes * nb, cs) == hipSuccess;
	if(ok)
		ok = hipEventCreateWithFlags(&S->ended, hipEventDisableTiming) == hipSuccess && hipStreamSynchronize(cs) == hipSuccess;
	ctx_lock.unlock();
	if(!ok)
	{
		serve_release(S);
		delete S;
		return fail(ctx, DMS_ERR_NOMEM, "dms_server_create: host or device allocation failed");
	}
	S->pos_req.assign(ring, nullptr);
	S->pos_off.assign(ring, 0);
	S->rounds.assign(R, ServeRoundBuf());
	for(uint32_t b = 0; b < max_rounds; b++)
		S->free_rounds.push_back(max_rounds - 1 - b);
	for(uint32_t q = 0; q < max_reqs; q++)
		S->free_reqs.push_back(max_reqs - 1 - q);
	S->driver = std::thread(serve_driver, S);
	*out = S;
	return DMS_OK;
}

int dms_server_submit(dms_server* S, dms_point* pt, const dms_job_desc* desc, dms_job** out)
{
	if(S == nullptr || pt == nullptr || desc == nullptr || out == nullptr)
		return fail(S ? S->ctx : nullptr, DMS_ERR_ARG, "dms_server_submit: null argument");
	*out = nullptr;
	if(pt->ctx != S->ctx)
		return fail(S->ctx, DMS_ERR_ARG, "dms_server_submit: point of another context");
	const std::string m = check_submit(*desc);
	if(!m.empty())
		return fail(S->ctx, DMS_ERR_ARG, "dms_server_submit: " + m);
	if(!same_launch_config(S->cfg, *desc))
		return fail(S->ctx, DMS_ERR_ARG,
		            "dms_server_submit: mode, precision, limits, r_max or n_bins differ from the server's launch descriptor");
	if(desc->record_trajectories)
		return fail(S->ctx, DMS_ERR_UNSUPPORTED, "dms_server_submit: served requests do not keep trajectory records");
	dms_job* job = new dms_job();
	job->pt = pt;
	job->desc = *desc;
	job->server = S;
	memset(&job->C, 0, sizeof(job->C));
	{
		std::lock_guard<std::mutex> lock(S->mu);
		if(S->closing)
		{
			delete job;
			return fail(S->ctx, DMS_ERR_STATE, "dms_server_submit: server is being destroyed");
		}
		S->pending.push_back(job);
	}
	S->cv.notify_one();
	*out = job;
	return DMS_OK;
}

int dms_server_get_stats(const dms_server* S, dms_server_stats* st)
{
	if(S == nullptr || st == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_server_get_stats: null argument");
	memset(st, 0, sizeof(*st));
	st->launches = S->n_launches.load();
	st->slots_published = S->n_published.load();
	st->slots_cancelled = S->n_cancelled.load();
	st->requests_done = S->n_done.load();
	st->requests_active = (uint32_t) S->n_active.load();
	st->running = (uint32_t) S->is_running.load();
	return DMS_OK;
}

void dms_server_destroy(dms_server* S)
{
	if(S == nullptr)
		return;
	{
		std::lock_guard<std::mutex> lock(S->mu);
		S->closing = true;
	}
	S->cv.notify_one();
	if(S->driver.joinable())
		S->driver.join();
	hipSetDevice(S->ctx->device);
	serve_release(S);
	delete S;
}

int dms_job_cancel(dms_job* job)
{
	if(job == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_job_cancel: null job");
	if(job->finished.load())
		return DMS_OK;
	job->cancel_requested.store(1);
	if(job->server != nullptr)
		return DMS_OK; /* the serving driver stops the request */
	dms_batch* b = job->batch.get();
	if(b == nullptr)
		return DMS_OK;
	b->cancelled.store(1);
	const int w = b->cancel_word.load();
	if(w >= 0)
		__atomic_store_n(job->pt->ctx->h_cancel + w, 1u, __ATOMIC_SEQ_CST);
	return DMS_OK;
}

int dms_job_poll(dms_job* job, dms_job_status* st)
{
	if(job == nullptr || st == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_job_poll: null argument");
	st->trajectories_done = job->done_traj.load();
	st->captured = job->done_captured.load();
	st->scatterings = job->done_scat.load();
	st->completed_outward_escapes = job->done_escaped.load();
	st->numerical_failures = job->done_failed.load();
	st->done = job->finished.load();
	st->pad = 0;
	return DMS_OK;
}

int dms_job_wait(dms_job* job, dms_job_status* st)
{
	if(job == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_job_wait: null argument");
	join_batch(job);
	if(st)
		dms_job_poll(job, st);
	if(job->status != DMS_OK)
		return fail(job->pt->ctx, job->status, "job failed: " + job->err);
	return DMS_OK;
}

int dms_job_read(dms_job* job, dms_counters* counters, double* hist, uint64_t* counts, size_t n_bins, dms_evap_event* events,
                 size_t event_capacity, size_t* n_events, dms_traj_record* records, size_t record_capacity, size_t* n_records)
{
	if(job == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_job_read: null job");
	if(!job->finished.load())
		return fail(job->pt->ctx, DMS_ERR_STATE, "dms_job_read: job has not finished (call dms_job_wait)");
	if(job->status != DMS_OK)
		return fail(job->pt->ctx, job->status, "dms_job_read: job failed: " + job->err);
	if(counters)
		*counters = job->C;
	const size_t nb = (size_t) job_view(job->desc).nbins;
	if(hist || counts)
	{
		if(n_bins != nb)
			return fail(job->pt->ctx, DMS_ERR_ARG, "dms_job_read: n_bins does not match the job");
		if(hist)
			memcpy(hist, job->hist.data(), sizeof(double) * 8 * nb);
		if(counts)
			for(size_t i = 0; i < 2 * nb; i++)
				counts[i] = job->counts[i];
	}
	if(n_events)
		*n_events = job->events.size();
	if(events)
		for(size_t i = 0; i < job->events.size() && i < event_capacity; i++)
			events[i] = job->events[i];
	if(n_records)
		*n_records = job->records.size();
	if(records)
		for(size_t i = 0; i < job->records.size() && i < record_capacity; i++)
			records[i] = job->records[i];
	return DMS_OK;
}

int dms_job_timing(const dms_job* job, double* trajectory_kernel_ms, double* launches, double* other_ms)
{
	if(job == nullptr)
		return fail(nullptr, DMS_ERR_ARG, "dms_job_timing: null job");
	if(trajectory_kernel_ms)
		*trajectory_kernel_ms = job->kernel_ms;
	if(launches)
		*launches = job->launches;
	if(other_ms)
		*other_ms = job->other_ms;
	return DMS_OK;
}

void dms_job_destroy(dms_job* job)
{
	if(job == nullptr)
		return;
	join_batch(job);
	delete job;
}

int dms_probe_math(dms_ctx* ctx, int fn, const double* x, const double* y, double* out, size_t n)
{
	if(ctx == nullptr || (n > 0 && (x == nullptr || out == nullptr)) || fn < 0 || fn > 9)
		return fail(ctx, DMS_ERR_ARG, "dms_probe_math: bad argument");
	if(ctx->device < 0)
		return fail(ctx, DMS_ERR_DEVICE, "dms_probe_math: context has no GPU");
	if(n == 0)
		return DMS_OK;
	std::lock_guard<std::mutex> lock(ctx->mu);
	hipSetDevice(ctx->device); /* callers may be on any host thread */
	double *dx = nullptr, *dy = nullptr, *dout = nullptr;
	DMS_HIP(ctx, hipMalloc((void**) &dx, sizeof(double) * n));
	DMS_HIP(ctx, hipMalloc((void**) &dy, sizeof(double) * n));
	DMS_HIP(ctx, hipMalloc((void**) &dout, sizeof(double) * n));
	hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice);
	hipMemcpy(dy, y ? y : x, sizeof(double) * n, hipMemcpyHostToDevice);
	(void) hipGetLastError();
	hipLaunchKernelGGL(dms_probe_kernel, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, ctx->stream, fn, dx, dy, dout, n);
	hipError_t e = hipGetLastError();
	if(e == hipSuccess)
		e = hipStreamSynchronize(ctx->stream);
	if(e == hipSuccess)
		e = hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost);
	hipFree(dx);
	hipFree(dy);
	hipFree(dout);
	if(e != hipSuccess)
		return fail(ctx, DMS_ERR_DEVICE, std::string("dms_probe_math: ") + hipGetErrorString(e));
	return DMS_OK;
}

} /* extern "C" */
