/*
 * gfx950 kernels of the solar dark-matter trajectory path (FP64).
 *
 *   dms_traj_kernel   persistent, one lane per trajectory (FP64; the mixed-precision
 *                     instantiation is dms_traj_kernel_mixed), lane refill from an
 *                     atomic work counter; Sample_Target's 63 nuclear + 1
 *                     electron rates are evaluated by the 64 lanes of the wave
 *                     together (one target per lane) for each scattering lane.
 *   dms_ic_kernel     Initial_Conditions + Hyperbolic_Kepler_Shift, one lane each.
 *   dms_grid_kernel   Interpolate_Total_DM_Scattering_Rate, one node per lane.
 *   dms_rates_kernel  point queries of the rate functions.
 *   dms_hist_*        per-bin sums over trajectory rows, fixed (deterministic) order.
 *   dms_probe_kernel  elementary functions on the device (for host comparison).
 */
#include <hip/hip_runtime.h>

#include "dms_engine.h"
#include "dms_halo.h"
#include "dms_kernels.hpp"

namespace
{

__device__ inline double readlane_d(double x, int l)
{
	const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
	const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
	return __hiloint2double(hi, lo);
}

/* lane l's value (l uniform over the wave): one v_readlane, no LDS round trip */
__device__ inline float readlane_f(float x, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l)); }

} // namespace

/* Live progress of a job (the heartbeat's per-rank counts): finished
 * trajectories, captured ones, completed outward escapes of uncaptured ones
 * and numerical failures, as the run accounting classifies them
 * (src/Data_Generation.cpp:764-850). */
__device__ inline void dms_count_finished(unsigned int* w, uint32_t flags, int reason, int capture_mode)
{
	atomicAdd(w, 1u);
	if(flags & DMS_F_CAPTURED)
		atomicAdd(w + 1, 1u);
	else if(reason == DMS_OUTWARD_ESCAPE)
		atomicAdd(w + 2, 1u);
	const bool numerical = reason == DMS_NUMERICAL_FAILURE || reason == DMS_NON_FINITE_STATE || reason == DMS_SPEED_LIMIT
	                       || reason == DMS_ENERGY_DRIFT_ESCAPE || reason == DMS_UNKNOWN
	                       || (!capture_mode && !(flags & DMS_F_CAPTURED) && (flags & DMS_F_REFLECTION_SHIFT_FAILED));
	if(numerical)
		atomicAdd(w + 3, 1u);
}

/* ---- serving launch ------------------------------------------------------ */

/* A structure in host memory read with system-scope loads (past every device
 * cache: the host rewrites request and slot entries while the launch runs). */
template <typename T>
__device__ __forceinline__ T dms_sys_load(const T* p)
{
	static_assert(sizeof(T) % 8 == 0, "8-byte words");
	unsigned long long w[sizeof(T) / 8];
	const unsigned long long* q = reinterpret_cast<const unsigned long long*>(p);
#pragma unroll
	for(int i = 0; i < (int) (sizeof(T) / 8); i++)
		w[i] = __hip_atomic_load(q + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
	T out;
	__builtin_memcpy(&out, w, sizeof(T));
	return out;
}

/* lane l's copy of a per-lane structure, broadcast to the wave */
template <typename T>
__device__ __forceinline__ T dms_readlane_struct(const T& v, int l)
{
	static_assert(sizeof(T) % 4 == 0, "4-byte words");
	int w[sizeof(T) / 4];
	__builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
	for(int i = 0; i < (int) (sizeof(T) / 4); i++)
		w[i] = __builtin_amdgcn_readlane(w[i], l);
	T out;
	__builtin_memcpy(&out, w, sizeof(T));
	return out;
}

/* constant-clock ticks (100 MHz) between reads of the host's published count */
#define DMS_SERVE_POLL_TICKS 200ull

struct DmsServePoll
{
	uint32_t tick, idle;
};

/* Hand-back of finished slots (lanes with `completing` set; the wave calls it
 * together).  The histogram rows of these trajectories are already in their
 * round's sums.  Order: row atomics, then the round count (agent fence); the
 * summary, then the completion-queue entry (system fence).  The slot that
 * completes a round copies the round's sums to its host mailbox, clears them
 * for the buffer's next round and stamps the mailbox. */
__device__ __forceinline__ void dms_serve_complete(const DmsTrajArgs& a, const DmsLane& L, uint32_t round, const DmsSummary& sm,
                                                   bool completing, int lane)
{
	if(__ballot(completing) == 0)
		return;
	const DmsServeView& V = a.V;
	__threadfence();
	bool closer = false;
	if(completing)
	{
		V.results[L.slot] = sm;
		if(round != DMS_SERVE_NO_ROUND)
		{
			const uint32_t size = __hip_atomic_load(V.round_size + round, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
			closer = atomicAdd(V.round_done + round, 1u) + 1u == size;
		}
	}
	__threadfence_system();
	if(completing)
	{
		const unsigned long long t = atomicAdd(V.counters + 1, 1ull);
		__hip_atomic_store(V.cq + (uint32_t) (t & V.ring_mask), L.slot + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
	}
	unsigned long long rm = __ballot(closer);
	while(rm)
	{
		const int s = __ffsll((long long) rm) - 1;
		rm &= rm - 1;
		const uint32_t b = (uint32_t) __builtin_amdgcn_readlane((int) round, s);
		const int nb = a.J.nbins;
		/* every slot of the round counted itself after its row atomics */
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
		unsigned long long* dh = reinterpret_cast<unsigned long long*>(a.R.hist + (size_t) b * 8 * (size_t) nb);
		unsigned long long* mh = reinterpret_cast<unsigned long long*>(V.mb_hist + (size_t) b * 8 * (size_t) nb);
		for(int i = lane; i < 8 * nb; i += 64)
		{
			mh[i] = __hip_atomic_load(dh + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			dh[i] = 0ull;
		}
		unsigned long long* dc = a.R.hcounts + (size_t) b * 2 * (size_t) nb;
		unsigned long long* mc = V.mb_counts + (size_t) b * 2 * (size_t) nb;
		for(int i = lane; i < 2 * nb; i += 64)
		{
			mc[i] = __hip_atomic_load(dc + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			dc[i] = 0ull;
		}
		if(lane == 0)
			__hip_atomic_store(V.round_done + b, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		__threadfence_system();
		if(lane == 0)
		{
			const unsigned long long g = __hip_atomic_load(V.round_gen + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
			__hip_atomic_store(V.round_stamp + b, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
		}
	}
}

/* Lane refill of a serving launch.  Lanes that need work take slot indices
 * from the counter at once and hold them (PH_WAIT) until the host has
 * published them; a published slot starts its trajectory with initial
 * conditions drawn in-lane (the computation of dms_ic_kernel: same Philox
 * stream, same bits).  A slot of a cancelled request, or one whose initial
 * shift fails, is handed back at once (`completing`, summary in sm).  After
 * the stop flag, lanes holding unpublished slots end.  A wave with nothing
 * but waiting lanes backs off between reads of the host's published count. */
template <bool MX>
__device__ __forceinline__ void dms_serve_refill(const DmsTrajArgs& a, DmsLaneCtx& c, DmsLane& L, uint32_t& round, DmsSummary& sm,
                                                 bool& completing, DmsServePoll& poll, int lane)
{
	const DmsServeView& V = a.V;
	const unsigned long long nmask = __ballot(L.phase == PH_NEED);
	if(nmask != 0)
	{
		const int leader = __ffsll((long long) nmask) - 1;
		unsigned long long base = 0;
		if(lane == leader)
			base = atomicAdd(V.counters, (unsigned long long) __popcll(nmask));
		base = ((unsigned long long) (uint32_t) __builtin_amdgcn_readlane((int) (base >> 32), leader) << 32)
		       | (uint32_t) __builtin_amdgcn_readlane((int) (uint32_t) base, leader);
		if(L.phase == PH_NEED)
		{
			L.id = V.base + base + (unsigned long long) __popcll(nmask & ((1ull << lane) - 1ull));
			L.phase = PH_WAIT;
		}
	}
	const unsigned long long wmask = __ballot(L.phase == PH_WAIT);
	if(wmask == 0)
	{
		poll.idle = 0;
		return;
	}
	const bool all_wait = __ballot((L.phase != PH_WAIT) & (L.phase != PH_DONE)) == 0;
	const int leader = __ffsll((long long) wmask) - 1;
	unsigned long long pub = 0, stop = 0;
	if(lane == leader)
	{
		/* The published count and the stop flag come from their device-memory
		 * mirror; at most one wave of the launch reads the host's copies per
		 * DMS_SERVE_POLL_TICKS of the constant clock (the wave that claims the
		 * read time), because reads of host memory from thousands of waiting
		 * waves slow the memory traffic of the running lanes.  The slot and
		 * request entries are read only after the count has arrived (their
		 * loads depend on it), with system-scope loads that do not hit in the
		 * device caches; the host writes them before it raises the count. */
		const unsigned long long now = wall_clock64();
		unsigned long long last = __hip_atomic_load(V.counters + 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if(now - last >= DMS_SERVE_POLL_TICKS
		   && __hip_atomic_compare_exchange_strong(V.counters + 4, &last, now, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
		                                           __HIP_MEMORY_SCOPE_AGENT))
		{
			pub = __hip_atomic_load(V.ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
			stop = __hip_atomic_load(V.ctl + 8, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
			atomicMax(V.counters + 2, pub);
			if(stop)
				__hip_atomic_store(V.counters + 3, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		}
		else
		{
			pub = __hip_atomic_load(V.counters + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			stop = __hip_atomic_load(V.counters + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		}
	}
	pub = ((unsigned long long) (uint32_t) __builtin_amdgcn_readlane((int) (pub >> 32), leader) << 32)
	      | (uint32_t) __builtin_amdgcn_readlane((int) (uint32_t) pub, leader);
	stop = (unsigned long long) (uint32_t) __builtin_amdgcn_readlane((int) (uint32_t) stop, leader);
	bool started = false;
	if(L.phase == PH_WAIT)
	{
		if(L.id < pub)
		{
			const uint32_t pos = (uint32_t) L.id & V.ring_mask;
			const DmsServeSlot d = dms_sys_load(V.slots + pos);
			const DmsServeReq* q = V.reqs + d.req;
			L.slot = pos;
			L.row = (uint32_t) (blockIdx.x * blockDim.x + threadIdx.x);
			L.job = d.req;
			L.id = d.id;
			round = d.round;
			started = true;
			if(__hip_atomic_load(&q->cancel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u)
			{
				sm.reason = DMS_SERVE_CANCELLED;
				sm.flags = 0;
				sm.n_scat = 0;
				sm.t_capture_sec = -1.0;
				sm.t_unbind_sec = -1.0;
				sm.n_rk = 0;
				sm.n_int = sm.n_ext = sm.n_rej = sm.n_tv = 0;
				L.phase = PH_NEED;
				completing = true;
			}
			else
			{
				c.P = dms_sys_load(&q->P);
				const DmsHaloView H = dms_sys_load(&q->H);
				const unsigned long long keys = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(&q->seed),
				                                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
				c.J.seed = (uint32_t) keys;
				c.J.point_id = (uint32_t) (keys >> 32);
				dm_rng rng;
				dm_rng_init(&rng, c.J.seed, c.J.point_id, L.id, DMS_STREAM_IC);
				dms_event e;
				dms_initial_condition(H, &rng, &e);
				const bool shift_ok = dms_kepler_shift(e.r, e.v, c.J.max_distance);
				dms_lane_start(L, c, e, shift_ok, nullptr, &sm);
				if(L.phase == PH_NEED) /* initial shift failed: recorded in sm */
					completing = true;
			}
		}
		else if(stop)
			L.phase = PH_DONE;
	}
	if(all_wait && __ballot(started) == 0)
	{
		/* nothing published for this wave yet: back off, up to ~50 us between
		 * reads of the mirror */
		if(poll.idle < 16)
			poll.idle++;
		for(uint32_t k = 0; k < poll.idle; k++)
			__builtin_amdgcn_s_sleep(127);
	}
	else
		poll.idle = 0;
}

/* ---- two-wave split of a lone trajectory ----------------------------------- */

__device__ __forceinline__ unsigned int dms_lds_load(const unsigned int* p)
{
	return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void dms_lds_store(unsigned int* p, unsigned int v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

/* Spin limits (each poll one LDS round trip, no sleep: the trajectory's waits
 * below sit on its critical path): a trajectory's wave that waits this long
 * for a helper's step gives the split up and goes on alone. */
#define DMS_PIPE_WAIT_POLLS (1u << 23)
/* polls for the second helper's rates of an entry before the post computes
 * its own */
#define DMS_PIPE_RATE_POLLS 512u

/* A helper wave serving slot sl (its lanes 0 and 1 compute, with the same
 * values except inside the RKF45 attempts, whose stages they evaluate in
 * pairs; lane 0 writes the ring): for each session the trajectory's wave
 * opens, it prepares the following steps, assuming each post continues, and
 * writes them to the ring.  A session that changes is noticed between the
 * preparation and every RKF45 attempt and between an attempt's stage pairs.
 * Returns when the trajectory's wave has finished all its lanes (want = 0). */
template <bool MX, bool MP>
__device__ __forceinline__ void dms_pipe_serve(const DmsTrajArgs& a, const DmsSolarView& S, DmsPipeSlot* sl, int lane)
{
	if(lane > 1)
		return;
	const bool second = lane == 1;
	DmsLane A;
	DmsStepEvent ev;
	DmsPointView P = a.P;
	uint32_t job = 0;
	unsigned int cur = 0, head = dms_lds_load(&sl->head);
	bool ended = true, redo_first = false;
	for(;;)
	{
		if(dms_lds_load(&sl->want) == 0u)
			return;
		const unsigned int s = dms_lds_load(&sl->session);
		if(s != cur)
		{
			cur = s;
			ended = (s == 0u);
			if(!ended)
			{
				DmsPipeSeed seed = sl->seed;
				/* the seed was read after the session word: if the session
				 * moved meanwhile, the next pass starts over (the fence keeps
				 * the seed's loads before the second look) */
				__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
				if(dms_lds_load(&sl->session) != s)
				{
					cur = 0u;
					continue;
				}
				dms_pipe_from_seed(A, seed);
				job = seed.job;
				if constexpr(MP)
					P = a.job_points[job];
				/* a redo session's first step starts where the rejected one did */
				redo_first = seed.redo != 0u;
				ev = redo_first ? seed.eb : seed.ev;
			}
		}
		if(ended)
		{
			__builtin_amdgcn_s_sleep(2);
			continue;
		}
		DmsPipeEntry e;
		e.eb = ev;
		if(!redo_first)
			dms_pipe_assume_next(A, ev);
		redo_first = false;
		e.r_before = A.r_before;
		e.job = job;
		if(dms_step_prep<MX>(A, P, S, a.J) >= 0)
		{
			e.kind = DMS_PIPE_PREP_END;
			ended = true;
		}
		else
		{
			int r = -1;
			bool moved = false;
			do
			{
				if(__hip_atomic_load(&sl->session, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != cur)
				{
					moved = true;
					break;
				}
				A.inner_iter++;
				A.rk_all++;
				/* an attempt is abandoned between its stage pairs once the
				 * session has moved */
				r = MX ? dms_rk45_attempt_mx<true>(A.prop, S, A.inner_iter, second, &sl->session, cur)
				       : dms_rk45_attempt<true>(A.prop, S, A.inner_iter, second, &sl->session, cur);
				if(r == -2)
				{
					moved = true;
					break;
				}
			} while(r < 0);
			if(moved)
				continue;
			dms_pipe_fill_entry<MX>(A, r, e);
			ev = e.ev;
		}
		e.session = cur;
		/* wait for a free ring entry */
		bool stale = false;
		while(head - dms_lds_load(&sl->tail) >= (unsigned int) DMS_PIPE_DEPTH)
		{
			if((dms_lds_load(&sl->session) != cur) | (dms_lds_load(&sl->want) == 0u))
			{
				stale = true;
				break;
			}
			/* the helpers' idle polls sleep in the FP64 kernel and spin in the
			 * mixed one: measured on lone trajectories, FP64 -5% per attempt
			 * with the trajectory's wave alone spinning, mixed -3% more with
			 * all three waves spinning */
			if constexpr(!MX)
				__builtin_amdgcn_s_sleep(1);
		}
		if(stale || dms_lds_load(&sl->session) != cur)
			continue;
		if(!second)
		{
			sl->e[head % DMS_PIPE_DEPTH] = e;
			dms_lds_store(&sl->head, head + 1);
		}
		head++;
	}
}

/* A second helper wave serving slot sl: for every entry the first helper
 * writes, in order, the post's rate queries at the optical-depth fractions and
 * the end point, one query per lane (dms_pipe_rate: the same evaluations
 * dms_rates_at_fractions makes one after the other; mixed precision
 * dms_pipe_rate_mx, the FP32 rates widened); then rhead passes the entry.
 * Entries of any session are evaluated (a stale one costs a little time,
 * nothing else).  Returns when the trajectory's wave has finished all its
 * lanes. */
template <bool MX, bool MP>
__device__ __forceinline__ void dms_pipe_rates(const DmsTrajArgs& a, const DmsSolarView& S, DmsPipeSlot* sl, int lane)
{
	constexpr int NQ = DMS_MAX_OPTICAL_DEPTH_PIECES + 2;
	unsigned int b = __builtin_amdgcn_readfirstlane(dms_lds_load(&sl->head));
	/* entries before b were written before this wave came: their have_rates
	 * stays 0 (the first helper writes it so), so their posts compute their own */
	dms_lds_store(&sl->rhead, b);
	for(;;)
	{
		if(__builtin_amdgcn_readfirstlane(dms_lds_load(&sl->want)) == 0u)
			return;
		if(__builtin_amdgcn_readfirstlane(dms_lds_load(&sl->head)) == b)
		{
			/* as the first helper's idle polls */
			if constexpr(!MX)
				__builtin_amdgcn_s_sleep(1);
			continue;
		}
		DmsPipeEntry* ep = &sl->e[b % DMS_PIPE_DEPTH];
		bool have = false;
		if(__builtin_amdgcn_readfirstlane(ep->kind) == DMS_PIPE_STEP)
		{
			DmsPointView P;
			if constexpr(MP)
				P = a.job_points[ep->job];
			else
				P = a.P;
			if constexpr(MX)
			{
				DmsPipeRateInputsMx q;
				have = dms_pipe_rate_inputs_mx(*ep, a.J, q);
				if(have & (lane < NQ))
					ep->rt[lane] = dms_pipe_rate_mx(*ep, q, P, S, lane);
			}
			else
			{
				DmsPipeRateInputs q;
				have = dms_pipe_rate_inputs(*ep, a.J, q);
				if(have & (lane < NQ))
					ep->rt[lane] = dms_pipe_rate(*ep, q, P, S, lane);
			}
		}
		if(lane == 0)
			ep->have_rates = have ? 1 : 0;
		b++;
		/* the rates and the flag before the progress word (every lane's stores) */
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
		dms_lds_store(&sl->rhead, b);
	}
}

/* A wave whose lanes have all finished: serve lone trajectories of the other
 * waves of the block until every wave of the block has finished. */
template <bool MX, bool MP>
__device__ __forceinline__ void dms_pipe_idle(const DmsTrajArgs& a, const DmsSolarView& S, DmsPipeBlock* pb, int wave, int lane)
{
	const int nwaves = (int) (blockDim.x >> 6);
	for(;;)
	{
		if(dms_lds_load(&pb->waves_done) >= (unsigned int) nwaves)
			return;
		/* first or second helper of a waiting wave */
		int claimed = -1, role = 0;
		if(lane == 0)
			for(int w = 0; w < nwaves && claimed < 0; w++)
			{
				DmsPipeSlot* sl = &pb->slot[w];
				if(w == wave || dms_lds_load(&sl->want) == 0u)
					continue;
				for(int r = 0; r < 2; r++)
				{
					unsigned int* word = r ? &sl->helper2 : &sl->helper;
					unsigned int expect = 0u;
					if(dms_lds_load(word) == 0u
					   && __hip_atomic_compare_exchange_strong(word, &expect, (unsigned int) (wave + 1), __ATOMIC_ACQ_REL,
					                                           __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP))
					{
						claimed = w;
						role = r;
						break;
					}
				}
			}
		claimed = __builtin_amdgcn_readfirstlane(claimed);
		role = __builtin_amdgcn_readfirstlane(role);
		if(claimed >= 0)
		{
			if(role == 0)
				dms_pipe_serve<MX, MP>(a, S, &pb->slot[claimed], lane);
			else
				dms_pipe_rates<MX, MP>(a, S, &pb->slot[claimed], lane);
		}
		else
			__builtin_amdgcn_s_sleep(8);
	}
}

template <bool MX, bool MP, bool SV = false>
__device__ __forceinline__ void dms_traj_body(const DmsTrajArgs& a)
{
	const int lane = threadIdx.x & 63;
	/* Stage the profile nodes and the Mass and escape-speed cubics in LDS: the
	 * RKF45 stages read Mass() six times per attempt in a dependent chain, so
	 * their latency sets the speed of a lone long trajectory. */
	extern __shared__ __attribute__((aligned(32))) unsigned char dms_lds[];
	const int nn = a.S.n;
	double* lx = reinterpret_cast<double*>(dms_lds);
	DmsCubic* lmass = reinterpret_cast<DmsCubic*>(dms_lds + dms_lds_mass_offset(nn));
	DmsCubic* lvesc = lmass + (nn - 1);
	for(int i = threadIdx.x; i < nn; i += blockDim.x)
		lx[i] = a.S.x[i];
	for(int i = threadIdx.x; i < nn - 1; i += blockDim.x)
	{
		lmass[i] = a.S.mass[i];
		lvesc[i] = a.S.vesc2[i];
	}
	/* the two-wave split: a lone trajectory's wave and a finished wave of the
	 * block prepare its steps together (not in a serving launch, whose lanes
	 * wait for published slots instead of finishing) */
	constexpr bool PIPE = !SV;
	DmsPipeBlock* pb = reinterpret_cast<DmsPipeBlock*>(dms_lds + dms_lds_pipe_offset(nn));
	const int wave = (int) (threadIdx.x >> 6);
	DmsPipeSlot* my = &pb->slot[wave];
	if constexpr(PIPE)
	{
		unsigned int* w = reinterpret_cast<unsigned int*>(pb);
		for(int i = threadIdx.x; i < (int) (sizeof(DmsPipeBlock) / sizeof(unsigned int)); i += blockDim.x)
			w[i] = 0u;
	}
	__syncthreads();
	DmsLaneCtx c;
	c.S = a.S;
	c.S.x = lx;
	c.S.mass = lmass;
	c.S.vesc2 = lvesc;
	c.P = a.P;
	c.J = a.J;
	c.R = a.R;
	DmsLane L;
	L.phase = PH_NEED;
	L.cur_bin = -1;
	L.slot = 0;
	L.row = 0;
	L.job = 0;
	L.id = 0;
	L.n_touched = 0;
	L.sc_target = -2;
	const uint32_t lane_gid = blockIdx.x * blockDim.x + threadIdx.x;
	const int n_iso = a.S.n_iso;
	/* serving launch: the slot's round buffer, the summary of a trajectory
	 * being handed back, and the wave's polling state */
	uint32_t sv_round = DMS_SERVE_NO_ROUND;
	DmsSummary sv_sum;
	DmsServePoll sv_poll = {0u, 0u};
	uint32_t cancel_tick = 0;
	/* two-wave split state (wave-uniform): lone = one live lane, the others
	 * finished; attached = a helper serves this wave; session = the open
	 * session (0: none); tail = ring entries read */
	bool lone = false, attached = false, held = false, seeded = false, held_rates = false;
	bool split_ok = PIPE && a.lone_serial == 0u;
	unsigned int session = 0u, sessions = 0u, tail = 0u;
	for(;;)
	{
		dms_lane_advance_rk<MX>(L, c);
		if constexpr(PIPE)
		{
			if(__builtin_expect(lone & !attached & split_ok, 0))
				attached = __builtin_amdgcn_readfirstlane(dms_lds_load(&my->helper)) != 0u;
			/* the lone lane accepted an attempt with no session open: the helper
			 * prepares the steps after this one from here */
			if(__builtin_expect(attached & (session == 0u), 0) && __ballot(L.phase == PH_POST) != 0)
			{
				session = ++sessions;
				if(L.phase == PH_POST)
				{
					/* the seed's stores stay after the previous session's end
					 * (session 0) and before this session's number: a helper
					 * that reads the seed and then still sees the same session
					 * read one whole seed */
					__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
					dms_pipe_make_seed<MX>(L, L.job, my->seed);
					dms_lds_store(&my->session, session);
				}
				seeded = true;
			}
		}
		int post_ret;
		/* a step taken from the helper: its entry (with the end event the post
		 * uses) stays in the ring until the post is done; a session's first
		 * post takes the event its seed holds */
		dms_lane_advance_post<MX>(L, c, &post_ret,
		                          held ? &my->e[tail % DMS_PIPE_DEPTH].ev : (seeded ? &my->seed.ev : nullptr),
		                          held_rates ? my->e[tail % DMS_PIPE_DEPTH].rt : nullptr);
		seeded = false;
		if constexpr(PIPE)
		{
			if(__builtin_expect(held, 0))
			{
				held = held_rates = false;
				tail++;
				if(lane == 0)
					dms_lds_store(&my->tail, tail);
			}
			/* the session lasts while the posts continue to the next step */
			if(__builtin_expect(session != 0u, 0)
			   && __ballot((L.phase == PH_PREP) & (post_ret == DMS_POST_NEXT)) == 0)
			{
				session = 0u;
				dms_lds_store(&my->session, 0u);
			}
			/* a step rejected on optical depth: the helper prepares it again
			 * (and the steps after it) from a redo seed, while this wave waits
			 * for the entry instead of running the preparation and the RKF45
			 * attempts itself */
			if(__builtin_expect(attached & (session == 0u), 0)
			   && __ballot((L.phase == PH_PREP) & dms_post_rejected(L, post_ret)) != 0)
			{
				session = ++sessions;
				if(L.phase == PH_PREP)
				{
					/* the seed's stores stay after the previous session's end,
					 * as for any seed */
					__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
					dms_pipe_make_redo_seed(L, L.job, my->seed);
					dms_lds_store(&my->session, session);
				}
			}
		}
		/* The wave-level services below (target sampling, scatter completion,
		 * capture counting, trajectory end, row flush, refill, start) act only
		 * on lanes in those phases: while every live lane is between RKF45
		 * attempts, as a lone long trajectory is, the wave skips them at once. */
		bool service;
		if constexpr(SV)
		{
			/* lanes waiting for published slots look for them on every 16th
			 * pass only, so that a lone trajectory beside them keeps the fast
			 * path */
			const bool busy = __ballot((L.phase != PH_DONE) & (L.phase != PH_PREP) & (L.phase != PH_RK) & (L.phase != PH_WAIT)) != 0;
			service = busy || (__ballot(L.phase == PH_WAIT) != 0 && ((++sv_poll.tick) & 15u) == 0u);
		}
		else
			service = __ballot((L.phase != PH_DONE) & (L.phase != PH_PREP) & (L.phase != PH_RK)) != 0;
		if(service)
		{
			/* Sample_Target (src/Simulation_Trajectory.cpp:1122-1158): for each
			 * scattering lane s the 64 lanes evaluate one target rate each. */
			/* (the rare branches of the loop are marked so that the compiler lays
			 * them out of the path a lone long trajectory executes) */
			unsigned long long m = __ballot(L.phase == PH_SCAT);
			while(__builtin_expect(m != 0, 0))
			{
				const int s = __ffsll((long long) m) - 1;
				m &= m - 1;
				const double r = readlane_d(L.sc_r, s);
				const double v = readlane_d(L.sc_v, s);
				const double xi = readlane_d(L.sc_xi, s);
				const int k = (lane < n_iso) ? lane : ((lane == DMS_MAX_TARGETS - 1) ? lane : -1);
				int target = -2;
				/* the scattering lane's parameter point (uniform over the wave):
				 * a serving launch's lanes hold their points in registers */
				DmsPointView Pv;
				if constexpr(SV)
					Pv = dms_readlane_struct(c.P, s);
				const DmsPointView& Ps = SV ? Pv : (MP ? a.job_points[__builtin_amdgcn_readlane((int) L.job, s)] : a.P);
				if constexpr(MX)
				{
					/* FP32 rates (1/s), same draw and order as below */
					float rate = 0.0f;
					if(k >= 0)
						rate = dms_rate_target_mx(Ps, c.S, r, v, k);
					const bool bad = (k >= 0) && (!mx_isfinite(rate) || rate < 0.0f);
					if(__ballot(bad) == 0)
					{
						/* lanes n_iso..62 hold +0 (no target): the fixed trip
						 * count adds them without changing the sum, and the
						 * constant lane indices let the readlanes run ahead of
						 * the dependent adds */
						float total = readlane_f(rate, DMS_MAX_TARGETS - 1);
#pragma unroll
						for(int i = 0; i < DMS_MAX_TARGETS - 1; i++)
							total += readlane_f(rate, i);
						if(mx_isfinite(total) && total > 0.0f)
						{
							const float q = rate * mx_rcp(total);
							const float xf = (float) xi;
							float sum = readlane_f(q, DMS_MAX_TARGETS - 1);
							if(sum > xf)
								target = -1;
							else
								for(int i = 0; i < n_iso; i++)
								{
									sum += readlane_f(q, i);
									if(sum > xf || i == n_iso - 1)
									{
										target = i;
										break;
									}
								}
						}
					}
					if(lane == s)
						L.sc_target = target;
					continue;
				}
				double rate = 0.0;
				if(k >= 0)
					rate = dms_rate_target(Ps, c.S, r, v, k);
				const bool bad = (k >= 0) && (!dm_isfinite(rate) || rate < 0.0);
				if(__ballot(bad) == 0)
				{
					/* same summation and comparison order as dms_choose_target */
					/* lanes n_iso..62 hold +0 (no target): adding them changes
					 * no bit of a sum of non-negative rates (a -0 start only
					 * turns into +0, which fails the test below as -0 does),
					 * and the fixed, unrolled trip count lets the readlanes
					 * run ahead of the dependent adds */
					const double re = readlane_d(rate, DMS_MAX_TARGETS - 1);
					double total = re;
#pragma unroll
					for(int i = 0; i < DMS_MAX_TARGETS - 1; i++)
						total = total + readlane_d(rate, i);
					if(dm_isfinite(total) && total > 0.0)
					{
						const double q = rate / total;
						double sum = readlane_d(q, DMS_MAX_TARGETS - 1);
						if(sum > xi)
							target = -1;
						else
							for(int i = 0; i < n_iso; i++)
							{
								sum += readlane_d(q, i);
								if(sum > xi || i == n_iso - 1)
								{
									target = i;
									break;
								}
							}
					}
				}
				if(lane == s)
					L.sc_target = target;
			}
			const int captured_before = L.captured;
			dms_lane_finish_scatter<MX>(L, c);
			/* capture-target chunk: count captures as they happen (the post-scatter
			 * test), so that the slot limit drops while the captured trajectories
			 * still orbit */
			if(a.fill_ctl && L.captured && !captured_before)
			{
				const unsigned int ncap = atomicAdd(a.fill_ctl, 1u) + 1u;
				if(ncap == a.cap_stop)
				{
					const unsigned int handed = __hip_atomic_load(a.next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
					atomicMin(a.fill_ctl + 1, handed + a.fill_extra);
				}
			}
			int flush_cls = -1;
			const bool finishing = L.phase == PH_FIN;
			if(__builtin_expect(finishing, 0))
			{
				uint32_t fl = 0;
				int reason = 0;
				if constexpr(SV)
				{
					const dms_event ic0 = {};
					flush_cls = dms_lane_finish_rows_pending(L, c, ic0, nullptr, &sv_sum);
				}
				else
				{
					flush_cls = dms_lane_finish_rows_pending(L, c, a.ics[L.slot], a.records ? a.records + L.slot : nullptr,
					                                         a.summ ? a.summ + L.slot : nullptr, &fl, &reason);
					if(a.job_done)
						dms_count_finished(a.job_done + DMS_PROGRESS_WORDS * L.job, fl, reason, c.J.terminate_on_capture);
				}
			}
			/* per-lane histogram rows of the finished trajectories: the wave adds
			 * each row's touched bins to its job's sums, 64 bins at a time */
			unsigned long long fm = __ballot(flush_cls >= 0);
			if(__builtin_expect(fm != 0, 0))
			{
				while(fm)
				{
					const int s = __ffsll((long long) fm) - 1;
					fm &= fm - 1;
					const uint32_t row_s = (uint32_t) __builtin_amdgcn_readlane((int) L.row, s);
					const uint32_t job_s = (uint32_t) __builtin_amdgcn_readlane((int) (SV ? sv_round : L.job), s);
					const uint32_t n_s = (uint32_t) __builtin_amdgcn_readlane((int) L.n_touched, s);
					const int cls_s = __builtin_amdgcn_readlane(flush_cls, s);
					dms_flush_rows(c.R, c.J.nbins, row_s, job_s, n_s, cls_s, (uint32_t) lane, 64u);
				}
				if(flush_cls >= 0)
					L.n_touched = 0;
				/* the cleared rows are reused by their lanes' next trajectories */
				__threadfence_block();
			}
			if constexpr(SV)
			{
				dms_serve_complete(a, L, sv_round, sv_sum, finishing, lane);
				bool handed_back = false;
				dms_serve_refill<MX>(a, c, L, sv_round, sv_sum, handed_back, sv_poll, lane);
				dms_serve_complete(a, L, sv_round, sv_sum, handed_back, lane);
			}
			else
			{
			/* lane refill: one atomic per wave, once refill_batch lanes are free
			 * (or no lane of the wave is left running), so that trajectory starts
			 * run together instead of one lane at a time */
			const unsigned long long nmask = __ballot(L.phase == PH_NEED);
			bool refill = nmask != 0;
			if(__builtin_expect(refill, 0) && (uint32_t) __popcll(nmask) < a.refill_batch
			   && __ballot((L.phase != PH_NEED) & (L.phase != PH_DONE)) != 0)
			{
				/* wait for more free lanes unless every slot is handed out */
				unsigned int handed = 0;
				if(lane == __ffsll((long long) nmask) - 1)
					handed = __hip_atomic_load(a.next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				refill = (unsigned int) __builtin_amdgcn_readlane((int) handed, __ffsll((long long) nmask) - 1) >= a.n;
			}
			if(__builtin_expect(refill, 0))
			{
				const int leader = __ffsll((long long) nmask) - 1;
				unsigned int base = 0, limit = a.n;
				if(lane == leader)
				{
					base = atomicAdd(a.next, (unsigned int) __popcll(nmask));
					if(a.fill_ctl)
						limit = __hip_atomic_load(a.fill_ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				}
				base = (unsigned int) __builtin_amdgcn_readlane((int) base, leader);
				limit = (unsigned int) __builtin_amdgcn_readlane((int) limit, leader);
				if(L.phase == PH_NEED)
				{
					const unsigned int rank = (unsigned int) __popcll(nmask & ((1ull << lane) - 1ull));
					const unsigned int idx = base + rank;
					if(idx < a.n && idx < limit)
					{
						L.slot = idx;
						/* job of the slot: last j with job_start[j] <= idx */
						uint32_t lo = 0, hi = a.n_jobs;
						while(hi - lo > 1)
						{
							const uint32_t mid = (lo + hi) >> 1;
							if(a.job_start[mid] <= idx)
								lo = mid;
							else
								hi = mid;
						}
						const uint32_t j = lo;
						L.job = j;
						if constexpr(MP)
						{
							c.P = a.job_points[j];
							c.J.seed = a.job_keys[2 * j];
							c.J.point_id = a.job_keys[2 * j + 1];
						}
						L.id = a.job_first_id[j] + (uint64_t) (idx - a.job_start[j]);
						L.row = a.R.per_lane ? lane_gid : idx;
						L.phase = PH_START;
					}
					else
						L.phase = PH_DONE;
				}
			}
			if(__builtin_expect(L.phase == PH_START, 0))
			{
				dms_lane_start(L, c, a.ics[L.slot], a.shift_ok == nullptr || a.shift_ok[L.slot] != 0,
				               a.records ? a.records + L.slot : nullptr, a.summ ? a.summ + L.slot : nullptr);
				if(L.phase == PH_NEED && a.job_done) /* initial shift failed: recorded, done */
					dms_count_finished(a.job_done + DMS_PROGRESS_WORDS * L.job, DMS_F_SHIFT_FAILED, DMS_NUMERICAL_FAILURE,
					                   c.J.terminate_on_capture);
			}
			} /* not a serving launch */
			if constexpr(PIPE)
			{
				/* one live lane left and every other lane finished for good (no
				 * slot left): ask for a helper wave */
				if(split_ok & !lone && __popcll(__ballot(L.phase != PH_DONE)) == 1)
				{
					lone = true;
					if(lane == 0)
						dms_lds_store(&my->want, 1u);
				}
			}
		}
		if constexpr(PIPE)
		{
			/* the next step of the open session: the helper's preparation and
			 * attempts take the place of this lane's (the post runs next) */
			if(__builtin_expect(session != 0u, 0))
			{
				bool got = false, timed_out = true;
				for(unsigned int polls = 0; polls < DMS_PIPE_WAIT_POLLS; polls++)
				{
					const unsigned int h = __builtin_amdgcn_readfirstlane(dms_lds_load(&my->head));
					if(h == tail)
						continue;
					const DmsPipeEntry* ep = &my->e[tail % DMS_PIPE_DEPTH];
					const unsigned int es = __builtin_amdgcn_readfirstlane(ep->session);
					const int kind = __builtin_amdgcn_readfirstlane(ep->kind);
					got = (es == session) & (kind == DMS_PIPE_STEP);
					if(got)
					{
						/* held until the post has run (the ring entry is not
						 * reused before tail passes it) */
						if(L.phase == PH_PREP)
						{
							dms_pipe_take(L, *ep);
							L.phase = PH_POST;
						}
						held = true;
						timed_out = false;
						/* the entry's rate queries from a second helper, once
						 * its progress word has passed the entry */
						held_rates = false;
						if(__builtin_amdgcn_readfirstlane(dms_lds_load(&my->helper2)) != 0u)
							for(unsigned int rp = 0; rp < DMS_PIPE_RATE_POLLS; rp++)
							{
								const unsigned int rh = __builtin_amdgcn_readfirstlane(dms_lds_load(&my->rhead));
								if((int) (rh - tail) > 0)
								{
									held_rates = __builtin_amdgcn_readfirstlane(ep->have_rates) != 0;
									break;
								}
							}
						break;
					}
					tail++;
					if(lane == 0)
						dms_lds_store(&my->tail, tail);
					if(es == session)
					{
						/* a preparation that ends the segment: this lane runs its
						 * own, which ends it the same way */
						timed_out = false;
						break;
					}
				}
				if(!got)
				{
					session = 0u;
					dms_lds_store(&my->session, 0u);
					/* no answer from the helper: go on alone for good */
					if(__builtin_expect(timed_out, 0))
						split_ok = attached = false;
				}
			}
		}
		dms_lane_advance_b<MX>(L, c);
		if constexpr(!SV)
		{
			/* cancellation: one read of the host's flag per wave every 1024
			 * iterations; every lane of the launch then ends where it is */
			if(__builtin_expect(a.cancel != nullptr, 0) && ((++cancel_tick) & 1023u) == 0u)
			{
				unsigned int cv = 0;
				if(lane == 0)
					cv = __hip_atomic_load(a.cancel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
				if(__builtin_amdgcn_readlane((int) cv, 0) != 0)
					L.phase = PH_DONE;
			}
		}
		if(__ballot(L.phase != PH_DONE) == 0)
			break;
	}
	if constexpr(PIPE)
	{
		/* this wave's lanes have all finished: release a helper, then help the
		 * block's other lone trajectories until every wave has finished */
		if(lane == 0)
		{
			dms_lds_store(&my->want, 0u);
			__hip_atomic_fetch_add(&pb->waves_done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
		}
		if(a.lone_serial == 0u)
			dms_pipe_idle<MX, MP>(a, c.S, pb, wave, lane);
	}
}

__global__ void __launch_bounds__(DMS_TRAJ_BLOCK) dms_traj_kernel(DmsTrajArgs a) { dms_traj_body<false, false>(a); }

__global__ void __launch_bounds__(DMS_TRAJ_BLOCK) dms_traj_kernel_mixed(DmsTrajArgs a) { dms_traj_body<true, false>(a); }

__global__ void __launch_bounds__(DMS_TRAJ_BLOCK) dms_traj_kernel_multi(DmsTrajArgs a) { dms_traj_body<false, true>(a); }

__global__ void __launch_bounds__(DMS_TRAJ_BLOCK) dms_traj_kernel_mixed_multi(DmsTrajArgs a) { dms_traj_body<true, true>(a); }

__global__ void __launch_bounds__(DMS_TRAJ_BLOCK) dms_traj_kernel_serve(DmsTrajArgs a) { dms_traj_body<false, true, true>(a); }

__global__ void __launch_bounds__(DMS_TRAJ_BLOCK) dms_traj_kernel_mixed_serve(DmsTrajArgs a) { dms_traj_body<true, true, true>(a); }

/* The kernels below run beside resident trajectory launches (a point's rate
 * table and initial conditions are built while another job's or the serving
 * launch's persistent kernel holds every compute unit).  That kernel leaves
 * 80 of a SIMD's 512 registers per lane free (256 VGPRs + its accumulation
 * registers), so a wave of these kernels fits beside it only with at most 80;
 * eight waves per SIMD caps them at 64. */
#define DMS_BESIDE_RESIDENT __attribute__((amdgpu_waves_per_eu(8)))

/* FP32 copy of the rate table in 1/s for the mixed-precision kernel */
__global__ void DMS_BESIDE_RESIDENT dms_grid_f32_kernel(const double* grid, size_t n, float* out)
{
	const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if(i < n)
		out[i] = (float) (grid[i] * DMS_sec);
}

__global__ void DMS_BESIDE_RESIDENT dms_ic_kernel(DmsHaloView H, uint32_t seed, uint32_t point_id, uint64_t first_id, uint32_t n, double R_final,
                              dms_event* out, int32_t* ok)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n)
		return;
	dm_rng rng;
	dm_rng_init(&rng, seed, point_id, first_id + i, DMS_STREAM_IC);
	dms_event e;
	dms_initial_condition(H, &rng, &e);
	int s = 1;
	if(R_final > 0.0)
		s = dms_kepler_shift(e.r, e.v, R_final) ? 1 : 0;
	out[i] = e;
	ok[i] = s;
}

__global__ void DMS_BESIDE_RESIDENT dms_grid_kernel(DmsPointView P, DmsSolarView S, int Nr, int Nv, double vmax, double* grid)
{
	const size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if(idx >= (size_t) Nr * (size_t) Nv)
		return;
	const int i = (int) (idx / (size_t) Nv);
	const int j = (int) (idx % (size_t) Nv);
	const double dv = (vmax - 0.0) / (double) (Nv - 1);
	const double radius = DMS_rSun * (double) i / (double) (Nr - 1);
	const double speed = 0.0 + j * dv;
	grid[idx] = dms_rate_computed(P, S, radius, speed);
}

__global__ void DMS_BESIDE_RESIDENT dms_rates_kernel(DmsPointView P, DmsSolarView S, int which, const double* r, const double* v, double* out, size_t n)
{
	const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n)
		return;
	double res;
	if(which == 0)
		res = dms_rate_total(P, S, r[i], v[i]);
	else if(which == 1)
		res = dms_rate_computed(P, S, r[i], v[i]);
	else if(which == 2)
		res = dms_rate_target(P, S, r[i], v[i], DMS_MAX_TARGETS - 1);
	else
		res = dms_rate_target(P, S, r[i], v[i], which - 3);
	out[i] = res;
}

/* Partial per-bin sums over trajectories [blk*TB, min(n, (blk+1)*TB)) in id order.
 * cls: 1 captured, 2 completed outward escape, 0 not counted.
 * partial layout: [blk][nbins][8] doubles, pcount: [blk][nbins][2]. */
__global__ void DMS_BESIDE_RESIDENT dms_hist_partial_kernel(const double* rows, const uint32_t* counts, const uint8_t* cls, uint32_t n, int nbins,
                                        uint32_t tb, double* partial, unsigned long long* pcount)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	const uint32_t blk = blockIdx.y;
	if(b >= nbins)
		return;
	const uint32_t i0 = blk * tb;
	const uint32_t i1 = (i0 + tb < n) ? i0 + tb : n;
	double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	unsigned long long c[2] = {0, 0};
	for(uint32_t i = i0; i < i1; i++)
	{
		const uint8_t k = cls[i];
		if(k == 0)
			continue;
		const double dt = rows[(size_t) i * 2 * nbins + b];
		const double v2 = rows[(size_t) i * 2 * nbins + nbins + b];
		const uint32_t cn = counts ? counts[(size_t) i * nbins + b] : 0u;
		const int o = (k == 1) ? 0 : 4;
		s[o + 0] += dt;
		s[o + 1] += v2;
		s[o + 2] += dt * dt;
		s[o + 3] += v2 * v2;
		c[k - 1] += cn;
	}
	double* p = partial + ((size_t) blk * nbins + b) * 8;
	for(int q = 0; q < 8; q++)
		p[q] = s[q];
	pcount[((size_t) blk * nbins + b) * 2 + 0] = c[0];
	pcount[((size_t) blk * nbins + b) * 2 + 1] = c[1];
}

/* hist[q][b] += sum over blocks (in block order); counts[k][b] likewise. */
__global__ void DMS_BESIDE_RESIDENT dms_hist_final_kernel(const double* partial, const unsigned long long* pcount, uint32_t nblk, int nbins,
                                      double* hist, unsigned long long* counts)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if(b >= nbins)
		return;
	double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	unsigned long long c0 = 0, c1 = 0;
	for(uint32_t blk = 0; blk < nblk; blk++)
	{
		const double* p = partial + ((size_t) blk * nbins + b) * 8;
		for(int q = 0; q < 8; q++)
			s[q] += p[q];
		c0 += pcount[((size_t) blk * nbins + b) * 2 + 0];
		c1 += pcount[((size_t) blk * nbins + b) * 2 + 1];
	}
	for(int q = 0; q < 8; q++)
		hist[(size_t) q * nbins + b] += s[q];
	counts[b] += c0;
	counts[nbins + b] += c1;
}

__global__ void dms_probe_kernel(int fn, const double* x, const double* y, double* out, size_t n)
{
	const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n)
		return;
	const double a = x[i];
	double r;
	switch(fn)
	{
		case 0: r = dm_sqrt(a); break;
		case 1: r = a / y[i]; break;
		case 2: r = dm_exp(a); break;
		case 3: r = dm_log(a); break;
		case 4: r = dm_sin(a); break;
		case 5: r = dm_cos(a); break;
		case 6: r = dm_erf(a); break;
		case 8: r = dm_div_shared(a, dm_recip_shared(y[i])); break;
		case 9: r = dm_div_const(a, y[i], 1.0 / y[i]); break;
		default: r = dm_acos(a); break;
	}
	out[i] = r;
}
