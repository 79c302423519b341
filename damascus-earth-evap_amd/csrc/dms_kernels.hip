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

template <bool MX, bool MP>
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
	for(;;)
	{
		dms_lane_advance_a<MX>(L, c);
		/* The wave-level services below (target sampling, scatter completion,
		 * capture counting, trajectory end, row flush, refill, start) act only
		 * on lanes in those phases: while every live lane is between RKF45
		 * attempts, as a lone long trajectory is, the wave skips them at once. */
		const bool service = __ballot((L.phase != PH_DONE) & (L.phase != PH_PREP) & (L.phase != PH_RK)) != 0;
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
				/* the scattering lane's parameter point (uniform over the wave) */
				const DmsPointView& Ps = MP ? a.job_points[__builtin_amdgcn_readlane((int) L.job, s)] : a.P;
				if constexpr(MX)
				{
					/* FP32 rates (1/s), same draw and order as below */
					float rate = 0.0f;
					if(k >= 0)
						rate = dms_rate_target_mx(Ps, a.S, r, v, k);
					const bool bad = (k >= 0) && (!mx_isfinite(rate) || rate < 0.0f);
					if(__ballot(bad) == 0)
					{
						float total = readlane_f(rate, DMS_MAX_TARGETS - 1);
						for(int i = 0; i < n_iso; i++)
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
					rate = dms_rate_target(Ps, a.S, r, v, k);
				const bool bad = (k >= 0) && (!dm_isfinite(rate) || rate < 0.0);
				if(__ballot(bad) == 0)
				{
					/* same summation and comparison order as dms_choose_target */
					const double re = readlane_d(rate, DMS_MAX_TARGETS - 1);
					double total = re;
					for(int i = 0; i < n_iso; i++)
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
			if(__builtin_expect(L.phase == PH_FIN, 0))
			{
				uint32_t fl = 0;
				int reason = 0;
				flush_cls = dms_lane_finish_rows_pending(L, c, a.ics[L.slot], a.records ? a.records + L.slot : nullptr,
				                                         a.summ ? a.summ + L.slot : nullptr, &fl, &reason);
				if(a.job_done)
					dms_count_finished(a.job_done + DMS_PROGRESS_WORDS * L.job, fl, reason, c.J.terminate_on_capture);
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
					const uint32_t job_s = (uint32_t) __builtin_amdgcn_readlane((int) L.job, s);
					const uint32_t n_s = (uint32_t) __builtin_amdgcn_readlane((int) L.n_touched, s);
					const int cls_s = __builtin_amdgcn_readlane(flush_cls, s);
					dms_flush_rows(c.R, c.J.nbins, row_s, job_s, n_s, cls_s, (uint32_t) lane, 64u);
				}
				if(flush_cls >= 0)
					L.n_touched = 0;
				/* the cleared rows are reused by their lanes' next trajectories */
				__threadfence_block();
			}
			/* lane refill: one atomic per wave */
			const unsigned long long nmask = __ballot(L.phase == PH_NEED);
			if(__builtin_expect(nmask != 0, 0))
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
		}
		dms_lane_advance_b<MX>(L, c);
		if(__ballot(L.phase != PH_DONE) == 0)
			break;
	}
}

__global__ void __launch_bounds__(DMS_TRAJ_BLOCK) dms_traj_kernel(DmsTrajArgs a) { dms_traj_body<false, false>(a); }

__global__ void __launch_bounds__(DMS_TRAJ_BLOCK) dms_traj_kernel_mixed(DmsTrajArgs a) { dms_traj_body<true, false>(a); }

__global__ void __launch_bounds__(DMS_TRAJ_BLOCK) dms_traj_kernel_multi(DmsTrajArgs a) { dms_traj_body<false, true>(a); }

__global__ void __launch_bounds__(DMS_TRAJ_BLOCK) dms_traj_kernel_mixed_multi(DmsTrajArgs a) { dms_traj_body<true, true>(a); }

/* FP32 copy of the rate table in 1/s for the mixed-precision kernel */
__global__ void dms_grid_f32_kernel(const double* grid, size_t n, float* out)
{
	const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if(i < n)
		out[i] = (float) (grid[i] * DMS_sec);
}

__global__ void dms_ic_kernel(DmsHaloView H, uint32_t seed, uint32_t point_id, uint64_t first_id, uint32_t n, double R_final,
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

__global__ void dms_grid_kernel(DmsPointView P, DmsSolarView S, int Nr, int Nv, double vmax, double* grid)
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

__global__ void dms_rates_kernel(DmsPointView P, DmsSolarView S, int which, const double* r, const double* v, double* out, size_t n)
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
__global__ void dms_hist_partial_kernel(const double* rows, const uint32_t* counts, const uint8_t* cls, uint32_t n, int nbins,
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
__global__ void dms_hist_final_kernel(const double* partial, const unsigned long long* pcount, uint32_t nblk, int nbins,
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
