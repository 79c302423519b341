/*
 * Host replay of the trajectory kernel's per-lane phase functions
 * (TEST INFRASTRUCTURE).
 *
 * Runs dms_lane_advance_a / finish_scatter / finish / start / advance_b from
 * damascus-earth-evap_amd/csrc/dms_engine.h one lane at a time on the CPU, with
 * the wave-cooperative Sample_Target replaced by its serial form
 * (dms_choose_target), the IC kernel's dms_initial_condition and the grid
 * kernel's dms_rate_computed.  The CPU tests compare its output with the
 * oracle bit for bit, which checks the kernel code paths without a GPU.
 */
#include <cstring>
#include <vector>

#include "../../damascus-earth-evap_amd/csrc/dms_engine.h"
#include "../../damascus-earth-evap_amd/csrc/dms_halo.h"
#include "../../damascus-earth-evap_amd/csrc/dms_host.hpp"

struct rp_ctx
{
	dms::SolarTables T;
};

struct rp_point
{
	rp_ctx* ctx;
	dms::PointParams P;
	std::vector<double> grid;
	std::vector<float> grid32;
	int Nr, Nv;
	DmsPointView view;
};

extern "C" {

rp_ctx* rp_create(const double* rows, size_t n_rows)
{
	rp_ctx* c = new rp_ctx();
	if(!dms::build_solar_tables(rows, n_rows, c->T).empty())
	{
		delete c;
		return nullptr;
	}
	return c;
}

void rp_free(rp_ctx* c) { delete c; }

double rp_solar_eval(rp_ctx* c, int which, double r)
{
	const DmsSolarView V = c->T.view();
	switch(which)
	{
		case 0: return dms_mass(V, r);
		case 1: return dms_temperature(V, r);
		case 2: return dms_local_escape_speed(V, r);
		case 3: return c->T.mass_density(r);
		default:
		{
			const int k = (which == 4) ? DMS_MAX_TARGETS - 1 : which - 100;
			if(!dm_isfinite(r))
				return dm_nan();
			if(r < 0.0)
				r = 0.0;
			if(r > DMS_rSun)
				return 0.0;
			const int j = dms_interval(V, r);
			return dms_density_at(V, j, r - V.x[j], k);
		}
	}
}

rp_point* rp_point_create(rp_ctx* c, const dms_point_desc* d)
{
	rp_point* p = new rp_point();
	p->ctx = c;
	if(!dms::build_point_params(c->T, *d, p->P).empty())
	{
		delete p;
		return nullptr;
	}
	const int Nr = d->interpolation_points;
	const int Nv = (d->speed_points == 0) ? Nr : d->speed_points;
	const bool grid = Nr >= 2 && Nv >= 2;
	p->Nr = grid ? Nr : 0;
	p->Nv = grid ? Nv : 0;
	DmsPointView& v = p->view;
	v.mass = p->P.mass;
	v.mu_e = p->P.mu_e;
	v.interaction = p->P.interaction;
	v.ff = p->P.ff;
	v.m_dark_photon = p->P.m_dark_photon;
	v.q_reference = p->P.q_reference;
	v.sigma_e = p->P.sigma_e;
	v.sigma_N = p->P.sigma_N;
	v.use_grid = grid;
	v.Nr = p->Nr;
	v.Nv = p->Nv;
	v.inv_dr = grid ? (double) (p->Nr - 1) / DMS_rSun : 0.0;
	v.inv_dv = grid ? (double) (p->Nv - 1) / DMS_V_MAX : 0.0;
	v.vmax = grid ? DMS_V_MAX : 0.0;
	v.grid = nullptr;
	v.grid32 = nullptr;
	if(grid)
	{
		const size_t cells = (size_t) p->Nr * (size_t) p->Nv;
		if(d->host_rate_grid)
			p->grid.assign(d->host_rate_grid, d->host_rate_grid + cells);
		else
		{
			/* same node construction as dms_grid_kernel */
			p->grid.resize(cells);
			const DmsSolarView S = c->T.view();
			const double dv = (DMS_V_MAX - 0.0) / (double) (p->Nv - 1);
			for(int i = 0; i < p->Nr; i++)
				for(int j = 0; j < p->Nv; j++)
				{
					const double radius = DMS_rSun * (double) i / (double) (p->Nr - 1);
					const double speed = 0.0 + j * dv;
					p->grid[(size_t) i * p->Nv + j] = dms_rate_computed(v, S, radius, speed);
				}
		}
		v.grid = p->grid.data();
		/* FP32 table in 1/s, as dms_grid_f32_kernel makes it */
		p->grid32.resize(p->grid.size());
		for(size_t k = 0; k < p->grid.size(); k++)
			p->grid32[k] = (float) (p->grid[k] * DMS_sec);
		v.grid32 = p->grid32.data();
	}
	return p;
}

void rp_point_free(rp_point* p) { delete p; }

double rp_rate(rp_point* p, int which, double r, double v)
{
	const DmsSolarView S = p->ctx->T.view();
	if(which == 0)
		return dms_rate_total(p->view, S, r, v);
	if(which == 1)
		return dms_rate_computed(p->view, S, r, v);
	if(which == 2)
		return dms_rate_target(p->view, S, r, v, DMS_MAX_TARGETS - 1);
	return dms_rate_target(p->view, S, r, v, which - 3);
}

int rp_grid(rp_point* p, double* out, size_t n)
{
	if(n < p->grid.size())
		return -1;
	memcpy(out, p->grid.data(), sizeof(double) * p->grid.size());
	return 0;
}

void rp_initial_conditions(rp_ctx* c, const dms_halo* h, uint32_t seed, uint32_t point_id, uint64_t first_id, size_t n,
                           double r_final_rsun, dms_event* out, int32_t* ok)
{
	DmsHaloView H;
	dms_halo_init(H, *h);
	const DmsSolarView V = c->T.view();
	H.vescR = dms_local_escape_speed(V, DMS_rSun);
	H.vesc_asym = dms_local_escape_speed(V, 1000.0 * DMS_AU);
	H.norm = dms_shm_moment(H, 1) + H.vescR * H.vescR * dms_shm_moment(H, -1);
	for(size_t i = 0; i < n; i++)
	{
		dm_rng rng;
		dm_rng_init(&rng, seed, point_id, first_id + i, DMS_STREAM_IC);
		dms_initial_condition(H, &rng, &out[i]);
		ok[i] = (r_final_rsun > 0.0) ? (dms_kepler_shift(out[i].r, out[i].v, r_final_rsun * DMS_rSun) ? 1 : 0) : 1;
	}
}

extern "C++" {
/* The wave-cooperative Sample_Target of the kernel, serially: the same
 * per-target formulas and the same summation and comparison order (FP64:
 * dms_choose_target; mixed: FP32 rates and reciprocal of the total). */
template <bool MX>
static void sample_target_serial(DmsLane& L, const DmsLaneCtx& c)
{
	if constexpr(!MX)
	{
		double rates[DMS_MAX_TARGETS];
		for(int k = 0; k < DMS_MAX_TARGETS; k++)
			rates[k] = 0.0;
		for(int k = 0; k < c.S.n_iso; k++)
			rates[k] = dms_rate_target(c.P, c.S, L.sc_r, L.sc_v, k);
		rates[DMS_MAX_TARGETS - 1] = dms_rate_target(c.P, c.S, L.sc_r, L.sc_v, DMS_MAX_TARGETS - 1);
		L.sc_target = dms_choose_target(rates, c.S.n_iso, L.sc_xi);
	}
	else
	{
		float rates[DMS_MAX_TARGETS];
		int target = -2;
		bool bad = false;
		for(int k = 0; k < DMS_MAX_TARGETS; k++)
		{
			rates[k] = 0.0f;
			if(k < c.S.n_iso || k == DMS_MAX_TARGETS - 1)
			{
				rates[k] = dms_rate_target_mx(c.P, c.S, L.sc_r, L.sc_v, k);
				bad = bad || !mx_isfinite(rates[k]) || rates[k] < 0.0f;
			}
		}
		if(!bad)
		{
			float total = rates[DMS_MAX_TARGETS - 1];
			for(int i = 0; i < c.S.n_iso; i++)
				total += rates[i];
			if(mx_isfinite(total) && total > 0.0f)
			{
				const float xf = (float) L.sc_xi;
				float sum = rates[DMS_MAX_TARGETS - 1] * mx_rcp(total);
				if(sum > xf)
					target = -1;
				else
					for(int i = 0; i < c.S.n_iso; i++)
					{
						sum += rates[i] * mx_rcp(total);
						if(sum > xf || i == c.S.n_iso - 1)
						{
							target = i;
							break;
						}
					}
			}
		}
		L.sc_target = target;
	}
}

} /* extern "C++" */

/* One trajectory through the kernel's phase drivers, with the wave-cooperative
 * Sample_Target done serially (same per-target formulas, same sum order). */
static void replay_lane(DmsLane& L, const DmsLaneCtx& c, const dms_event& ic, bool shift_ok, dms_traj_record* rec, DmsSummary* sum)
{
	dms_lane_start(L, c, ic, shift_ok, rec, sum);
	if(L.phase == PH_NEED)
		return;
	dms_lane_advance_b(L, c);
	while(L.phase != PH_NEED)
	{
		dms_lane_advance_a(L, c);
		if(L.phase == PH_SCAT)
			sample_target_serial<false>(L, c);
		dms_lane_finish_scatter(L, c);
		if(L.phase == PH_FIN)
		{
			dms_lane_finish(L, c, ic, rec, sum);
			break;
		}
		dms_lane_advance_b(L, c);
	}
}

/* The mixed-precision phase drivers, serially (FP32 Sample_Target in the
 * kernel's summation order). */
static void replay_lane_mx(DmsLane& L, const DmsLaneCtx& c, const dms_event& ic, bool shift_ok, dms_traj_record* rec, DmsSummary* sum)
{
	dms_lane_start(L, c, ic, shift_ok, rec, sum);
	if(L.phase == PH_NEED)
		return;
	dms_lane_advance_b<true>(L, c);
	while(L.phase != PH_NEED)
	{
		dms_lane_advance_a<true>(L, c);
		if(L.phase == PH_SCAT)
			sample_target_serial<true>(L, c);
		dms_lane_finish_scatter<true>(L, c);
		if(L.phase == PH_FIN)
		{
			dms_lane_finish(L, c, ic, rec, sum);
			break;
		}
		dms_lane_advance_b<true>(L, c);
	}
}

static void fill_job_view(DmsJobView& J, const dms_job_desc* job)
{
	J.max_time_steps = job->max_free_steps;
	J.max_scatterings = job->max_scatterings;
	J.rk_budget = job->rk_attempt_budget;
	J.max_distance = job->r_max_rsun * DMS_rSun;
	J.terminate_on_capture = (job->mode == DMS_MODE_CAPTURE);
	J.nbins = job->n_bins > 0 ? job->n_bins : 2000;
	J.bin_width_km = (2.0 * DMS_R_SUN_KM) / J.nbins;
	J.seed = job->seed;
	J.point_id = job->point_id;
	J.first_id = job->first_trajectory_id;
	J.mixed = (job->precision == DMS_MIXED) ? 1 : 0;
}

/* Serial replay of the per-lane histogram mode: nlanes scratch rows reused
 * round-robin, trajectories finished in id order, class sums flushed at each
 * trajectory end (Parameter-point jobs).  hist: 8 x nbins, hcounts: 2 x nbins. */
int rp_lane_hist(rp_point* p, const dms_job_desc* job, const dms_event* ics, const int32_t* ok, size_t n, size_t nlanes,
                 double* hist, unsigned long long* hcounts)
{
	DmsLaneCtx c;
	c.S = p->ctx->T.view();
	c.P = p->view;
	fill_job_view(c.J, job);
	const size_t nb = (size_t) c.J.nbins;
	if(c.J.terminate_on_capture || nlanes == 0 || nb > 65536)
		return -1;
	std::vector<DmsBinAcc> bins(nb * nlanes);
	memset(bins.data(), 0, sizeof(DmsBinAcc) * bins.size());
	std::vector<uint16_t> touched(nb * nlanes, 0);
	memset(hist, 0, sizeof(double) * 8 * nb);
	memset(hcounts, 0, sizeof(unsigned long long) * 2 * nb);
	c.R.rows = nullptr;
	c.R.counts = nullptr;
	c.R.bins = bins.data();
	c.R.touched = touched.data();
	c.R.per_lane = 1;
	c.R.hist = hist;
	c.R.hcounts = hcounts;
	for(size_t i = 0; i < n; i++)
	{
		DmsLane L;
		memset(&L, 0, sizeof(L));
		L.cur_bin = -1;
		L.slot = (uint32_t) i;
		L.row = (uint32_t) (i % nlanes);
		L.job = 0;
		L.id = c.J.first_id + i;
		DmsSummary s;
		replay_lane(L, c, ics[i], ok[i] != 0, nullptr, &s);
	}
	for(size_t k = 0; k < nb * nlanes; k++)
		if(bins[k].dt != 0.0 || bins[k].v2dt != 0.0 || bins[k].cnt != 0u)
			return -2; /* scratch not cleared after a flush */
	return 0;
}

/* Serial replay of dms_traj_kernel for n given initial events. */
int rp_simulate(rp_point* p, const dms_job_desc* job, const dms_event* ics, size_t n, dms_traj_record* recs, double* rows)
{
	DmsLaneCtx c;
	c.S = p->ctx->T.view();
	c.P = p->view;
	DmsJobView& J = c.J;
	fill_job_view(J, job);
	c.R.rows = J.terminate_on_capture ? nullptr : rows;
	c.R.counts = nullptr;
	c.R.bins = nullptr;
	c.R.touched = nullptr;
	c.R.per_lane = 0;
	c.R.hist = nullptr;
	c.R.hcounts = nullptr;
	if(rows)
		memset(rows, 0, sizeof(double) * 2 * (size_t) J.nbins * n);
	for(size_t i = 0; i < n; i++)
	{
		DmsLane L;
		memset(&L, 0, sizeof(L));
		L.cur_bin = -1;
		L.slot = (uint32_t) i;
		L.row = (uint32_t) i;
		L.job = 0;
		L.id = J.first_id + i;
		if(c.J.mixed)
			replay_lane_mx(L, c, ics[i], true, &recs[i], nullptr);
		else
			replay_lane(L, c, ics[i], true, &recs[i], nullptr);
	}
	return 0;
}

extern "C++" {
/* The two-wave split of a lone trajectory, serially: from the step whose
 * RKF45 attempts the trajectory's own lane has just run, a helper lane (all
 * fields it is not given hold garbage) prepares up to `depth` further steps
 * assuming each post continues, and the post takes them (dms_pipe_take).  A
 * post that rejects its step on optical depth opens a redo session: a fresh
 * helper lane prepares that step again from the redo seed (and the steps after
 * it), and the next post takes its first step.  Every other post exit
 * discards the helper's steps.  Returns the last post's value, or
 * DMS_POST_NEXT when the helper's preparation ended the segment (the lane
 * then runs its own preparation, which ends it the same way).  *taken counts
 * the steps taken from the helper, *dropped the ones discarded, *redos the
 * redo sessions. */
template <bool MX>
static int pipe_session(DmsLane& L, const DmsLaneCtx& c, int depth, uint64_t* taken, uint64_t* dropped, uint64_t* redos)
{
	int r;
	do
	{
		L.inner_iter++;
		L.rk_all++;
		r = MX ? dms_rk45_attempt_mx(L.prop, c.S, L.inner_iter) : dms_rk45_attempt(L.prop, c.S, L.inner_iter);
		L.rk_ok = (r >= 0) ? r : L.rk_ok;
	} while(r < 0);
	DmsPipeSeed seed;
	dms_pipe_make_seed<MX>(L, L.job, seed);
	DmsLane A;
	memset(&A, 0xC3, sizeof(A));
	dms_pipe_from_seed(A, seed);
	DmsStepEvent ev = seed.ev;
	bool redo_first = false;
	std::vector<DmsPipeEntry> ring;
	size_t head = 0;
	bool helper_ended = false;
	auto fill = [&]() {
		while(!helper_ended && ring.size() - head < (size_t) depth)
		{
			DmsPipeEntry e;
			memset(&e, 0xA7, sizeof(e));
			if(redo_first)
				e.eb = seed.eb;
			else
			{
				e.eb = ev;
				dms_pipe_assume_next(A, ev);
			}
			redo_first = false;
			e.r_before = A.r_before;
			if(!dms_pipe_prepare_step<MX>(A, c.P, c.S, c.J, e))
				helper_ended = true;
			else
				ev = e.ev;
			ring.push_back(e);
		}
	};
	DmsPipeEntry held;                 /* the step taken from the helper, as the kernel holds its ring entry */
	const DmsStepEvent* pre = &seed.ev; /* the first post takes the seed's event */
	const double* pre_rates = nullptr;  /* and computes its own rates */
	for(;;)
	{
		fill();
		const int pr = dms_step_post<MX>(L, c.P, c.S, c.J, c.R, pre, pre_rates);
		if(pr != DMS_POST_NEXT)
		{
			*dropped += ring.size() - head;
			if(!dms_post_rejected(L, pr))
				return pr;
			/* a redo session from a fresh helper */
			ring.clear();
			head = 0;
			helper_ended = false;
			dms_pipe_make_redo_seed(L, L.job, seed);
			memset(&A, 0xC3, sizeof(A));
			dms_pipe_from_seed(A, seed);
			redo_first = true;
			(*redos)++;
			fill();
		}
		held = ring[head++];
		if(held.kind == DMS_PIPE_PREP_END)
		{
			*dropped += ring.size() - head;
			return DMS_POST_NEXT;
		}
		dms_pipe_take(L, held);
		pre = &held.ev;
		/* the step's rate queries evaluated from the entry alone, as the
		 * kernel's second helper wave does (one query per lane there) */
		pre_rates = nullptr;
		if constexpr(MX)
		{
			DmsPipeRateInputsMx q;
			if(dms_pipe_rate_inputs_mx(held, c.J, q))
			{
				for(int k = 0; k < DMS_MAX_OPTICAL_DEPTH_PIECES + 2; k++)
					held.rt[k] = dms_pipe_rate_mx(held, q, c.P, c.S, k);
				pre_rates = held.rt;
			}
		}
		else
		{
			DmsPipeRateInputs q;
			if(dms_pipe_rate_inputs(held, c.J, q))
			{
				for(int k = 0; k < DMS_MAX_OPTICAL_DEPTH_PIECES + 2; k++)
					held.rt[k] = dms_pipe_rate(held, q, c.P, c.S, k);
				pre_rates = held.rt;
			}
		}
		(*taken)++;
	}
}

template <bool MX>
static void replay_lane_pipe(DmsLane& L, const DmsLaneCtx& c, const dms_event& ic, bool shift_ok, dms_traj_record* rec,
                             DmsSummary* sum, int depth, uint64_t* taken, uint64_t* dropped, uint64_t* redos)
{
	dms_lane_start(L, c, ic, shift_ok, rec, sum);
	if(L.phase == PH_NEED)
		return;
	dms_lane_advance_b<MX>(L, c);
	while(L.phase != PH_NEED)
	{
		if(L.phase == PH_RK)
		{
			const int pr = pipe_session<MX>(L, c, depth, taken, dropped, redos);
			L.phase = ((pr == -2) | (pr == DMS_POST_NEXT)) ? (int) PH_PREP
			                                                : dms_after_propagation(L, (pr == -3) ? (int) DMS_SCATTER : pr);
		}
		dms_lane_advance_a<MX>(L, c); /* the scatter hand-off when the session ended in one */
		if(L.phase == PH_SCAT)
			sample_target_serial<MX>(L, c);
		dms_lane_finish_scatter<MX>(L, c);
		if(L.phase == PH_FIN)
		{
			dms_lane_finish(L, c, ic, rec, sum);
			break;
		}
		dms_lane_advance_b<MX>(L, c);
	}
}

} /* extern "C++" */

/* rp_simulate with every step loop run through the two-wave split;
 * stats[0] steps taken from the helper, stats[1] steps discarded, stats[2]
 * redo sessions. */
int rp_simulate_pipe(rp_point* p, const dms_job_desc* job, const dms_event* ics, size_t n, dms_traj_record* recs,
                     double* rows, int depth, uint64_t* stats)
{
	DmsLaneCtx c;
	c.S = p->ctx->T.view();
	c.P = p->view;
	DmsJobView& J = c.J;
	fill_job_view(J, job);
	c.R.rows = J.terminate_on_capture ? nullptr : rows;
	c.R.counts = nullptr;
	c.R.bins = nullptr;
	c.R.touched = nullptr;
	c.R.per_lane = 0;
	c.R.hist = nullptr;
	c.R.hcounts = nullptr;
	if(depth < 1)
		return -1;
	if(rows)
		memset(rows, 0, sizeof(double) * 2 * (size_t) J.nbins * n);
	stats[0] = stats[1] = stats[2] = 0;
	for(size_t i = 0; i < n; i++)
	{
		DmsLane L;
		memset(&L, 0, sizeof(L));
		L.cur_bin = -1;
		L.slot = (uint32_t) i;
		L.row = (uint32_t) i;
		L.job = 0;
		L.id = J.first_id + i;
		if(c.J.mixed)
			replay_lane_pipe<true>(L, c, ics[i], true, &recs[i], nullptr, depth, &stats[0], &stats[1], &stats[2]);
		else
			replay_lane_pipe<false>(L, c, ics[i], true, &recs[i], nullptr, depth, &stats[0], &stats[1], &stats[2]);
	}
	return 0;
}

/* dm_div_const over n dividends for one divisor b and reciprocal y (the
 * engine passes 1.0 / b). */
void rp_div_const(const double* a, size_t n, double b, double y, double* out)
{
	for(size_t i = 0; i < n; i++)
		out[i] = dm_div_const(a[i], b, y);
}

} /* extern "C" */
