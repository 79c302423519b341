/* CPU ORACLE -- TEST INFRASTRUCTURE ONLY (see dms_oracle.h).
 * Batch driver: Simulation_Data::Generate_Data per-trajectory accounting and
 * round logic restated from src/Data_Generation.cpp:59-176,488-972 for one
 * rank.  Trajectories are simulated speculatively in id order by a pool of
 * threads (one independent Philox stream per trajectory), then folded into the
 * run totals strictly in id order, so the totals do not depend on the thread
 * count. */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle_internal.h"

typedef struct
{
	const or_sim_cfg* cfg;
	const dms_job_desc* job;
	const dms_halo* halo;
	const dms_event* given; /* NULL: generate initial conditions */
	uint64_t first_id;
	size_t n;
	dms_traj_record* rec;
	double* rows;      /* n x 2 x nbins, may be NULL */
	uint32_t* counts;  /* n x nbins, may be NULL */
	int32_t* shift_ok;
	size_t next;
	pthread_mutex_t lock;
	int status;
} or_batch;

static void run_one(or_batch* B, size_t i)
{
	const or_sim_cfg* cfg = B->cfg;
	const int nb = cfg->nbins;
	double* dt = B->rows ? B->rows + i * 2 * (size_t) nb : NULL;
	double* v2 = B->rows ? dt + nb : NULL;
	uint32_t* cn = B->counts ? B->counts + i * (size_t) nb : NULL;
	if(dt)
	{
		memset(dt, 0, sizeof(double) * 2 * (size_t) nb);
	}
	if(cn)
		memset(cn, 0, sizeof(uint32_t) * (size_t) nb);
	const uint64_t id = B->first_id + i;
	dms_event ic;
	int ok = 1;
	if(B->given)
		ic = B->given[i];
	else
	{
		dm_rng rng;
		dm_rng_init(&rng, B->job->seed, B->job->point_id, id, DMS_STREAM_IC);
		or_make_initial_condition(cfg->P->S, B->halo, &rng, &ic);
		ok = or_kepler_shift_impl(&ic, cfg->max_distance);
	}
	if(B->shift_ok)
		B->shift_ok[i] = ok;
	if(!ok)
	{
		/* src/Data_Generation.cpp:749-761 */
		dms_traj_record* R = &B->rec[i];
		memset(R, 0, sizeof(*R));
		R->initial = ic;
		R->final = ic;
		R->trajectory_id = id;
		R->reason = DMS_NUMERICAL_FAILURE;
		R->flags = DMS_F_SHIFT_FAILED;
		R->t_capture_sec = -1.0;
		R->t_last_bound_sec = -1.0;
		R->t_final_unbinding_scatter_sec = dm_nan();
		R->t_boundary_escape_sec = dm_nan();
		R->t_termination_sec = -1.0;
		R->r_first_negative_km = -1.0;
		R->E_first_negative_eV = dm_nan();
		R->dE_first_negative_from_prev_eV = dm_nan();
		return;
	}
	OR_OPS_PUSH(OR_OPS_TRAJECTORY);
	const int st = or_simulate_one(cfg, &ic, B->job->seed, B->job->point_id, id, &B->rec[i], dt, v2, cn);
	OR_OPS_POP();
	if(st != DMS_OK)
		B->status = st;
}

static void* batch_worker(void* arg)
{
	or_batch* B = (or_batch*) arg;
	for(;;)
	{
		pthread_mutex_lock(&B->lock);
		const size_t i = B->next;
		B->next += 8;
		pthread_mutex_unlock(&B->lock);
		if(i >= B->n)
			break;
		const size_t e = (i + 8 < B->n) ? i + 8 : B->n;
		for(size_t k = i; k < e; k++)
			run_one(B, k);
	}
	return NULL;
}

static void run_batch(or_batch* B, int n_threads)
{
	B->next = 0;
	B->status = DMS_OK;
	pthread_mutex_init(&B->lock, NULL);
	if(n_threads <= 1)
		batch_worker(B);
	else
	{
		if(n_threads > 256)
			n_threads = 256;
		pthread_t th[256];
		for(int k = 0; k < n_threads; k++)
			pthread_create(&th[k], NULL, batch_worker, B);
		for(int k = 0; k < n_threads; k++)
			pthread_join(th[k], NULL);
	}
	pthread_mutex_destroy(&B->lock);
}

static void cfg_from_job(or_sim_cfg* cfg, const or_point* P, const dms_job_desc* job)
{
	memset(cfg, 0, sizeof(*cfg));
	cfg->P = P;
	cfg->max_time_steps = job->max_free_steps;
	cfg->max_scatterings = job->max_scatterings;
	cfg->max_distance = job->r_max_rsun * OR_rSun;
	cfg->terminate_on_capture = (job->mode == DMS_MODE_CAPTURE);
	cfg->rk_budget = job->rk_attempt_budget;
	cfg->nbins = job->n_bins > 0 ? job->n_bins : 2000;
	cfg->bin_width_km = (2.0 * OR_R_SUN_KM) / cfg->nbins;
}

int or_simulate_events(const or_point* P, const dms_job_desc* job, const dms_event* initial, size_t n,
                       dms_traj_record* records, double* hist_rows, uint32_t* count_rows, int n_threads)
{
	or_sim_cfg cfg;
	cfg_from_job(&cfg, P, job);
	or_batch B;
	memset(&B, 0, sizeof(B));
	B.cfg = &cfg;
	B.job = job;
	B.halo = &job->halo;
	B.given = initial;
	B.first_id = job->first_trajectory_id;
	B.n = n;
	B.rec = records;
	B.rows = hist_rows;
	B.counts = count_rows;
	run_batch(&B, n_threads);
	return B.status;
}

/* Normal_Mode_MPI_Sync_Interval, src/Data_Generation.cpp:59-74 */
static uint64_t sync_interval(double sigma_cm2)
{
	if(!dm_isfinite(sigma_cm2) || sigma_cm2 <= 0.0)
		return 128;
	if(sigma_cm2 >= 1.0e-35)
		return 64;
	if(sigma_cm2 >= 1.0e-36)
		return 128;
	if(sigma_cm2 >= 1.0e-37)
		return 1024;
	if(sigma_cm2 >= 1.0e-38)
		return 8192;
	if(sigma_cm2 >= 1.0e-39)
		return 65536;
	return 1048576;
}

static int is_numerical(int r)
{
	return r == DMS_NUMERICAL_FAILURE || r == DMS_NON_FINITE_STATE || r == DMS_SPEED_LIMIT || r == DMS_ENERGY_DRIFT_ESCAPE || r == DMS_UNKNOWN;
}
static int is_truncation(int r) { return r == DMS_WALL_TIME_LIMIT || r == DMS_MAX_FREE_STEPS || r == DMS_MAX_SCATTERINGS; }

static double now_sec(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

int or_run_job(const or_point* P, const dms_job_desc* job, int n_threads, dms_counters* C, double* hist8,
               uint64_t* counts2, dms_evap_event* events, size_t event_capacity, size_t* n_events,
               dms_traj_record* records, size_t record_capacity, size_t* n_records)
{
	or_sim_cfg cfg;
	cfg_from_job(&cfg, P, job);
	const int nb = cfg.nbins;
	const int capture_mode = (job->mode == DMS_MODE_CAPTURE);
	memset(C, 0, sizeof(*C));
	if(hist8)
		memset(hist8, 0, sizeof(double) * 8 * (size_t) nb);
	if(counts2)
		memset(counts2, 0, sizeof(uint64_t) * 2 * (size_t) nb);
	if(n_events)
		*n_events = 0;
	if(n_records)
		*n_records = 0;
	const uint64_t target = job->capture_target == 0 ? UINT64_MAX : job->capture_target;
	const uint64_t max_traj = job->max_trajectories == 0 ? UINT64_MAX : job->max_trajectories;
	const int unlimited = job->max_trajectories == 0;
	double sp, se;
	or_point_cross_sections(P, &sp, &se);
	const uint64_t interval = capture_mode ? 32 : sync_interval(sp);
	C->sync_interval = (int32_t) interval;

	const size_t chunk = 4096;
	or_batch B;
	memset(&B, 0, sizeof(B));
	B.cfg = &cfg;
	B.job = job;
	B.halo = &job->halo;
	B.rec = (dms_traj_record*) malloc(sizeof(dms_traj_record) * chunk);
	const int want_rows = !capture_mode && (hist8 != NULL || counts2 != NULL);
	B.rows = want_rows ? (double*) malloc(sizeof(double) * 2 * (size_t) nb * chunk) : NULL;
	B.counts = want_rows ? (uint32_t*) malloc(sizeof(uint32_t) * (size_t) nb * chunk) : NULL;
	B.shift_ok = (int32_t*) malloc(sizeof(int32_t) * chunk);
	size_t have = 0, cursor = 0;
	uint64_t next_id = job->first_trajectory_id;

	const double t0 = now_sec();
	uint64_t local_total = 0, captured = 0;
	int stop = 0; /* SimulationStopReason */
	int status = DMS_OK;
	while(captured < target)
	{
		const uint64_t remaining = target - captured;
		uint64_t capacity = 0;
		if(local_total < max_traj)
			capacity = (max_traj - local_total < interval) ? max_traj - local_total : interval;
		if(capacity == 0)
		{
			stop = 1;
			break;
		}
		const uint64_t batch = capture_mode ? (remaining < capacity ? remaining : capacity) : capacity;
		C->sync_rounds++;
		C->final_round_trajectories = batch;
		for(uint64_t k = 0; k < batch; k++)
		{
			if(cursor == have)
			{
				B.first_id = next_id;
				B.n = chunk;
				/* chunks never exceed the buffers; a round longer than a chunk
				 * takes several (cursor == have refills) */
				if(!unlimited && max_traj - local_total < chunk)
					B.n = (size_t) (max_traj - local_total);
				run_batch(&B, n_threads);
				if(B.status != DMS_OK)
					status = B.status;
				C->simulated_trajectories += B.n;
				next_id += B.n;
				have = B.n;
				cursor = 0;
			}
			const size_t i = cursor++;
			const dms_traj_record* R = &B.rec[i];
			const int shift_ok = B.shift_ok[i];
			if(!shift_ok)
			{
				C->initial_shift_failures++;
				C->numerical_failures++;
			}
			local_total++;
			C->trajectories++;
			C->total_scatterings += R->n_scatterings;
			C->reason_counts[R->reason]++;
			C->n_rk_attempts += R->n_rk_attempts;
			C->n_steps_interior += R->n_steps_interior;
			C->n_steps_exterior += R->n_steps_exterior;
			C->n_tau_rejections += R->n_tau_rejections;
			C->n_tv_iterations += R->n_tv_iterations;
			if(R->n_rk_attempts > C->max_rk_attempts)
				C->max_rk_attempts = R->n_rk_attempts;
			const int completed_escape = R->reason == DMS_OUTWARD_ESCAPE;
			if(shift_ok && is_numerical(R->reason))
				C->numerical_failures++;
			if(is_truncation(R->reason))
				C->computational_truncations++;
			const double* dt = want_rows ? B.rows + i * 2 * (size_t) nb : NULL;
			const double* v2 = want_rows ? dt + nb : NULL;
			const uint32_t* cn = want_rows ? B.counts + i * (size_t) nb : NULL;
			if(R->flags & DMS_F_CAPTURED)
			{
				C->captured++;
				captured++;
				if(!capture_mode)
				{
					if(R->flags & DMS_F_EVENT_OBSERVED)
						C->complete_evaporation++;
					else
						C->invalid_survival_captured++;
					if(hist8)
						for(int b = 0; b < nb; b++)
						{
							hist8[0 * nb + b] += dt[b];
							hist8[1 * nb + b] += v2[b];
							hist8[2 * nb + b] += dt[b] * dt[b];
							hist8[3 * nb + b] += v2[b] * v2[b];
						}
					if(counts2)
						for(int b = 0; b < nb; b++)
							counts2[b] += cn[b];
					/* Build_Evaporation_Record + Make_Compact_Evaporation_Event, :118-176 */
					if(dm_isfinite(R->t_capture_sec))
					{
						const int nie = (R->flags & DMS_F_NUMERICALLY_INVALID_ESCAPE) || R->reason == DMS_ENERGY_DRIFT_ESCAPE;
						const int sv = (R->flags & DMS_F_SURVIVAL_VALID) && !nie;
						const int eo = sv && (R->flags & DMS_F_EVENT_OBSERVED);
						const double lu = eo ? (R->t_final_unbinding_scatter_sec - R->t_capture_sec) : -1.0;
						const double lifetime = (dm_isfinite(lu) && lu >= 0.0) ? lu : -1.0;
						if(sv && eo && dm_isfinite(lifetime) && lifetime >= 0.0 && n_events)
						{
							if(*n_events < event_capacity && events)
							{
								dms_evap_event* E = &events[*n_events];
								E->rank = 0;
								E->pad = 0;
								E->trajectory_id = C->trajectories;
								E->completion_wall_time_sec = now_sec() - t0;
								E->lifetime_unbinding_sec = lifetime;
							}
							(*n_events)++;
						}
					}
				}
			}
			else if(completed_escape)
			{
				C->completed_outward_escapes++;
				if(!capture_mode)
				{
					if(hist8)
						for(int b = 0; b < nb; b++)
						{
							hist8[4 * nb + b] += dt[b];
							hist8[5 * nb + b] += v2[b];
							hist8[6 * nb + b] += dt[b] * dt[b];
							hist8[7 * nb + b] += v2[b] * v2[b];
						}
					if(counts2)
						for(int b = 0; b < nb; b++)
							counts2[nb + b] += cn[b];
					if(R->flags & DMS_F_FREE)
						C->free_particles++;
					else if(R->flags & DMS_F_REFLECTED)
					{
						C->reflected_particles++;
						if(R->flags & DMS_F_REFLECTION_SHIFT_FAILED)
						{
							C->final_reflection_shift_failures++;
							C->numerical_failures++;
						}
					}
				}
			}
			if(records && n_records && *n_records < record_capacity)
				records[(*n_records)++] = *R;
		}
		/* run-level fuses, :893-938 */
		if(C->trajectories > 0)
		{
			const double shift_fraction = (double) C->initial_shift_failures / (double) C->trajectories;
			if(shift_fraction > 1.0e-2)
			{
				stop = 3;
				break;
			}
			const double invalid_fraction = (double) (C->numerical_failures + C->computational_truncations) / (double) C->trajectories;
			if(unlimited && invalid_fraction > 1.0e-2)
			{
				stop = 4;
				break;
			}
		}
	}
	C->capture_target_overshoot = (captured > target) ? captured - target : 0;
	if(captured < target && stop == 0)
		stop = 2;
	C->early_stop_reason = stop;
	C->computing_time_sec = now_sec() - t0;
	free(B.rec);
	free(B.rows);
	free(B.counts);
	free(B.shift_ok);
	return status;
}
