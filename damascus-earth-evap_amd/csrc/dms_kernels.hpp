/* Kernel argument blocks and launch entry points (private to libdmsun). */
#ifndef DMS_KERNELS_HPP
#define DMS_KERNELS_HPP

#include <hip/hip_runtime.h>

#include "dms_engine.h"
#include "dms_halo.h"

#define DMS_TRAJ_BLOCK 256

/* Dynamic LDS of dms_traj_kernel: nodes x[n], then Mass and v_esc^2 cubics. */
__host__ __device__ inline size_t dms_lds_mass_offset(int n) { return ((size_t) n * sizeof(double) + 31) & ~(size_t) 31; }
__host__ __device__ inline size_t dms_lds_bytes(int n) { return dms_lds_mass_offset(n) + 2 * (size_t) (n - 1) * sizeof(DmsCubic); }

#define DMS_PROGRESS_WORDS 4

struct DmsTrajArgs
{
	DmsSolarView S;
	DmsPointView P;
	DmsJobView J;
	const dms_event* ics;
	const int32_t* shift_ok;
	uint32_t n;
	dms_traj_record* records;
	DmsSummary* summ;
	DmsRows R;
	unsigned int* next;
	/* jobs sharing the launch: slots [job_start[j], job_start[j+1]) belong to job j,
	 * whose trajectory ids start at job_first_id[j] */
	uint32_t n_jobs;
	const uint32_t* job_start;
	const uint64_t* job_first_id;
	/* live progress per job, DMS_PROGRESS_WORDS counters each (finished,
	 * captured, completed escapes, numerical failures), read by the host while
	 * the launch runs; may be null */
	unsigned int* job_done;
	/* jobs of different parameter points (multi-point launch): per-job point
	 * views and Philox keys {seed, point_id}; null for one point */
	const DmsPointView* job_points;
	const uint32_t* job_keys;
	/* capture-target chunks (may be null): fill_ctl[0] counts the launch's
	 * trajectories that got captured; when it reaches cap_stop, fill_ctl[1] (the
	 * slot limit, initially n) drops to the slots handed out so far plus
	 * fill_extra, and no slot at or past it is started */
	unsigned int* fill_ctl;
	uint32_t cap_stop, fill_extra;
};

__global__ void dms_traj_kernel(DmsTrajArgs a);
__global__ void dms_traj_kernel_mixed(DmsTrajArgs a);
__global__ void dms_traj_kernel_multi(DmsTrajArgs a);
__global__ void dms_traj_kernel_mixed_multi(DmsTrajArgs a);
__global__ void dms_grid_f32_kernel(const double* grid, size_t n, float* out);
__global__ void dms_ic_kernel(DmsHaloView H, uint32_t seed, uint32_t point_id, uint64_t first_id, uint32_t n, double R_final,
                              dms_event* out, int32_t* ok);
__global__ void dms_grid_kernel(DmsPointView P, DmsSolarView S, int Nr, int Nv, double vmax, double* grid);
__global__ void dms_rates_kernel(DmsPointView P, DmsSolarView S, int which, const double* r, const double* v, double* out, size_t n);
__global__ void dms_hist_partial_kernel(const double* rows, const uint32_t* counts, const uint8_t* cls, uint32_t n, int nbins,
                                        uint32_t tb, double* partial, unsigned long long* pcount);
__global__ void dms_hist_final_kernel(const double* partial, const unsigned long long* pcount, uint32_t nblk, int nbins,
                                      double* hist, unsigned long long* counts);
__global__ void dms_probe_kernel(int fn, const double* x, const double* y, double* out, size_t n);

#endif
