/* Kernel argument blocks and launch entry points (private to libdmsun). */
#ifndef DMS_KERNELS_HPP
#define DMS_KERNELS_HPP

#include <hip/hip_runtime.h>

#include <cstddef>

#include "dms_engine.h"
#include "dms_halo.h"

#define DMS_TRAJ_BLOCK 256

#define DMS_TRAJ_WAVES (DMS_TRAJ_BLOCK / 64)

/* Two-wave split of a lone trajectory (dms_engine.h): one exchange area per
 * wave of a block, in LDS.  The trajectory's wave (the only live lane of its
 * wave) asks for a helper (want); a wave of the block whose lanes have all
 * finished claims it (helper, by compare-and-swap) and serves it until that
 * wave's lanes have all finished too.  Each session (a step whose post the
 * trajectory's wave is about to run) starts from seed; the helper writes the
 * steps it prepares to the ring e[] (head, written by the helper) and the
 * trajectory's wave reads them (tail); entries of an older session are
 * skipped.  A second finished wave may claim helper2: it evaluates each
 * entry's rate queries (one per lane) behind the first helper (rhead), and the
 * trajectory's post takes them when rhead has passed the entry. */
#define DMS_PIPE_DEPTH 2

struct DmsPipeSlot
{
	unsigned int want;    /* 1 while the wave's lone trajectory would take a helper */
	unsigned int helper;  /* helper wave + 1 once claimed, 0 = none */
	unsigned int session; /* current session, 0 = none */
	unsigned int head;    /* entries written (helper) */
	unsigned int tail;    /* entries read (trajectory's wave) */
	unsigned int helper2; /* second helper wave + 1 (the entries' rate queries), 0 = none */
	unsigned int rhead;   /* entries whose rate queries the second helper has evaluated */
	unsigned int pad;
	DmsPipeSeed seed;
	DmsPipeEntry e[DMS_PIPE_DEPTH];
};

struct DmsPipeBlock
{
	unsigned int waves_done; /* waves whose lanes have all finished */
	unsigned int pad[3];
	DmsPipeSlot slot[DMS_TRAJ_WAVES];
};

/* Dynamic LDS of dms_traj_kernel: nodes x[n], then Mass and v_esc^2 cubics,
 * then the two-wave exchange areas. */
__host__ __device__ inline size_t dms_lds_mass_offset(int n) { return ((size_t) n * sizeof(double) + 31) & ~(size_t) 31; }
__host__ __device__ inline size_t dms_lds_pipe_offset(int n)
{
	return (dms_lds_mass_offset(n) + 2 * (size_t) (n - 1) * sizeof(DmsCubic) + 15) & ~(size_t) 15;
}
__host__ __device__ inline size_t dms_lds_bytes(int n) { return dms_lds_pipe_offset(n) + sizeof(DmsPipeBlock); }

#define DMS_PROGRESS_WORDS 4

/* Serving launch: one persistent launch that runs while a host thread appends
 * trajectories of admitted requests as slots.  Everything the host and the
 * running launch exchange lives in host memory mapped for the device
 * (fine-grained, read and written with system-scope accesses), so that no copy
 * or kernel has to run beside the launch.
 *
 * Slot idx (handed out in order from `next`, offset by `base`) sits at ring
 * position idx & ring_mask; the host publishes slots by writing their
 * descriptors and then raising ctl[0] (slots published).  A finished slot's
 * summary goes to results[position], then position + 1 to the completion
 * queue cq.  Histogram rows are summed per sync round: round buffer b holds the
 * sums of one round of one request in device memory (R.hist/R.hcounts with job
 * index b); the slot that completes the round copies them to the host mailbox
 * b, clears them and stamps round_stamp[b] with round_gen[b]. */
#define DMS_SERVE_NO_ROUND 0xffffffffu
#define DMS_SERVE_CANCELLED (-2)

struct DmsServeReq
{
	DmsPointView P;
	DmsHaloView H;
	uint32_t seed, point_id;
	uint32_t cancel; /* 1: slots not yet started end at once (summary reason DMS_SERVE_CANCELLED) */
	uint32_t pad;
};

static_assert(sizeof(DmsPointView) % 8 == 0 && sizeof(DmsHaloView) % 8 == 0, "serving loads read 8-byte words");
static_assert(offsetof(DmsServeReq, seed) % 8 == 0 && sizeof(DmsServeReq) % 8 == 0, "serving loads read 8-byte words");

struct DmsServeSlot
{
	uint32_t req;   /* request table index */
	uint32_t round; /* round buffer, or DMS_SERVE_NO_ROUND */
	uint64_t id;    /* trajectory id (Philox counter) */
};

struct DmsServeView
{
	const unsigned long long* ctl; /* host: [0] slots published, [8] stop */
	unsigned long long base;       /* slot index of this launch's first handed-out slot */
	uint32_t ring_mask;
	const DmsServeSlot* slots;     /* host, ring */
	const DmsServeReq* reqs;       /* host, request table */
	DmsSummary* results;           /* host, ring */
	uint32_t* cq;                  /* host, ring: ring position + 1 of finished slots */
	unsigned long long* counters;  /* device: [0] slots handed out, [1] completion tickets,
	                                * [2] published count and [3] stop flag as last read from
	                                * the host, [4] constant-clock time of that read */
	const uint32_t* round_size;    /* host */
	const unsigned long long* round_gen; /* host */
	unsigned long long* round_stamp;     /* host */
	double* mb_hist;               /* host [round][8][nbins] */
	unsigned long long* mb_counts; /* host [round][2][nbins] */
	uint32_t* round_done;          /* device: finished slots per round buffer */
};

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
	/* lanes a wave lets go free before it takes new slots (1: at once); a wave
	 * with no lane left running refills whatever it has free */
	uint32_t refill_batch;
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
	/* host memory mapped for the device (may be null): nonzero ends the launch;
	 * each wave reads it every 1024 loop iterations */
	const unsigned int* cancel;
	/* serving launch (dms_traj_kernel_serve*) */
	DmsServeView V;
	/* 1: a lone trajectory runs on its own wave only (no helper wave) */
	uint32_t lone_serial;
};

__global__ void dms_traj_kernel(DmsTrajArgs a);
__global__ void dms_traj_kernel_mixed(DmsTrajArgs a);
__global__ void dms_traj_kernel_multi(DmsTrajArgs a);
__global__ void dms_traj_kernel_mixed_multi(DmsTrajArgs a);
__global__ void dms_traj_kernel_serve(DmsTrajArgs a);
__global__ void dms_traj_kernel_mixed_serve(DmsTrajArgs a);
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
