/*
 * Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11) and the per-trajectory
 * uniform stream used by the trajectory path.
 *
 * Stream layout (one independent stream per trajectory and purpose):
 *   key     = { seed (32 bit), point id }
 *   counter = { trajectory id low, trajectory id high, block, stream }
 * Each block yields two uniforms in [0,1) with 53 random bits:
 *   U = (word >> 11) * 2^-53, word = (x1<<32|x0) then (x3<<32|x2).
 * stream 0 feeds the simulation (the reference's Trajectory_Simulator::PRNG
 * draws, src/Simulation_Trajectory.cpp:795,1144,1193-1219,1238), stream 1
 * feeds Initial_Conditions (src/Simulation_Utilities.cpp:186-242).
 *
 * Valid C99 and HIP C++; integer-only, so host and device agree bit for bit.
 */
#ifndef DMS_PHILOX_H
#define DMS_PHILOX_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define DMP_FN static inline __host__ __device__
#else
#define DMP_FN static inline
#endif

#define DMS_STREAM_SIM 0u
#define DMS_STREAM_IC 1u

DMP_FN void dm_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
	uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
	uint32_t k0 = key[0], k1 = key[1];
	for(int round = 0; round < 10; round++)
	{
		const uint64_t p0 = (uint64_t) 0xD2511F53u * c0;
		const uint64_t p1 = (uint64_t) 0xCD9E8D57u * c2;
		const uint32_t hi0 = (uint32_t) (p0 >> 32), lo0 = (uint32_t) p0;
		const uint32_t hi1 = (uint32_t) (p1 >> 32), lo1 = (uint32_t) p1;
		const uint32_t n0 = hi1 ^ c1 ^ k0;
		const uint32_t n2 = hi0 ^ c3 ^ k1;
		c0 = n0;
		c1 = lo1;
		c2 = n2;
		c3 = lo0;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	out[0] = c0;
	out[1] = c1;
	out[2] = c2;
	out[3] = c3;
}

typedef struct
{
	uint32_t key0, key1;
	uint32_t traj_lo, traj_hi;
	uint32_t stream;
	uint32_t block;
	uint32_t has_cached;
	double cached;
} dm_rng;

DMP_FN void dm_rng_init(dm_rng* g, uint32_t seed, uint32_t point_id, uint64_t trajectory_id, uint32_t stream)
{
	g->key0 = seed;
	g->key1 = point_id;
	g->traj_lo = (uint32_t) trajectory_id;
	g->traj_hi = (uint32_t) (trajectory_id >> 32);
	g->stream = stream;
	g->block = 0;
	g->has_cached = 0;
	g->cached = 0.0;
}

/* Next uniform in [0, 1). */
DMP_FN double dm_uniform(dm_rng* g)
{
	if(g->has_cached)
	{
		g->has_cached = 0;
		return g->cached;
	}
	const uint32_t ctr[4] = {g->traj_lo, g->traj_hi, g->block, g->stream};
	const uint32_t key[2] = {g->key0, g->key1};
	uint32_t o[4];
	dm_philox4x32_10(ctr, key, o);
	g->block++;
	const uint64_t a = ((uint64_t) o[1] << 32) | o[0];
	const uint64_t b = ((uint64_t) o[3] << 32) | o[2];
	g->cached = (double) (b >> 11) * 0x1p-53;
	g->has_cached = 1;
	return (double) (a >> 11) * 0x1p-53;
}

#endif
