/*
 * Op-counting build of the oracle (MEASUREMENT INFRASTRUCTURE; see
 * opcount.hpp): the oracle's own sources, compiled with `double` meaning a
 * counted double, plus or_opcount_read to fetch the counts.  Built by this
 * directory's Makefile into liboracle_opcount.so; count_ops.py drives it.
 */
#include <math.h>
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "detmath.h"
#include "philox.h"
#include "opcount.hpp"

thread_local OpState g_ops = {};

#define double cdouble
#include "../oracle_solar.c"
#include "../oracle_engine.c"
#include "../oracle_ic.c"
#include "../oracle_driver.c"
#undef double

/* ops[OR_OPS_PARTS] then step_iterations[OR_OPS_PARTS] of the calling thread;
 * reset clears them. */
extern "C" void or_opcount_read(uint64_t* out, int reset)
{
	for(int i = 0; i < OR_OPS_PARTS; i++)
	{
		out[i] = g_ops.ops[i];
		out[OR_OPS_PARTS + i] = g_ops.step_iterations[i];
	}
	if(reset)
		g_ops = OpState{};
}
