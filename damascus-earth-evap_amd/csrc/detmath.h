/*
 * Deterministic elementary functions for the solar dark-matter trajectory path.
 *
 * Every function here is built from IEEE-754 correctly rounded operations
 * (+, -, *, /, sqrt), integer/bit manipulation and exact conversions only.
 * Compiled with -ffp-contract=off, the host (gcc/g++, x86-64 SSE2) and the
 * device (hipcc, gfx950) evaluate them to the same bits, which is what makes
 * trajectory-level replay between the CPU restatement and the FP64 kernel
 * bit-exact.  Accuracy against correctly rounded results is <= 2-3 ulp on the
 * domains the path uses (exp, log, sin, cos, erf, acos), measured by the
 * unit tests against 50-digit references.
 *
 * The polynomial tables live in detmath_tables.h (generated).
 * This header is valid C99 and HIP C++.
 */
#ifndef DMS_DETMATH_H
#define DMS_DETMATH_H

#include <stdint.h>
#include "detmath_tables.h"

#if defined(__HIPCC__) || defined(__HIP__)
#define DM_FN static inline __host__ __device__
#else
#define DM_FN static inline
#endif

#define DM_DBL_MIN 0x1p-1022
#define DM_SQRT2 0x1.6a09e667f3bcdp+0

DM_FN uint64_t dm_bits(double x)
{
	uint64_t u;
	__builtin_memcpy(&u, &x, 8);
	return u;
}

DM_FN double dm_from_bits(uint64_t u)
{
	double x;
	__builtin_memcpy(&x, &u, 8);
	return x;
}

DM_FN int dm_isfinite(double x) { return (dm_bits(x) & 0x7ff0000000000000ULL) != 0x7ff0000000000000ULL; }
DM_FN int dm_isnan(double x) { return x != x; }
DM_FN double dm_nan(void) { return dm_from_bits(0x7ff8000000000000ULL); }
DM_FN double dm_inf(void) { return dm_from_bits(0x7ff0000000000000ULL); }
DM_FN double dm_fabs(double x) { return dm_from_bits(dm_bits(x) & 0x7fffffffffffffffULL); }
/* Same NaN semantics as std::min / std::max. */
DM_FN double dm_min(double a, double b) { return (b < a) ? b : a; }
DM_FN double dm_max(double a, double b) { return (a < b) ? b : a; }
DM_FN double dm_sqrt(double x) { return __builtin_sqrt(x); }
#ifdef __cplusplus
/* Several correctly rounded quotients a / b with one divisor b.  The gfx950
 * IEEE division is: scale the operands (v_div_scale), y = rcp(b) refined by
 * two FMA Newton steps, q = a y, r = fma(-b, q, a), fma(r, y, q), then fix up
 * special operands.  Where the scaling and the fix-up are the identity -- b in
 * [2^-200, 2^200] and a zero or in [2^-800, 2^330]: operands, reciprocal,
 * quotient and remainder normal, exponent gap far below the 768 at which the
 * hardware scales -- y depends on b alone, so it is computed once
 * (dm_recip_shared) and each quotient (dm_div_shared) costs three dependent
 * operations instead of eleven, giving the same bits as a / b.  Callers check
 * that domain; the host build (CPU replay of the kernel code) divides. */
struct DmRecip
{
	double b, y;
};
DM_FN DmRecip dm_recip_shared(double b)
{
	DmRecip R;
	R.b = b;
#if defined(__HIP_DEVICE_COMPILE__)
	double y = __builtin_amdgcn_rcp(b);
	double e = __builtin_fma(-b, y, 1.0);
	y = __builtin_fma(y, e, y);
	e = __builtin_fma(-b, y, 1.0);
	R.y = __builtin_fma(y, e, y);
#else
	R.y = 1.0 / b;
#endif
	return R;
}
DM_FN double dm_div_shared(double a, const DmRecip& R)
{
#if defined(__HIP_DEVICE_COMPILE__)
	const double q = a * R.y;
	const double r = __builtin_fma(-R.b, q, a);
	return __builtin_fma(r, R.y, q);
#else
	return a / R.b;
#endif
}
#endif

/* a / b for a constant b > 0 with y = 1.0 / b (the correctly rounded
 * reciprocal, folded by the compiler): q = a y lies within one ulp of a / b,
 * r = fma(-b, q, a) is exact, and fma(r, y, q) is the correctly rounded
 * quotient (Markstein's theorem) where nothing overflows or underflows:
 * |a| zero or in [2^-900, 2^900] with |a / b| in the same range.  Three
 * operations instead of the division's eleven; infinite and NaN quotients
 * pass through from q.  Callers pass no -0 (whose quotient would be +0). */
DM_FN double dm_div_const(double a, double b, double y)
{
	const double q = a * y;
	const double r = __builtin_fma(-b, q, a);
	const double c = __builtin_fma(r, y, q);
	return dm_isfinite(q) ? c : q;
}

DM_FN double dm_clamp_unit(double v) { return dm_max(0.0, dm_min(1.0, v)); }
DM_FN double dm_clamp_cos(double v) { return dm_max(-1.0, dm_min(1.0, v)); }

/* x * 2^k, one rounding at most (only when the result is subnormal). */
DM_FN double dm_scale2(double x, int k)
{
	if(k > 1023)
	{
		x = x * 0x1p+1023;
		k -= 1023;
		if(k > 1023)
			k = 1023;
	}
	else if(k < -1022)
	{
		int k1 = k + 54;
		if(k1 < -1022)
			k1 = -1022;
		return (x * dm_from_bits((uint64_t) (k1 + 1023) << 52)) * 0x1p-54;
	}
	return x * dm_from_bits((uint64_t) (k + 1023) << 52);
}

/* exp(x): Cody-Waite reduction by ln2 and a degree-13 Taylor kernel. */
DM_FN double dm_exp(double x)
{
	if(x != x)
		return x;
	if(x > 709.782712893384)
		return dm_inf();
	if(x < -745.1332191019412)
		return 0.0;
	const int k = (int) (x * DM_INV_LN2 + (x < 0.0 ? -0.5 : 0.5));
	const double kd = (double) k;
	const double r = (x - kd * DM_LN2_HI) - kd * DM_LN2_LO;
	double q = DM_EXP_C[11];
	for(int i = 10; i >= 0; i--)
		q = q * r + DM_EXP_C[i];
	const double p = 1.0 + (r + r * r * q);
	return dm_scale2(p, k);
}

/* log(x): x = 2^e m, m in (sqrt(1/2), sqrt(2)], log(m) = 2 atanh(f/(2+f)). */
DM_FN double dm_log(double x)
{
	if(x != x)
		return x;
	if(x < 0.0)
		return dm_nan();
	if(x == 0.0)
		return -dm_inf();
	if(!dm_isfinite(x))
		return x;
	uint64_t u = dm_bits(x);
	int e = (int) ((u >> 52) & 0x7ff);
	if(e == 0)
	{
		x = x * 0x1p+54;
		u = dm_bits(x);
		e = (int) ((u >> 52) & 0x7ff) - 54;
	}
	e -= 1023;
	double m = dm_from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
	if(m > DM_SQRT2)
	{
		m = m * 0.5;
		e += 1;
	}
	const double f = m - 1.0;
	const double s = f / (2.0 + f);
	const double z = s * s;
	double R = DM_LOG_C[10];
	for(int i = 9; i >= 0; i--)
		R = R * z + DM_LOG_C[i];
	const double lm = f - s * (f - z * R);
	const double ed = (double) e;
	return ed * DM_LN2_HI + (ed * DM_LN2_LO + lm);
}

/* Reduction by pi/2 with a four-piece constant (exact products for |k| < 2^30). */
DM_FN double dm_reduce_pio2(double x, int* quadrant)
{
	const double t = x * DM_2_OVER_PI;
	const int64_t k = (int64_t) (t + (t < 0.0 ? -0.5 : 0.5));
	const double kd = (double) k;
	*quadrant = (int) (k & 3);
	return (((x - kd * DM_PIO2_1) - kd * DM_PIO2_2) - kd * DM_PIO2_3) - kd * DM_PIO2_4;
}

DM_FN double dm_sin_kernel(double r)
{
	const double z = r * r;
	double S = DM_SIN_C[7];
	for(int i = 6; i >= 0; i--)
		S = S * z + DM_SIN_C[i];
	return r + r * z * S;
}

DM_FN double dm_cos_kernel(double r)
{
	const double z = r * r;
	double C = DM_COS_C[7];
	for(int i = 6; i >= 0; i--)
		C = C * z + DM_COS_C[i];
	const double hz = 0.5 * z;
	const double w = 1.0 - hz;
	return w + (((1.0 - w) - hz) + z * (z * C));
}

DM_FN double dm_sin(double x)
{
	if(!dm_isfinite(x))
		return dm_nan();
	int q;
	const double r = dm_reduce_pio2(x, &q);
	switch(q)
	{
		case 0: return dm_sin_kernel(r);
		case 1: return dm_cos_kernel(r);
		case 2: return -dm_sin_kernel(r);
		default: return -dm_cos_kernel(r);
	}
}

/* sin(x) and cos(x) from one reduction: the same bits as dm_sin and dm_cos. */
DM_FN void dm_sincos(double x, double* sn, double* cs)
{
	if(!dm_isfinite(x))
	{
		*sn = dm_nan();
		*cs = dm_nan();
		return;
	}
	int q;
	const double r = dm_reduce_pio2(x, &q);
	const double s = dm_sin_kernel(r);
	const double c = dm_cos_kernel(r);
	switch(q)
	{
		case 0: *sn = s; *cs = c; break;
		case 1: *sn = c; *cs = -s; break;
		case 2: *sn = -s; *cs = -c; break;
		default: *sn = -c; *cs = s; break;
	}
}

DM_FN double dm_cos(double x)
{
	if(!dm_isfinite(x))
		return dm_nan();
	int q;
	const double r = dm_reduce_pio2(x, &q);
	switch(q)
	{
		case 0: return dm_cos_kernel(r);
		case 1: return -dm_sin_kernel(r);
		case 2: return -dm_cos_kernel(r);
		default: return dm_sin_kernel(r);
	}
}

/* erf(x): Maclaurin series below 0.84375, 1 - exp(-x^2) erfcx(x) with
 * piecewise degree-16 Taylor expansions of erfcx up to 6, 1 beyond. */
DM_FN double dm_erf(double x)
{
	if(x != x)
		return x;
	const double ax = dm_fabs(x);
	if(ax < 0.84375)
	{
		const double z = x * x;
		double p = DM_ERF_SMALL_C[18];
		for(int i = 17; i >= 0; i--)
			p = p * z + DM_ERF_SMALL_C[i];
		return x * p;
	}
	double res;
	if(ax >= 6.0)
		res = 1.0;
	else
	{
		int seg = (ax < 1.25) ? 0 : (int) ((ax - 1.25) * 2.0) + 1;
		if(seg > DM_ERFCX_NSEG - 1)
			seg = DM_ERFCX_NSEG - 1;
		const double t = ax - DM_ERFCX_CENTRE[seg];
		double p = DM_ERFCX_C[seg][DM_ERFCX_DEG];
		for(int j = DM_ERFCX_DEG - 1; j >= 0; j--)
			p = p * t + DM_ERFCX_C[seg][j];
		res = 1.0 - dm_exp(-(ax * ax)) * p;
	}
	return (x < 0.0) ? -res : res;
}

DM_FN double dm_asin_kernel(double y)
{
	const double z = y * y;
	double p = DM_ASIN_C[27];
	for(int i = 26; i >= 0; i--)
		p = p * z + DM_ASIN_C[i];
	return y + y * z * p;
}

DM_FN double dm_acos(double x)
{
	if(x != x || x > 1.0 || x < -1.0)
		return dm_nan();
	if(x > 0.5)
		return 2.0 * dm_asin_kernel(dm_sqrt(0.5 * (1.0 - x)));
	if(x < -0.5)
		return DM_PI - 2.0 * dm_asin_kernel(dm_sqrt(0.5 * (1.0 + x)));
	return DM_PI_2 - dm_asin_kernel(x);
}

/* x^(1/4) as two square roots (used for the RKF45 step-size factor). */
DM_FN double dm_pow_quarter(double x) { return dm_sqrt(dm_sqrt(x)); }

#endif
