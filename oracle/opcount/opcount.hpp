/*
 * Op-counting build of the oracle -- MEASUREMENT INFRASTRUCTURE, not product
 * code.  opcount.cpp compiles the unchanged oracle sources as C++ with every
 * `double` replaced by `cdouble`, a double that counts the floating-point
 * operations performed on it:
 *
 *   + - * / and sqrt  1 each
 *   exp log sin cos acos erf  1 each (a transcendental counts once, whatever
 *                     its implementation costs)
 *   pow(x, 1/4)       2 (two square roots, as the reference computes it)
 *   comparisons, negation, fabs, min, max, conversions  0
 *
 * Literal-only constant expressions (3.0 / 32.0) are evaluated in plain
 * double and count nothing, as they are compile-time constants in the
 * reference too.  The oracle's attribution marks (oracle_internal.h) route
 * each operation to the part of a trajectory it belongs to: an RKF45 attempt,
 * a free-flight step inside or outside the Sun (its preparation and
 * post-processing), another step iteration (an optical-depth retry or a
 * trajectory's last iteration), a scattering, or the trajectory's own
 * bookkeeping (start, segment start, end).
 */
#ifndef DMS_OPCOUNT_HPP
#define DMS_OPCOUNT_HPP

#include <cstdint>
#include <type_traits>

enum
{
	OR_OPS_RK = 0,
	OR_OPS_INTERIOR = 1,
	OR_OPS_EXTERIOR = 2,
	OR_OPS_STEP_OTHER = 3,
	OR_OPS_SCATTER = 4,
	OR_OPS_TRAJECTORY = 5,
	OR_OPS_OUTSIDE = 6,
	OR_OPS_PARTS = 7
};

struct OpState
{
	uint64_t ops[OR_OPS_PARTS];
	uint64_t step_iterations[OR_OPS_PARTS]; /* step iterations by class */
	int stack[16];
	int depth;
	int in_step, step_class;
	uint64_t pending;
};

extern thread_local OpState g_ops;

inline void opc_count(uint64_t k)
{
	OpState& s = g_ops;
	const int top = s.depth > 0 ? s.stack[s.depth - 1] : OR_OPS_OUTSIDE;
	if(top == OR_OPS_RK || top == OR_OPS_SCATTER)
		s.ops[top] += k;
	else if(s.in_step)
		s.pending += k;
	else
		s.ops[top] += k;
}

inline void opc_push(int part)
{
	OpState& s = g_ops;
	if(s.depth < 16)
		s.stack[s.depth] = part;
	s.depth++;
}

inline void opc_pop() { g_ops.depth--; }

inline void opc_step_end()
{
	OpState& s = g_ops;
	if(s.in_step)
	{
		s.ops[s.step_class] += s.pending;
		s.step_iterations[s.step_class]++;
	}
	s.pending = 0;
	s.in_step = 0;
	s.step_class = OR_OPS_STEP_OTHER;
}

inline void opc_step_begin()
{
	opc_step_end();
	g_ops.in_step = 1;
}

inline void opc_step_class(int part) { g_ops.step_class = part; }

#define OR_OPCOUNT 1
#define OR_OPS_PUSH(part) opc_push(part)
#define OR_OPS_POP() opc_pop()
#define OR_OPS_STEP_BEGIN() opc_step_begin()
#define OR_OPS_STEP_CLASS(part) opc_step_class(part)
#define OR_OPS_STEP_END() opc_step_end()

struct cdouble
{
	double v;
	cdouble() = default;
	constexpr cdouble(double x) : v(x) {}
	template <class T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
	constexpr cdouble(T x) : v((double) x)
	{
	}
	explicit constexpr operator double() const { return v; }
	template <class T, typename std::enable_if<std::is_arithmetic<T>::value && !std::is_same<T, double>::value, int>::type = 0>
	explicit constexpr operator T() const
	{
		return (T) v;
	}
	cdouble operator-() const { return cdouble(-v); }
	cdouble operator+() const { return *this; }
	cdouble& operator+=(cdouble o)
	{
		opc_count(1);
		v += o.v;
		return *this;
	}
	cdouble& operator-=(cdouble o)
	{
		opc_count(1);
		v -= o.v;
		return *this;
	}
	cdouble& operator*=(cdouble o)
	{
		opc_count(1);
		v *= o.v;
		return *this;
	}
	cdouble& operator/=(cdouble o)
	{
		opc_count(1);
		v /= o.v;
		return *this;
	}
};

static_assert(std::is_trivially_copyable<cdouble>::value && sizeof(cdouble) == sizeof(double), "a double in layout");

inline cdouble operator+(cdouble a, cdouble b)
{
	opc_count(1);
	return cdouble(a.v + b.v);
}
inline cdouble operator-(cdouble a, cdouble b)
{
	opc_count(1);
	return cdouble(a.v - b.v);
}
inline cdouble operator*(cdouble a, cdouble b)
{
	opc_count(1);
	return cdouble(a.v * b.v);
}
inline cdouble operator/(cdouble a, cdouble b)
{
	opc_count(1);
	return cdouble(a.v / b.v);
}
inline bool operator<(cdouble a, cdouble b) { return a.v < b.v; }
inline bool operator>(cdouble a, cdouble b) { return a.v > b.v; }
inline bool operator<=(cdouble a, cdouble b) { return a.v <= b.v; }
inline bool operator>=(cdouble a, cdouble b) { return a.v >= b.v; }
inline bool operator==(cdouble a, cdouble b) { return a.v == b.v; }
inline bool operator!=(cdouble a, cdouble b) { return a.v != b.v; }
inline bool operator!(cdouble a) { return !a.v; }

/* the deterministic elementary functions on counted values */
inline int dm_isfinite(cdouble x) { return dm_isfinite(x.v); }
inline cdouble dm_fabs(cdouble x) { return cdouble(dm_fabs(x.v)); }
inline cdouble dm_min(cdouble a, cdouble b) { return cdouble(dm_min(a.v, b.v)); }
inline cdouble dm_max(cdouble a, cdouble b) { return cdouble(dm_max(a.v, b.v)); }
inline cdouble dm_clamp_unit(cdouble x) { return cdouble(dm_clamp_unit(x.v)); }
inline cdouble dm_clamp_cos(cdouble x) { return cdouble(dm_clamp_cos(x.v)); }
inline cdouble dm_sqrt(cdouble x)
{
	opc_count(1);
	return cdouble(dm_sqrt(x.v));
}
inline cdouble dm_pow_quarter(cdouble x)
{
	opc_count(2);
	return cdouble(dm_pow_quarter(x.v));
}
#define OPC_TRANSCENDENTAL(name)              \
	inline cdouble name(cdouble x)            \
	{                                         \
		opc_count(1);                         \
		return cdouble(name(x.v));            \
	}
OPC_TRANSCENDENTAL(dm_exp)
OPC_TRANSCENDENTAL(dm_log)
OPC_TRANSCENDENTAL(dm_sin)
OPC_TRANSCENDENTAL(dm_cos)
OPC_TRANSCENDENTAL(dm_acos)
OPC_TRANSCENDENTAL(dm_erf)
#undef OPC_TRANSCENDENTAL

#endif
