#!/usr/bin/env python3
"""Generate detmath_tables.h: constants and polynomial tables for the
deterministic elementary functions in detmath.h.

Every coefficient is computed in 60-digit arithmetic with mpmath and written
as an exact hexadecimal floating-point literal, so the C (gcc) and HIP
(clang, device) compilations read bit-identical constants.

Run from this directory:  python3 gen_detmath_tables.py > detmath_tables.h
"""
import mpmath as mp

mp.mp.dps = 60


def hx(v):
    return float(v).hex()


def split_bits(v, bits):
    """Round v to a double whose significand has only `bits` leading bits."""
    v = mp.mpf(v)
    if v == 0:
        return mp.mpf(0)
    e = mp.floor(mp.log(abs(v), 2))
    scale = mp.mpf(2) ** (e - bits + 1)
    return mp.nint(v / scale) * scale


out = []
emit = out.append
emit("/* Generated by gen_detmath_tables.py -- do not edit. */")
emit("#ifndef DMS_DETMATH_TABLES_H")
emit("#define DMS_DETMATH_TABLES_H")
emit("#ifdef __cplusplus")
emit("#define DM_TABLE static constexpr")
emit("#else")
emit("#define DM_TABLE static const")
emit("#endif")

pi = mp.pi
emit("#define DM_PI            %s" % hx(pi))
emit("#define DM_PI_2          %s" % hx(pi / 2))
emit("#define DM_SQRT_PI       %s" % hx(mp.sqrt(pi)))
emit("#define DM_2_SQRTPI      %s" % hx(2 / mp.sqrt(pi)))

# ln2 split for exp/log: hi part with 32 significant bits so k*hi is exact.
ln2 = mp.log(2)
ln2_hi = split_bits(ln2, 32)
ln2_lo = ln2 - ln2_hi
emit("#define DM_LN2_HI        %s" % hx(ln2_hi))
emit("#define DM_LN2_LO        %s" % hx(ln2_lo))
emit("#define DM_INV_LN2       %s" % hx(1 / ln2))

# pi/2 in four pieces of 23 significant bits (last piece full precision):
# k * piece is exact for |k| < 2^30.
r = pi / 2
parts = []
for i in range(3):
    p = split_bits(r, 23)
    parts.append(p)
    r = r - p
parts.append(r)
for i, p in enumerate(parts):
    emit("#define DM_PIO2_%d        %s" % (i + 1, hx(p)))
emit("#define DM_2_OVER_PI     %s" % hx(2 / pi))

# exp(r) on |r| <= ln2/2: Taylor coefficients 1/n!, n = 2..13.
emit("DM_TABLE double DM_EXP_C[12] = {")
emit(",\n".join("    %s" % hx(1 / mp.factorial(n)) for n in range(2, 14)))
emit("};")

# log: 2*atanh(s) = 2s + sum_{n>=1} 2 s^(2n+1)/(2n+1); coefficients 2/(2n+1), n=1..11.
emit("DM_TABLE double DM_LOG_C[11] = {")
emit(",\n".join("    %s" % hx(mp.mpf(2) / (2 * n + 1)) for n in range(1, 12)))
emit("};")

# sin(r) = r + r^3 * sum_{n>=0} S_n r^(2n), S_n = (-1)^(n+1)/(2n+3)!, n = 0..7
emit("DM_TABLE double DM_SIN_C[8] = {")
emit(",\n".join("    %s" % hx((-1) ** (n + 1) / mp.factorial(2 * n + 3)) for n in range(8)))
emit("};")
# cos(r) = 1 - r^2/2 + r^4 * sum_{n>=0} C_n r^(2n), C_n = (-1)^n/(2n+4)!, n = 0..7
emit("DM_TABLE double DM_COS_C[8] = {")
emit(",\n".join("    %s" % hx((-1) ** n / mp.factorial(2 * n + 4)) for n in range(8)))
emit("};")

# asin(y) = y + y^3 * sum_{n>=0} A_n y^(2n) on |y| <= 0.5,
# A_n = (2n+2)! / (4^(n+1) ((n+1)!)^2 (2n+3)), n = 0..27
def asin_coef(m):  # coefficient of y^(2m+1)
    return mp.factorial(2 * m) / (mp.mpf(4) ** m * mp.factorial(m) ** 2 * (2 * m + 1))
emit("DM_TABLE double DM_ASIN_C[28] = {")
emit(",\n".join("    %s" % hx(asin_coef(n + 1)) for n in range(28)))
emit("};")

# erf(x) = x * sum_n E_n x^(2n) on |x| < 0.84375, E_n = 2/sqrt(pi) (-1)^n / (n! (2n+1)), n = 0..18
emit("DM_TABLE double DM_ERF_SMALL_C[19] = {")
emit(",\n".join("    %s" % hx(2 / mp.sqrt(pi) * (-1) ** n / (mp.factorial(n) * (2 * n + 1)))
                for n in range(19)))
emit("};")

# erfcx(x) = exp(x^2) erfc(x): Taylor expansions of degree 16 about the
# centres of [0.84375,1.25], [1.25,1.75], ..., [5.75,6.25].
edges = [0.84375] + [1.25 + 0.5 * i for i in range(11)]


def erfcx(x):
    return mp.exp(x * x) * mp.erfc(x)


def taylor_erfcx(c, deg):
    # y' = 2 x y - 2/sqrt(pi);  y^(n+1) = 2 x y^(n) + 2 n y^(n-1)  (n >= 1)
    y = [erfcx(c), 2 * c * erfcx(c) - 2 / mp.sqrt(pi)]
    for n in range(1, deg):
        y.append(2 * c * y[n] + 2 * n * y[n - 1])
    return [y[n] / mp.factorial(n) for n in range(deg + 1)]


DEG = 16
emit("#define DM_ERFCX_NSEG %d" % (len(edges) - 1))
emit("#define DM_ERFCX_DEG  %d" % DEG)
emit("DM_TABLE double DM_ERFCX_CENTRE[%d] = {" % (len(edges) - 1))
emit(",\n".join("    %s" % hx((a + b) / 2) for a, b in zip(edges[:-1], edges[1:])))
emit("};")
emit("DM_TABLE double DM_ERFCX_C[%d][%d] = {" % (len(edges) - 1, DEG + 1))
rows = []
for a, b in zip(edges[:-1], edges[1:]):
    co = taylor_erfcx(mp.mpf((a + b) / 2), DEG)
    rows.append("    {" + ", ".join(hx(v) for v in co) + "}")
emit(",\n".join(rows))
emit("};")

# 10-point Gauss-Legendre rule on [-1, 1] (nodes ascending)
nodes = sorted(mp.polyroots(mp.taylor(lambda x: mp.legendre(10, x), 0, 10)[::-1],
                            maxsteps=200, extraprec=200))
nodes = [mp.re(n) for n in nodes]
weights = [2 / ((1 - x * x) * mp.diff(lambda t: mp.legendre(10, t), x) ** 2) for x in nodes]
emit("DM_TABLE double DM_GL10_X[10] = {" + ", ".join(hx(x) for x in nodes) + "};")
emit("DM_TABLE double DM_GL10_W[10] = {" + ", ".join(hx(w) for w in weights) + "};")

emit("#endif")
print("\n".join(out))
