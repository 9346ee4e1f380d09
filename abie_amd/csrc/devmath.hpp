// devmath.hpp -- scalar arithmetic shared by the force and GR15 kernels.
//
// The library is compiled with -ffp-contract=off: every product and sum below
// rounds on its own unless an fma() is written out.  That is what makes the
// GR15 state updates bit-identical to the reference C (gcc -O3 -std=c99 on
// x86-64 never contracts).
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#define ABIE_HD __host__ __device__ __forceinline__

namespace abie {

// x / c, correctly rounded, for a constant c with rc = RN(1/c):
// q = RN(x*rc) is within one ulp of x/c, the residual x - q*c is exact in an
// FMA, and q + residual*rc rounds to RN(x/c) (Markstein, "IA-64 and
// elementary functions", 2000, Th. 8.10).  The residual is formed negated so
// that x = -0 keeps its sign.  Finite x only (x = +-inf gives NaN).
ABIE_HD double div_const(double x, double c, double rc) {
    const double q = x * rc;
    const double r = fma(q, c, -x);
    return fma(-r, rc, q);
}

// ---------------------------------------------------------------------------
// Double-double arithmetic (Dekker 1971; Knuth TAOCP vol. 2 §4.2.2): a value
// is hi + lo with |lo| <= ulp(hi)/2.  Used only for the step-size exponent,
// once per GR15 attempt, so speed is irrelevant; accuracy ~2^-100.
struct dd {
    double hi, lo;
};

ABIE_HD dd quick_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
ABIE_HD dd two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
ABIE_HD dd two_prod(double a, double b) {
    const double p = a * b;
    return {p, fma(a, b, -p)};
}
ABIE_HD dd dd_add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    const double e = s.lo + (a.lo + b.lo);
    return quick_two_sum(s.hi, e);
}
ABIE_HD dd dd_mul(dd a, dd b) {
    dd p = two_prod(a.hi, b.hi);
    const double e = p.lo + (a.hi * b.lo + a.lo * b.hi);
    return quick_two_sum(p.hi, e);
}
ABIE_HD dd dd_mul_d(dd a, double b) {
    dd p = two_prod(a.hi, b);
    return quick_two_sum(p.hi, p.lo + a.lo * b);
}
ABIE_HD dd dd_div_d(dd a, double b) {
    const double q1 = a.hi / b;
    const dd p = two_prod(q1, b);
    const double s = ((a.hi - p.hi) - p.lo) + a.lo;
    return quick_two_sum(q1, s / b);
}

// exp in double-double: x = k ln2 + r, r/2^10 by a 14-term Taylor series,
// then 10 squarings.
ABIE_HD dd dd_exp(dd x) {
    const dd ln2 = {0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};
    const double k = rint(x.hi / ln2.hi);
    dd r = dd_add(x, dd_mul_d({-ln2.hi, -ln2.lo}, k));
    r.hi = ldexp(r.hi, -10);
    r.lo = ldexp(r.lo, -10);
    dd term = {1.0, 0.0}, sum = {1.0, 0.0};
#pragma unroll 1
    for (int i = 1; i <= 14; ++i) {
        term = dd_div_d(dd_mul(term, r), (double)i);
        sum = dd_add(sum, term);
    }
#pragma unroll 1
    for (int i = 0; i < 10; ++i) sum = dd_mul(sum, sum);
    return {ldexp(sum.hi, (int)k), ldexp(sum.lo, (int)k)};
}

// log in double-double for x > 0 finite: one Newton step on exp from the
// double-precision log, y1 = y0 + (x e^{-y0} - 1).
ABIE_HD dd dd_log(double x) {
    const double y0 = log(x);
    const dd e = dd_exp({-y0, 0.0});
    const dd t = dd_mul_d(e, x);
    const dd d = dd_add(t, {-1.0, 0.0});
    return dd_add({y0, 0.0}, d);
}

// x^y rounded to nearest for x >= 0 (the only case GR15 needs: the step-size
// exponents 1/7 and 1/16 of integrator_gauss_radau15.c:48,316).  The
// double-double result is accurate to ~2^-100, so the final rounding agrees
// with any correctly rounded pow except within 2^-100 of a rounding boundary.
ABIE_HD double pow_pos(double x, double y) {
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return 0.0;
    if (x == 1.0) return 1.0;
    if (x > DBL_MAX) return x;  // +inf
    const dd l = dd_log(x);
    const dd z = dd_mul_d(l, y);
    const dd r = dd_exp(z);
    return r.hi + r.lo;
}

// x^(1/7) for x >= 0, to within an ulp or two, for the step controllers that
// are not bound to the reference's bits (the massive-only ones).  The binary
// exponent is split off exactly, x = z 2^(7q) with z in [2^-1, 2^6); an FP32
// seed for z^(1/7) (relative error ~2^-23) is refined by two Newton steps
// y <- (6 y + z / y^6) / 7 on y^7 = z, each of which squares the relative
// error (times 3): 2^-23 -> 2^-44 -> FP64 rounding.  About 40 instructions
// against the double-double pow_pos's thousands.  0, +inf and NaN map to
// themselves.
ABIE_HD double root7(double x) {
    if (!(x > 0.0) || x > DBL_MAX) return (x < 0.0) ? __builtin_nan("") : x;
    int e = 0;
    const double m = frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
    int q = e / 7, r = e - 7 * q;   // e = 7 q + r
    if (r < 0) {
        r += 7;
        q -= 1;
    }
    const double z = ldexp(m, r);  // [0.5, 64)
    double y = (double)exp2f(log2f((float)z) * (1.0f / 7.0f));
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double y3 = y * y * y;
        y = (6.0 * y + z / (y3 * y3)) / 7.0;
    }
    return ldexp(y, q);
}

}  // namespace abie
