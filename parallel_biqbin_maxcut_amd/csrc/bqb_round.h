/* bqb_round.h -- FP64 arithmetic rounded toward +inf / -inf, built from
 * round-to-nearest operations and error-free transformations (usable on the
 * host and on the device; no change of the hardware rounding mode).
 *
 *   a + b: s = fl(a + b) and its exact error e (Knuth's TwoSum,
 *          a + b = s + e); the upward result is s if e <= 0, else the next
 *          double above s (|e| <= ulp(s)/2, so a + b lies in (s, next(s)]).
 *   a * b: p = fl(a b) and its exact error e = fma(a, b, -p); the same rule.
 *   a / k: q = fl(a / k) and the exact remainder r = fma(-q, k, a).
 * Each result equals the correctly rounded directed result, except where a
 * product or quotient lies below 2^-960: there the error need not be
 * representable, and the neighbour of the rounded-to-nearest result is
 * returned (still a bound on the correct side, one ulp wider at most).
 * Overflow is not handled (the bounds never come near it).  Operations are
 * written so that no contraction into fma can occur. */
#define BQB_TINY 0x1p-960
#ifndef BQB_ROUND_H
#define BQB_ROUND_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#ifdef __HIPCC__
#define BQB_RHD __host__ __device__ static inline
#else
#define BQB_RHD static inline
#endif

/* Contraction into fma is switched off inside each function only (a
 * file-scope pragma would reach every function of the including file). */
#if defined(__clang__)
#define BQB_NOCONTRACT _Pragma("clang fp contract(off)")
#else
#define BQB_NOCONTRACT
#endif

BQB_RHD uint64_t bqb_bits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
BQB_RHD double bqb_from_bits(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

/* next double toward +inf / -inf (infinities and NaN unchanged) */
BQB_RHD double bqb_next_up(double x)
{
    if (x != x || x == INFINITY) return x;
    if (x == 0.0) return bqb_from_bits(1);                 /* smallest subnormal */
    const uint64_t u = bqb_bits(x);
    return bqb_from_bits(x > 0.0 ? u + 1 : u - 1);
}
BQB_RHD double bqb_next_down(double x) { return -bqb_next_up(-x); }

BQB_RHD double bqb_two_sum_err(double a, double b, double s)
{
    BQB_NOCONTRACT
    const double bb = s - a;
    const double ab = s - bb;
    return (a - ab) + (b - bb);
}

BQB_RHD double bqb_add_ru(double a, double b)
{
    BQB_NOCONTRACT
    const double s = a + b;
    return bqb_two_sum_err(a, b, s) > 0.0 ? bqb_next_up(s) : s;
}
BQB_RHD double bqb_add_rd(double a, double b)
{
    BQB_NOCONTRACT
    const double s = a + b;
    return bqb_two_sum_err(a, b, s) < 0.0 ? bqb_next_down(s) : s;
}
BQB_RHD double bqb_sub_ru(double a, double b) { return bqb_add_ru(a, -b); }
BQB_RHD double bqb_sub_rd(double a, double b) { return bqb_add_rd(a, -b); }

BQB_RHD double bqb_mul_ru(double a, double b)
{
    BQB_NOCONTRACT
    const double p = a * b;
    if (fabs(p) < BQB_TINY) return (a == 0.0 || b == 0.0) ? p : bqb_next_up(p);
    return fma(a, b, -p) > 0.0 ? bqb_next_up(p) : p;
}
BQB_RHD double bqb_mul_rd(double a, double b)
{
    BQB_NOCONTRACT
    const double p = a * b;
    if (fabs(p) < BQB_TINY) return (a == 0.0 || b == 0.0) ? p : bqb_next_down(p);
    return fma(a, b, -p) < 0.0 ? bqb_next_down(p) : p;
}

/* a / k for k > 0 */
BQB_RHD double bqb_div_ru(double a, double k)
{
    BQB_NOCONTRACT
    const double q = a / k;
    if (fabs(q) < BQB_TINY) return a == 0.0 ? q : bqb_next_up(q);
    return fma(-q, k, a) > 0.0 ? bqb_next_up(q) : q;
}
BQB_RHD double bqb_div_rd(double a, double k)
{
    BQB_NOCONTRACT
    const double q = a / k;
    if (fabs(q) < BQB_TINY) return a == 0.0 ? q : bqb_next_down(q);
    return fma(-q, k, a) < 0.0 ? bqb_next_down(q) : q;
}

#endif
