/* Checks the directed-rounding helpers of bqb_round.h against the hardware
 * rounding modes of the host (fesetround), on random operands of mixed
 * magnitudes and on hand-picked edge cases.  Prints the number of
 * mismatches (0 expected) and exits with it. */
#include <fenv.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include "bqb_round.h"

#pragma STDC FENV_ACCESS ON

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint64_t next(void)
{
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double rnd(void)
{
    double m = (double)(next() >> 11) * 0x1.0p-53 * 2.0 - 1.0;
    int e = (int)(next() % 80) - 40;
    return ldexp(m, e);
}

static volatile double va, vb;
static double hw(int mode, int op)
{
    fesetround(mode);
    double r = op == 0 ? va + vb : op == 1 ? va * vb : va / vb;
    fesetround(FE_TONEAREST);
    return r;
}

/* 0 if got is the hardware result, or (tiny) a bound on its side at most
 * one step beyond it */
static int check(double got, double ref, int tiny, int dir)
{
    if (got == ref) return 0;
    if (!tiny) return 1;
    if (dir > 0) return !(got > ref && got == bqb_next_up(ref));
    return !(got < ref && got == bqb_next_down(ref));
}

int main(void)
{
    long bad = 0;
    const double edge[] = {0.0, -0.0, 1.0, -1.0, 0x1p-1074, -0x1p-1074, 0x1.fffffffffffffp-1, 3.0, 6.0, 1e300, -1e-300,
                           0x1.0000000000001p0, 1.0 / 3.0};
    const int ne = sizeof edge / sizeof edge[0];
    for (long it = 0; it < 2000000 + ne * ne; ++it) {
        double a, b;
        if (it < ne * ne) { a = edge[it / ne]; b = edge[it % ne]; }
        else { a = rnd(); b = rnd(); }
        va = a; vb = b;
        if (bqb_add_ru(a, b) != hw(FE_UPWARD, 0)) ++bad;
        if (bqb_add_rd(a, b) != hw(FE_DOWNWARD, 0)) ++bad;
        /* products and quotients below 2^-960: a bound on the correct side,
         * at most one step beyond the hardware's directed result */
        const int tiny_p = fabs(a * b) < BQB_TINY;
        bad += check(bqb_mul_ru(a, b), hw(FE_UPWARD, 1), tiny_p, +1);
        bad += check(bqb_mul_rd(a, b), hw(FE_DOWNWARD, 1), tiny_p, -1);
        if (b > 0.0) {
            const int tiny_q = fabs(a / b) < BQB_TINY;
            bad += check(bqb_div_ru(a, b), hw(FE_UPWARD, 2), tiny_q, +1);
            bad += check(bqb_div_rd(a, b), hw(FE_DOWNWARD, 2), tiny_q, -1);
        }
    }
    printf("%ld\n", bad);
    return bad != 0;
}
