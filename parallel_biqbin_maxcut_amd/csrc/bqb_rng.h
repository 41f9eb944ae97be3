// bqb_rng.h -- counter-based random streams (SplitMix64 finaliser, Steele,
// Lea & Flood 2014).  Every random draw of a node is a pure function of
// (node seed, stream tags, counter), so a chain or hyperplane can be
// evaluated by any lane in any order and reproduced exactly on the CPU.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define BQB_HD __host__ __device__ __forceinline__
#else
#define BQB_HD static inline
#endif

BQB_HD uint64_t bqb_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
BQB_HD uint64_t bqb_key(uint64_t a, uint64_t b)
{
    return bqb_mix64(bqb_mix64(a ^ 0x5851F42D4C957F2DULL) + 0x9E3779B97F4A7C15ULL * (b + 1));
}
BQB_HD uint64_t bqb_draw(uint64_t key, uint64_t ctr)
{
    return bqb_mix64(key + 0x9E3779B97F4A7C15ULL * (ctr + 1));
}
BQB_HD double bqb_u01(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }
BQB_HD int bqb_rmod(uint64_t r, int m) { return (int)((uint32_t)(r >> 32) % (uint32_t)m); }

/* The same residue for a divisor fixed over many draws (0 < m < 2^16):
 * the quotient of x < 2^32 by m estimated with a double reciprocal is off
 * by at most 2^-20 before truncation, so it is the true quotient or one
 * less, and one correction makes the residue exact. */
typedef struct bqb_divm { double rd; uint32_t m; } bqb_divm;
BQB_HD bqb_divm bqb_divm_make(int m)
{
    bqb_divm d;
    d.rd = 1.0 / (double)m;
    d.m = (uint32_t)m;
    return d;
}
BQB_HD int bqb_rmod_by(uint64_t r, bqb_divm d)
{
    const uint32_t x = (uint32_t)(r >> 32);
    const uint32_t q = (uint32_t)((double)x * d.rd);
    uint32_t rem = x - q * d.m;
    if (rem >= d.m) rem -= d.m;
    return (int)rem;
}

/* stream tags of one node (heuristic call h, separation round c) */
#define BQB_KEY_HEUR(seed, h) bqb_key((seed), 1000u + (uint64_t)(h))
#define BQB_KEY_PENT(seed, c) bqb_key((seed), 2000u + (uint64_t)(c))
#define BQB_KEY_HEPT(seed, c) bqb_key((seed), 3000u + (uint64_t)(c))
#define BQB_GW_STRIDE 1024
