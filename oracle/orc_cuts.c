/*
 * orc_cuts.c -- oracle restatement of the cut operators and cut separation:
 *   src/operators.c:51-132  (op_B)       src/operators.c:140-481 (op_Bt)
 *   src/cutting_planes.c:20-127 (triangle evaluation and scan)
 *   src/cutting_planes.c:224-309, 416-503 (pentagonal/heptagonal SA loops)
 *   src/qap_simulated_annealing.c:12-190
 * plus the counter-based RNG that replaces glibc rand().
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Every k-gonal cut (k = 3, 5, 7) is written as B_c(X) = coef * sum_{a<b}
 * s_ab X(p_a, p_b) <= 1 with s_ab = -e_a e_b, e a +-1 vector fixed by the
 * cut type and coef = 1, 1/2, 1/3.  The pairs are visited in the order
 * (0,1),(0,2),...,(1,2),... which is the order of the terms written out in
 * operators.c, so the floating-point sums are formed in the same order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define BIG_NUMBER 1e+9
#define SA_MAX_STAGES 150   /* qap_simulated_annealing.c:91 is unbounded */

/* ---------------- RNG: SplitMix64 finaliser, keyed counters ---------------- */
uint64_t orc_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
uint64_t orc_key(uint64_t a, uint64_t b)
{
    return orc_mix64(orc_mix64(a ^ 0x5851F42D4C957F2DULL) + 0x9E3779B97F4A7C15ULL * (b + 1));
}
uint64_t orc_draw(uint64_t key, uint64_t ctr)
{
    return orc_mix64(key + 0x9E3779B97F4A7C15ULL * (ctr + 1));
}
static inline double u01(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }
static inline int rint_mod(uint64_t r, int m) { return (int)((uint32_t)(r >> 32) % (uint32_t)m); }

/* ---------------- sign pattern of a cut ---------------- */
/* e vector of a triangle type (cutting_planes.c:24-36). */
static void tri_e(int type, int e[3])
{
    e[0] = e[1] = e[2] = 1;
    if (type == 2) e[2] = -1;
    else if (type == 3) e[1] = -1;
    else if (type == 4) e[0] = -1;
}
/* e vector of a pentagonal/heptagonal type: the first type-1 entries are -1
 * (cutting_planes.c:239-241, 431-434). */
static void kgon_e(int type, int k, int *e)
{
    for (int a = 0; a < k; ++a) e[a] = (a < type - 1) ? -1 : 1;
}

static double kgon_sum(const double *X, int N, const int *p, const int *e, int k)
{
    double s = 0.0;
    int first = 1;
    for (int a = 0; a < k; ++a)
        for (int b = a + 1; b < k; ++b) {
            double v = X[p[a] + p[b] * N];
            double t = (-e[a] * e[b] > 0) ? v : -v;
            if (first) { s = t; first = 0; }
            else s = s + t;
        }
    return s;
}

void orc_op_B(int n, const orc_tri *tri, int ntri, const orc_pent *pent, int npent,
              const orc_hept *hept, int nhept, double *y, const double *X)
{
    int e[7], p[7];
    for (int c = 0; c < ntri; ++c) {
        tri_e(tri[c].type, e);
        p[0] = tri[c].i; p[1] = tri[c].j; p[2] = tri[c].k;
        y[c] -= kgon_sum(X, n, p, e, 3);
    }
    for (int c = 0; c < npent; ++c) {
        kgon_e(pent[c].type, 5, e);
        y[ntri + c] -= 0.5 * kgon_sum(X, n, pent[c].perm, e, 5);
    }
    for (int c = 0; c < nhept; ++c) {
        kgon_e(hept[c].type, 7, e);
        y[ntri + npent + c] -= 1.0 / 3.0 * kgon_sum(X, n, hept[c].perm, e, 7);
    }
}

static void kgon_scatter(double *X, int N, const int *p, const int *e, int k, double temp)
{
    for (int a = 0; a < k; ++a)
        for (int b = a + 1; b < k; ++b) {
            double h = (e[a] * e[b] > 0) ? temp : -temp;   /* X -= s_ab * temp */
            X[p[a] + p[b] * N] += h;
            X[p[b] + p[a] * N] += h;
        }
}

void orc_op_Bt(int n, const orc_tri *tri, int ntri, const orc_pent *pent, int npent,
               const orc_hept *hept, int nhept, double *X, const double *t)
{
    int e[7], p[7];
    for (int c = 0; c < ntri; ++c) {
        tri_e(tri[c].type, e);
        p[0] = tri[c].i; p[1] = tri[c].j; p[2] = tri[c].k;
        kgon_scatter(X, n, p, e, 3, 0.5 * t[c]);
    }
    for (int c = 0; c < npent; ++c) {
        kgon_e(pent[c].type, 5, e);
        kgon_scatter(X, n, pent[c].perm, e, 5, 0.25 * t[ntri + c]);
    }
    for (int c = 0; c < nhept; ++c) {
        kgon_e(hept[c].type, 7, e);
        kgon_scatter(X, n, hept[c].perm, e, 7, 1.0 / 6.0 * t[ntri + npent + c]);
    }
}

/* ---------------- triangle separation ---------------- */
double orc_eval_triangle(const double *XX, int N, int type, int ii, int jj, int kk)
{
    switch (type) {
    case 1: return -XX[ii + jj * N] - XX[ii + kk * N] - XX[jj + kk * N] - 1.0;
    case 2: return -XX[ii + jj * N] + XX[ii + kk * N] + XX[jj + kk * N] - 1.0;
    case 3: return XX[ii + jj * N] - XX[ii + kk * N] + XX[jj + kk * N] - 1.0;
    default: return XX[ii + jj * N] + XX[ii + kk * N] - XX[jj + kk * N] - 1.0;
    }
}

/* cutting_planes.c:48-127: full scan, keep at most list_cap violations above
 * the threshold, replacing the current least violated entry when full. */
double orc_get_violated_triangles(const double *X, int N, double threshold, int list_cap,
                                  orc_tri *List, int *ListSize)
{
    int size = 0, least = 0;
    double leastv = BIG_NUMBER, maxall = -BIG_NUMBER;
    for (int type = 1; type <= 4; ++type)
        for (int ii = 0; ii < N; ++ii)
            for (int jj = 0; jj < ii; ++jj)
                for (int kk = 0; kk < jj; ++kk) {
                    double v = orc_eval_triangle(X, N, type, ii, jj, kk);
                    maxall = (v > maxall) ? v : maxall;
                    if (v > threshold) {
                        if (size < list_cap) {
                            List[size].type = type; List[size].i = ii;
                            List[size].j = jj; List[size].k = kk;
                            List[size].value = v;
                            if (size == 0 || v < leastv) { leastv = v; least = size; }
                            ++size;
                        } else if (v > leastv) {
                            List[least].type = type; List[least].i = ii;
                            List[least].j = jj; List[least].k = kk;
                            List[least].value = v;
                            leastv = BIG_NUMBER;
                            for (int c = 0; c < size; ++c)
                                if (List[c].value < leastv) { leastv = List[c].value; least = c; }
                        }
                    }
                }
    *ListSize = size;
    return maxall;
}

/* ---------------- pentagonal / heptagonal ---------------- */
static const int H5[3][25] = {
    {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {1, -1, -1, -1, -1, -1, 1, 1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1, 1, 1},
    {1, 1, -1, -1, -1, 1, 1, -1, -1, -1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1}};
static int H7[4][49];
static int H7_ready = 0;

const int *orc_H_pent(int type) { return H5[type - 1]; }
const int *orc_H_hept(int type)
{
    if (!H7_ready) {
        /* H = e e' with the first type-1 entries of e equal to -1
         * (cutting_planes.c:431-438). */
        for (int t = 1; t <= 4; ++t) {
            int e[7];
            kgon_e(t, 7, e);
            for (int a = 0; a < 7; ++a)
                for (int b = 0; b < 7; ++b)
                    H7[t - 1][b + 7 * a] = e[a] * e[b];
        }
        H7_ready = 1;
    }
    return H7[type - 1];
}

/* qap_simulated_annealing.c:12-190.  The chain draws its random numbers from
 * the keyed counter stream (key, 0), (key, 1), ... in the order the reference
 * calls rand(). */
double orc_qap_sa(const int *H, int k, const double *X, int n, int *ineq, uint64_t key,
                  int *status)
{
    uint64_t ctr = 0;
    int inner_iter = n;
    double reduce_temp = 0.6, increase_iter = 1.1;
    double t = k * k;
    double sumX = n;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            sumX += 2 * fabs(X[j + i * n]);
    t *= sumX;
    t /= (n * (n - 1));

    double cost = BIG_NUMBER;
    int *perm = malloc(sizeof(int) * n);
    int *perm_best = calloc(n, sizeof(int));
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = n - 1; i > 0; --i) {
        int r = rint_mod(orc_draw(key, ctr++), i);
        int tmp = perm[i]; perm[i] = perm[r]; perm[r] = tmp;
    }
    double t1 = t;
    int m1 = inner_iter;
    double sol = k;
    for (int i = 0; i < k; ++i)
        for (int j = i + 1; j < k; ++j)
            sol += 2 * H[j + i * k] * X[perm[j] + perm[i] * n];

    int not_done = 1, stages = 0;
    while (not_done > 0) {
        if (stages++ >= SA_MAX_STAGES) { if (status) *status = ORC_ERR_SA_STAGES; break; }
        not_done = 0;
        for (int it = 0; it < m1; ++it) {
            int i1 = rint_mod(orc_draw(key, ctr++), k);
            int i2 = rint_mod(orc_draw(key, ctr++), n);
            if (i2 < i1) { int tmp = i1; i1 = i2; i2 = tmp; }
            double delta = 0.0;
            for (int i = 0; i < k; ++i)
                delta += H[i + k * i1] * (X[perm[i] + n * perm[i2]] - X[perm[i] + n * perm[i1]]);
            if (i2 < k)
                for (int i = 0; i < k; ++i)
                    delta += H[i + k * i2] * (X[perm[i] + n * perm[i1]] - X[perm[i] + n * perm[i2]]);
            delta *= 2.0;
            double tv1 = 2.0 - 2.0 * X[perm[i2] + n * perm[i1]];
            double tv2 = 1.0;
            if (i2 < k) tv2 += -2 * H[i1 + k * i2] + 1.0;
            delta += tv1 * tv2;
            double sol_temp = sol + delta;
            int accept;
            if (delta > 0) {
                double dt1 = delta / t1;
                if (dt1 > 5) accept = 0;
                else accept = (u01(orc_draw(key, ctr++)) < exp(-dt1)) ? 1 : 0;
            } else
                accept = 1;
            if (accept) {
                if (fabs(delta) > 0.001) not_done = 1;
                int tmp = perm[i1]; perm[i1] = perm[i2]; perm[i2] = tmp;
                sol = sol_temp;
                if (sol < cost) {
                    cost = sol;
                    memcpy(perm_best, perm, sizeof(int) * n);
                }
            }
        }
        t1 *= reduce_temp;
        m1 = (int)round(m1 * increase_iter);
    }
    for (int i = 0; i < k; ++i) ineq[i] = perm_best[i];
    free(perm);
    free(perm_best);
    return cost;
}

/* cutting_planes.c:224-309 (k=5) and 416-503 (k=7).  The list keeps at most
 * list_cap violated cuts; the replacement branch keeps the reference's
 * behaviour exactly (its rescan starts from BIG_NUMBER with a '>' test, so the
 * replacement slot never moves after the first replacement). */
#define DEFINE_KGON_SEP(NAME, TYPE_T, K, NTYPES, HFUN)                                    \
double NAME(const double *X, int N, int trials, int list_cap, TYPE_T *List,              \
            int *ListSize, uint64_t key, int *status)                                     \
{                                                                                         \
    int size = 0, least = 0;                                                              \
    double leastv = -BIG_NUMBER, minall = BIG_NUMBER;                                     \
    int perm[K];                                                                          \
    for (int trial = 0; trial < trials; ++trial)                                          \
        for (int type = 1; type <= NTYPES; ++type) {                                      \
            uint64_t ck = orc_key(key, (uint64_t)(trial * NTYPES + type - 1));             \
            double v = orc_qap_sa(HFUN(type), K, X, N, perm, ck, status);                 \
            minall = (v < minall) ? v : minall;                                           \
            if (v < 1.0) {                                                                \
                if (size < list_cap) {                                                    \
                    List[size].type = type; List[size].value = v;                         \
                    for (int i = 0; i < K; ++i) List[size].perm[i] = perm[i];             \
                    if (size == 0 || v > leastv) { leastv = v; least = size; }            \
                    ++size;                                                               \
                } else if (v < leastv) {                                                  \
                    List[least].type = type; List[least].value = v;                       \
                    for (int i = 0; i < K; ++i) List[least].perm[i] = perm[i];            \
                    leastv = BIG_NUMBER;                                                  \
                    for (int c = 0; c < size; ++c)                                        \
                        if (List[c].value > leastv) { leastv = List[c].value; least = c; } \
                }                                                                         \
            }                                                                             \
        }                                                                                 \
    *ListSize = size;                                                                     \
    return minall;                                                                        \
}

DEFINE_KGON_SEP(orc_get_violated_pent, orc_pent, 5, 3, orc_H_pent)
DEFINE_KGON_SEP(orc_get_violated_hept, orc_hept, 7, 4, orc_H_hept)
