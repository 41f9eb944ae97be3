/*
 * orc_heur.c -- oracle restatement of the primal heuristic:
 *   src/heuristic.c:19-115 (runHeuristic_unpacked), 119-186 (GW_heuristic),
 *   195-296 (mc_1opt), 304-317 (update_best); src/bab_functions.c:96-107.
 * Random numbers come from the keyed counter stream (see oracle.h).
 * TEST INFRASTRUCTURE ONLY.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_internal.h"

#define MAX_HEUR_ROUNDS 1000
#define GW_STRIDE 1024   /* draws reserved per hyperplane (N <= GW_STRIDE) */

static inline double u01(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }

double orc_evaluate_solution(const double *SP_L, int n0, const int *sol)
{
    int N = n0 - 1;
    double val = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            val += SP_L[j + i * n0] * sol[i] * sol[j];
    return val;
}

double orc_mc_1opt(int *x, const double *L, int N)
{
    double *Lx = calloc(N, sizeof(double));
    double *d = calloc(N, sizeof(double));
    double *delta = calloc(N, sizeof(double));
    int *I = calloc(N, sizeof(int));
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            Lx[i] += L[j + i * N] * x[j];
    double cost = 0.0;
    for (int i = 0; i < N; ++i) {
        d[i] = L[i + i * N];
        cost += x[i] * Lx[i];
        delta[i] = d[i] - x[i] * Lx[i];
    }
    double best = -ORC_BIG_NUMBER;
    int index = 0;
    for (int i = 0; i < N; ++i)
        if (delta[i] > best) { best = delta[i]; index = i; }
    int flips = 0;
    while (best > 0.001 && flips < 100 * N + 100) {
        ++flips;
        int num_I = 0;
        for (int j = 0; j < N; ++j)
            if (fabs(L[index + N * j]) > 0.001) I[num_I++] = j;
        if (x[index] > 0)
            for (int i = 0; i < num_I; ++i) Lx[I[i]] -= 2 * L[index + I[i] * N];
        else
            for (int i = 0; i < num_I; ++i) Lx[I[i]] += 2 * L[index + I[i] * N];
        x[index] *= -1;
        cost += 4 * best;
        for (int i = 0; i < N; ++i) delta[i] = d[i] - x[i] * Lx[i];
        best = -ORC_BIG_NUMBER;
        index = 0;
        for (int i = 0; i < N; ++i)
            if (delta[i] > best) { best = delta[i]; index = i; }
    }
    free(Lx); free(d); free(delta); free(I);
    return cost;
}

/* GW_heuristic (heuristic.c:119-186) with Z the lower Cholesky factor. */
static double gw(orc_ws *w, const double *Zc, const int *xfixed, const int *solX, int *x,
                 int num, uint64_t key)
{
    int N = w->n, n0 = w->n0;
    int *temp_x = malloc(sizeof(int) * N);
    int *sol = malloc(sizeof(int) * (n0 - 1));
    double *v = malloc(sizeof(double) * N);
    double best = -ORC_BIG_NUMBER;
    for (int count = 0; count < num; ++count) {
        for (int i = 0; i < N; ++i)
            v[i] = u01(orc_draw(key, (uint64_t)count * GW_STRIDE + i)) - 0.5;
        for (int i = 0; i < N; ++i) {
            double sca = 0.0;
            for (int j = 0; j < N; ++j) sca += v[j] * Zc[j * N + i];
            temp_x[i] = (sca < 0) ? -1 : 1;
        }
        orc_mc_1opt(temp_x, w->L, N);
        int index = 0;
        for (int i = 0; i < n0 - 1; ++i) {
            if (xfixed[i]) sol[i] = solX[i];
            else { sol[i] = (temp_x[index] + 1) / 2; ++index; }
        }
        double hv = orc_evaluate_solution(w->SP_L, n0, sol);     /* update_best */
        if (best < hv) {
            memcpy(x, sol, sizeof(int) * (n0 - 1));
            best = hv;
        }
    }
    free(temp_x); free(sol); free(v);
    return best;
}

/* runHeuristic_unpacked (heuristic.c:19-115) on the workspace's X and PP. */
double orc_run_heuristic(orc_ws *w, const int *xfixed, const int *solX, int *x, uint64_t key)
{
    int n = w->n, N = w->n0 - 1, nn = n * n;
    double *xh = malloc(sizeof(double) * n);
    int *temp_x = malloc(sizeof(int) * N);
    for (int i = 0; i < n; ++i)
        xh[i] = 2 * (int)(orc_draw(key, (uint64_t)i) >> 63) - 1;
    int index = 0;
    for (int i = 0; i < N; ++i) {
        if (xfixed[i]) temp_x[i] = solX[i];
        else { temp_x[i] = (int)((xh[index] + 1) / 2.0); ++index; }
    }
    double fh = orc_evaluate_solution(w->SP_L, w->n0, temp_x);
    int done = 0, round = 0;
    double heur_val = -ORC_BIG_NUMBER;
    memcpy(w->Z, w->X, sizeof(double) * nn);
    while (done < 2) {
        ++done;
        if (round >= MAX_HEUR_ROUNDS) break;
        uint64_t rkey = orc_key(key, (uint64_t)round + 1);
        ++round;
        if (orc_chol_lower(w->Z, n)) { w->status = ORC_ERR_HEUR_CHOL; break; }
        heur_val = gw(w, w->Z, xfixed, solX, x, w->n0, rkey);
        if (heur_val > fh) {
            done = 0;
            fh = heur_val;
            index = 0;
            for (int i = 0; i < N; ++i)
                if (!xfixed[i]) { xh[index] = 2 * x[i] - 1; ++index; }
            xh[n - 1] = -1.0;
        }
        double constant = 0.3 + 0.6 * u01(orc_draw(rkey, (uint64_t)w->n0 * GW_STRIDE));
        double alpha = 1.0 - constant;
        for (int i = 0; i < nn; ++i) w->Z[i] = alpha * w->X[i];
        for (int j = 0; j < n; ++j)            /* dsyr('L'): lower triangle += c xh xh' */
            for (int i = j; i < n; ++i)
                w->Z[i + j * n] = w->Z[i + j * n] + constant * xh[j] * xh[i];
    }
    free(xh); free(temp_x);
    return heur_val;
}

/* Test helper: run the heuristic of one node on a given X (n x n) and
 * subproblem matrix PP_L; x_out[n0-1] receives the last round's best cut and
 * the return value is its value.  Stream key as the node bound uses it. */
double orc_heuristic_on(const double *SP_L, int n0, const int *xfixed, const int *solX,
                        const double *PP_L, const double *X, int n, uint64_t key, int *x_out,
                        int *status)
{
    orc_params p;
    orc_params_default(&p);
    orc_ws *w = orc_ws_create(n0, 4, 4, 4, 4, &p);
    w->SP_L = SP_L;
    w->n = n;
    memcpy(w->L, PP_L, sizeof(double) * n * n);
    memcpy(w->X, X, sizeof(double) * n * n);
    w->status = 0;
    orc_run_heuristic(w, xfixed, solX, x_out, key);
    double v = orc_evaluate_solution(SP_L, n0, x_out);
    if (status) *status = w->status;
    orc_ws_destroy(w);
    return v;
}
