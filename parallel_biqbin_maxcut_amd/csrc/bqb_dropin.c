/*
 * bqb_dropin.c -- the reference's node-evaluation entry points with their
 * own signatures, so a B&B host that calls Evaluate() (src/evaluate.c:11,
 * called from src/bab_functions.c:51 and :237) can link this library in
 * place of evaluate.c, bounding.c, bundle.c, cutting_planes.c, heuristic.c,
 * ipm_mc_pk.c, operators.c and qap_simulated_annealing.c.
 *
 * The incumbent is read and raised through the host's own Bab_LBGet()
 * (src/heap.c:15) and updateSolution() (src/bab_functions.c:114), the
 * parameters and the root gap through its globals `params` and `diff`
 * (src/global_var.h:6, :12).  All four are weak references and are used
 * only when the host defines all four; otherwise the library keeps its own
 * incumbent and root gap and uses the default parameters.
 *
 * The node is bounded on the GPU as a batch of one; its random streams are
 * keyed by (2020, evaluation number), 2020 being the reference's srand seed
 * (src/bab_functions.c:80).  A numerical failure terminates the process with
 * exit code 10, as the reference's MPI_Abort(MPI_COMM_WORLD, 10) does.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bqb_batch.h"
#include "bqb_rng.h"
#include "bqb_runtime.h"

extern double Bab_LBGet(void) __attribute__((weak));
extern int updateSolution(int *x) __attribute__((weak));
extern BiqBinParameters params __attribute__((weak));
extern double diff __attribute__((weak));

static struct {
    bqb_ctx *ctx;
    const double *SP_L;
    int n0;
    BiqBinParameters p;
    uint64_t evals;
    double LB;          /* used when the host has no Bab_LBGet */
    double diff;        /* used when the host has no `diff` */
    int sol[BQB_NMAX];
} g_drop;

/* The host is the reference's B&B when it defines all four symbols (the
 * generic names `params` and `diff` alone prove nothing). */
static int host_linked(void) { return Bab_LBGet && updateSolution && &params && &diff; }

static bqb_ctx *context_for(const BqbProblem *SP)
{
    BiqBinParameters p;
    if (host_linked()) p = params;
    else bqb_params_default(&p);
    if (g_drop.ctx && (g_drop.SP_L != SP->L || g_drop.n0 != SP->n || memcmp(&p, &g_drop.p, sizeof p))) {
        bqb_ctx_destroy(g_drop.ctx);
        g_drop.ctx = NULL;
    }
    if (!g_drop.ctx) {
        g_drop.ctx = bqb_ctx_create(SP->n, SP->L, &p, 1, 0, 0, 0, 0);
        if (!g_drop.ctx) {
            fprintf(stderr, "Error: bqb_ctx_create: %s\n", bqb_last_error());
            exit(10);
        }
        g_drop.SP_L = SP->L;
        g_drop.n0 = SP->n;
        g_drop.p = p;
    }
    return g_drop.ctx;
}

/* src/evaluate.c:123-138 */
double getFixedValue(BqbBabNode *node, BqbProblem *SP)
{
    const int n0 = SP->n, B = n0 - 1;
    double v = 0.0;
    for (int i = 0; i < B; ++i)
        for (int j = 0; j < B; ++j)
            if (node->xfixed[i] && node->xfixed[j])
                v += SP->L[j + (size_t)i * n0] * node->solX[i] * node->solX[j];
    return v;
}

/* src/evaluate.c:38-115: PP->L is the 1/4-scaled objective of the free
 * variables plus the homogenising row and column; PP->n its size.  (The
 * device builds the same matrix itself; this keeps PP consistent for a host
 * that reads it.) */
void createSubproblem(BqbBabNode *node, BqbProblem *SP, BqbProblem *PP)
{
    const int n0 = SP->n, B = n0 - 1;
    int n = 1;
    for (int i = 0; i < B; ++i) n += node->xfixed[i] ? 0 : 1;
    const int m = n - 1;
    int *freev = malloc(sizeof(int) * (m > 0 ? m : 1));
    if (!freev) { fprintf(stderr, "Error: memory allocation in createSubproblem\n"); exit(10); }
    for (int i = 0, a = 0; i < B; ++i)
        if (!node->xfixed[i]) freev[a++] = i;
    double *L = PP->L;
    double corner = 0.0;
    for (int a = 0; a < m; ++a) {
        const int i = freev[a];
        double fixed_row = 0.0;        /* sum of L(i, j) over j fixed to 1 */
        for (int j = 0; j < B; ++j)
            if (node->xfixed[j] && node->solX[j] == 1) fixed_row += SP->L[i + (size_t)j * n0];
        double row = 0.0;
        for (int b = 0; b < m; ++b) {
            const double v = SP->L[freev[b] + (size_t)i * n0];
            L[b + (size_t)a * n] = v;
            row += v;
        }
        L[a + (size_t)a * n] += 2 * fixed_row;
        row += 2 * fixed_row;
        L[m + (size_t)a * n] = row;
        L[a + (size_t)m * n] = row;
        corner += row;
    }
    L[m + (size_t)m * n] = corner;
    for (size_t e = 0; e < (size_t)n * n; ++e) L[e] *= 0.25;
    PP->n = n;
    free(freev);
}

/* src/bounding.c:24-342 for one node (the subproblem is rebuilt on the
 * device from SP and the node's fixings). */
double SDPbound(BqbBabNode *node, BqbProblem *SP, BqbProblem *PP, int rank)
{
    bqb_ctx *c = context_for(SP);
    const int B = SP->n - 1;
    const int host = host_linked();
    const double LB = host ? Bab_LBGet() : g_drop.LB;
    double *dp = host ? &diff : &g_drop.diff;
    const int is_root = rank == 0;
    const uint64_t seed = bqb_key(2020, g_drop.evals++);
    double bound, LB_out, diff_out;
    int improved, status, stats[8];
    int *xbest = malloc(sizeof(int) * B);
    if (!xbest) { fprintf(stderr, "Error: memory allocation in SDPbound\n"); exit(10); }
    if (bqb_bound_batch(c, 1, node->xfixed, node->solX, &LB, dp, &is_root, &seed, &bound,
                        node->fracsol, xbest, &LB_out, &improved, &diff_out, &status, stats)) {
        fprintf(stderr, "Error: node bound: %s\n", bqb_last_error());
        exit(10);
    }
    if (status) {
        fprintf(stderr, "Error: node bound failed with status %d\n", status);
        exit(10);
    }
    if (improved) {
        if (host) updateSolution(xbest);
        else if (LB_out > g_drop.LB) {
            g_drop.LB = LB_out;
            memcpy(g_drop.sol, xbest, sizeof(int) * B);
        }
    }
    if (is_root) *dp = diff_out;
    PP->n = stats[7];
    free(xbest);
    return bound;
}

/* src/evaluate.c:11-20 */
double Evaluate(BqbBabNode *node, BqbProblem *SP, BqbProblem *PP, int rank)
{
    createSubproblem(node, SP, PP);
    return SDPbound(node, SP, PP, rank);
}

/* The library-internal incumbent (hosts without Bab_LBGet/updateSolution). */
double bqb_dropin_incumbent(int *sol)
{
    if (sol && g_drop.n0 > 1) memcpy(sol, g_drop.sol, sizeof(int) * (g_drop.n0 - 1));
    return g_drop.LB;
}

/* Release the cached device context of the drop-in entry points. */
void bqb_dropin_release(void)
{
    if (g_drop.ctx) bqb_ctx_destroy(g_drop.ctx);
    memset(&g_drop, 0, sizeof g_drop);
}
