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

/* Cut records of the reference (src/biqbin.h:81-107) and its global cut
 * lists (src/global_var.h:42-51), read by fct_eval. */
typedef struct RefTri { int i, j, k, type; double value, y; } RefTri;
typedef struct RefPent { int type; int perm[5]; double value, y; } RefPent;
typedef struct RefHept { int type; int perm[7]; double value, y; } RefHept;
extern RefTri *Cuts __attribute__((weak));
extern RefPent *Pent_Cuts __attribute__((weak));
extern RefHept *Hepta_Cuts __attribute__((weak));

/* The host's heuristic, when it has one (src/heuristic.c:11-17): the
 * override of the reference's Python build (src/wrapper.cpp:210-226) or a
 * runHeuristic_unpacked of its own (src/biqbin_cpp_api.h:65).  Without
 * either, the device GW heuristic runs. */
extern double wrapped_heuristic(BqbProblem *P0, BqbProblem *P, BqbBabNode *node, int *x) __attribute__((weak));
extern double runHeuristic_unpacked(double *P0_L, int P0_N, double *P_L, int P_N, int *node_xfixed,
                                    int *node_sol_X, int *x) __attribute__((weak));

static double call_wrapped(double *P0_L, int P0_N, double *P_L, int P_N, int *xfixed, int *solX, int *x,
                           void *user)
{
    static BqbBabNode node;
    (void)user;
    BqbProblem P0 = {P0_L, P0_N, 0, 0, 0, 0}, P = {P_L, P_N, 0, 0, 0, 0};
    memset(&node, 0, sizeof node);
    memcpy(node.xfixed, xfixed, sizeof(int) * (P0_N - 1));
    memcpy(node.solX, solX, sizeof(int) * (P0_N - 1));
    return wrapped_heuristic(&P0, &P, &node, x);
}

static double call_unpacked(double *P0_L, int P0_N, double *P_L, int P_N, int *xfixed, int *solX, int *x,
                            void *user)
{
    (void)user;
    return runHeuristic_unpacked(P0_L, P0_N, P_L, P_N, xfixed, solX, x);
}

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
        bqb_heuristic_fn h = wrapped_heuristic ? call_wrapped : runHeuristic_unpacked ? call_unpacked : NULL;
        if (h && bqb_ctx_set_heuristic(g_drop.ctx, h, NULL)) {
            fprintf(stderr, "Error: %s\n", bqb_last_error());
            exit(10);
        }
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
    int improved, status, stats[BQB_NSTATS];
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

/* fct_eval's own context: any matrix up to its tile, cut capacities grown
 * on demand. */
static struct { bqb_ctx *ctx; int n0, tri, pent, hept; } g_fct;

static bqb_ctx *fct_context(int n, int ntri, int npent, int nhept)
{
    if (g_fct.ctx && bqb_ctx_ld(g_fct.ctx) >= n && ntri <= g_fct.tri && npent <= g_fct.pent && nhept <= g_fct.hept)
        return g_fct.ctx;
    if (g_fct.ctx) bqb_ctx_destroy(g_fct.ctx);
    g_fct.ctx = NULL;
    const int n0 = n > g_fct.n0 ? n : g_fct.n0;
    g_fct.tri = ntri > 8192 ? ntri : 8192;
    g_fct.pent = npent > 2048 ? npent : 2048;
    g_fct.hept = nhept > 2048 ? nhept : 2048;
    double *zero = calloc((size_t)n0 * n0, sizeof(double));
    BiqBinParameters p;
    bqb_params_default(&p);
    if (zero) g_fct.ctx = bqb_ctx_create(n0, zero, &p, 1, 0, g_fct.tri, g_fct.pent, g_fct.hept);
    free(zero);
    g_fct.n0 = n0;
    return g_fct.ctx;
}

/* src/bundle.c:242-275 with the host's global cut lists (Cuts, Pent_Cuts,
 * Hepta_Cuts, counts in PP): X = IPM solution for PP->L - B'(gamma),
 * g = 1 - B(X), returns phi + sum(gamma). */
double fct_eval(const BqbProblem *PP, double *dual_gamma, double *X, double *g)
{
    const int nt = PP->NIneq, np = PP->NPentIneq, nh = PP->NHeptaIneq;
    if ((nt && !(&Cuts && Cuts)) || (np && !(&Pent_Cuts && Pent_Cuts)) || (nh && !(&Hepta_Cuts && Hepta_Cuts))) {
        fprintf(stderr, "Error: fct_eval needs the host's cut lists (Cuts, Pent_Cuts, Hepta_Cuts)\n");
        exit(10);
    }
    int *tri = malloc(sizeof(int) * 4 * (nt ? nt : 1));
    int *pent = malloc(sizeof(int) * 6 * (np ? np : 1));
    int *hept = malloc(sizeof(int) * 8 * (nh ? nh : 1));
    if (!tri || !pent || !hept) { fprintf(stderr, "Error: memory allocation in fct_eval\n"); exit(10); }
    for (int q = 0; q < nt; ++q) {
        tri[4 * q] = Cuts[q].i; tri[4 * q + 1] = Cuts[q].j; tri[4 * q + 2] = Cuts[q].k; tri[4 * q + 3] = Cuts[q].type;
    }
    for (int q = 0; q < np; ++q) { pent[6 * q] = Pent_Cuts[q].type; memcpy(pent + 6 * q + 1, Pent_Cuts[q].perm, 5 * sizeof(int)); }
    for (int q = 0; q < nh; ++q) { hept[8 * q] = Hepta_Cuts[q].type; memcpy(hept + 8 * q + 1, Hepta_Cuts[q].perm, 7 * sizeof(int)); }
    bqb_ctx *c = fct_context(PP->n, nt, np, nh);
    double f = 0.0;
    int status = 0;
    if (!c || bqb_fct_eval(c, PP->n, PP->L, nt, tri, np, pent, nh, hept, dual_gamma, X, g, &f, &status)) {
        fprintf(stderr, "Error: fct_eval: %s\n", bqb_last_error());
        exit(10);
    }
    if (status) { fprintf(stderr, "Error: fct_eval: IPM status %d\n", status); exit(10); }
    free(tri); free(pent); free(hept);
    return f;
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
    if (g_fct.ctx) bqb_ctx_destroy(g_fct.ctx);
    memset(&g_fct, 0, sizeof g_fct);
}
