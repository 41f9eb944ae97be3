/* A B&B host in the shape of the reference's (src/main.c, bab_functions.c,
 * heap.c): it owns the incumbent (Bab_LBGet / updateSolution), the
 * parameters (`params`), the root gap (`diff`) and the global cut lists
 * (Cuts, Pent_Cuts, Hepta_Cuts), and calls the bounding entry points by
 * their reference signatures -- here served by libbiqbin_gpu.so.  With
 * -DOVERRIDE it also defines the heuristic override of the reference's
 * Python build (wrapped_heuristic), which must replace the device GW.
 *
 * Usage: dropin_host graph.rudy   -- prints one JSON object per step. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "biqbin_gpu.h"

BiqBinParameters params;
double diff;
typedef struct { int i, j, k, type; double value, y; } Triangle_Inequality;
typedef struct { int type; int perm[5]; double value, y; } Pentagonal_Inequality;
typedef struct { int type; int perm[7]; double value, y; } Heptagonal_Inequality;
Triangle_Inequality *Cuts;
Pentagonal_Inequality *Pent_Cuts;
Heptagonal_Inequality *Hepta_Cuts;

static double best = 0.0;
static int n0g, updates = 0, heur_calls = 0;
static const double *SPLg;
static int best_x[BQB_NMAX];

double Bab_LBGet(void) { return best; }

int updateSolution(int *x)
{
    ++updates;
    const double v = bqb_evaluate_solution(n0g, SPLg, x);
    if (v > best) {
        best = v;
        memcpy(best_x, x, sizeof(int) * (n0g - 1));
        return 1;
    }
    return 0;
}

#ifdef OVERRIDE
/* A deliberately plain heuristic: every free variable on side 1. */
double wrapped_heuristic(BqbProblem *P0, BqbProblem *P, BqbBabNode *node, int *x)
{
    (void)P;
    ++heur_calls;
    for (int i = 0; i < P0->n - 1; ++i) x[i] = node->xfixed[i] ? node->solX[i] : 1;
    return bqb_evaluate_solution(P0->n, P0->L, x);
}
#endif

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    int n0 = 0;
    double *adj = bqb_read_rudy(argv[1], &n0);
    if (!adj) return 3;
    double *SPL = malloc(sizeof(double) * n0 * n0);
    if (bqb_build_L(n0, adj, SPL)) return 4;
    n0g = n0;
    SPLg = SPL;
    bqb_params_default(&params);
    params.TriIneq = 10 * n0;               /* adjust_TriIneq (process_input.c:135-136) */
    BqbProblem SP = {SPL, n0, 0, 0, 0, 0};
    BqbProblem PP = {malloc(sizeof(double) * n0 * n0), n0, 0, 0, 0, 0};
    static BqbBabNode node;
    memset(&node, 0, sizeof node);
    const double root = Evaluate(&node, &SP, &PP, 0);
    printf("{\"step\": \"root\", \"bound\": %.17g, \"LB\": %.17g, \"updates\": %d, \"heur_calls\": %d, \"diff\": %.17g, "
           "\"n\": %d, \"frac0\": %.17g}\n", root, best, updates, heur_calls, diff, PP.n, node.fracsol[0]);
    for (int v = 0; v <= 1; ++v) {
        static BqbBabNode ch;
        memset(&ch, 0, sizeof ch);
        ch.xfixed[0] = 1;
        ch.solX[0] = v;
        const double b = Evaluate(&ch, &SP, &PP, 1);
        printf("{\"step\": \"child%d\", \"bound\": %.17g, \"LB\": %.17g, \"updates\": %d, \"heur_calls\": %d, \"n\": %d}\n",
               v, b, best, updates, heur_calls, PP.n);
    }
    /* fct_eval on the root subproblem with the host's global cut lists */
    memset(&node, 0, sizeof node);
    createSubproblem(&node, &SP, &PP);
    const int n = PP.n;
    Cuts = calloc(3, sizeof *Cuts);
    Pent_Cuts = calloc(1, sizeof *Pent_Cuts);
    Hepta_Cuts = calloc(1, sizeof *Hepta_Cuts);
    const int tri[3][4] = {{2, 1, 0, 1}, {5, 3, 1, 2}, {n - 1, 4, 2, 4}};
    for (int c = 0; c < 3; ++c) { Cuts[c].i = tri[c][0]; Cuts[c].j = tri[c][1]; Cuts[c].k = tri[c][2]; Cuts[c].type = tri[c][3]; }
    Pent_Cuts[0].type = 2;
    for (int a = 0; a < 5; ++a) Pent_Cuts[0].perm[a] = 3 * a + 1;
    Hepta_Cuts[0].type = 3;
    for (int a = 0; a < 7; ++a) Hepta_Cuts[0].perm[a] = 2 * a;
    PP.NIneq = 3;
    PP.NPentIneq = 1;
    PP.NHeptaIneq = 1;
    double gamma[5] = {0.125, 0.5, 0.0625, 0.25, 0.2};
    double *X = malloc(sizeof(double) * n * n), g[5];
    const double f = fct_eval(&PP, gamma, X, g);
    printf("{\"step\": \"fct_eval\", \"n\": %d, \"f\": %.17g, \"g\": [%.17g, %.17g, %.17g, %.17g, %.17g]}\n", n, f,
           g[0], g[1], g[2], g[3], g[4]);
    return 0;
}
