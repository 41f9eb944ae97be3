/* orc_internal.h -- oracle workspace layout (TEST INFRASTRUCTURE ONLY).
 * Holds what the reference keeps in the process globals of src/global_var.h. */
#ifndef BQB_ORC_INTERNAL_H
#define BQB_ORC_INTERNAL_H
#include "oracle.h"

#define ORC_BIG_NUMBER 1e+9

/* Constants of the reference's algorithm.  They can be overridden at build
 * time only so that the tests can show that their checks notice a changed
 * constant (tests/test_oracle_pins.py builds such variants); the oracle
 * itself is always built with these values. */
#ifndef ORC_SERIOUS_FRAC
#define ORC_SERIOUS_FRAC 0.05   /* serious step test, bundle.c:93 */
#endif
#ifndef ORC_T_STEP
#define ORC_T_STEP 1.01         /* t *= 1.01 (serious) / t /= 1.01 (null), bundle.c:112, 172 */
#endif
#ifndef ORC_PURGE_LAMBDA
#define ORC_PURGE_LAMBDA 0.01   /* bundle purge lambda < 0.01 max(lambda), bundle.c:118, 177 */
#endif
#ifndef ORC_T_NOCONV
#define ORC_T_NOCONV 0.95       /* lambda-eta loop without convergence, bundle.c:362-365 */
#endif
#ifndef ORC_LE_MAXIT
#define ORC_LE_MAXIT 50         /* bundle.c:362 */
#endif
#ifndef ORC_LE_TOL
#define ORC_LE_TOL 1e-5         /* bundle.c:358 */
#endif
#ifndef ORC_CUT_PURGE
#define ORC_CUT_PURGE 1e-5      /* purge of inactive cuts, cutting_planes.c:151, 336, 529 */
#endif
#ifndef ORC_T_REBUILD
#define ORC_T_REBUILD 1.05      /* t after a cut change, bounding.c:318 */
#endif

struct orc_ws {
    orc_params p;
    int n0;                    /* vertices of the original problem (SP->n) */
    int max_bundle, max_tri, max_pent, max_hept;
    /* subproblem (PP) */
    double *L;                 /* PP->L, n x n */
    int n;
    int NIneq, NPentIneq, NHeptaIneq, bundle;
    /* primal */
    double *X, *Z, *X_test, *X_bundle;
    /* dual */
    double *gamma, *dgamma, *gamma_test, *eta, *g, *lambda, *F;
    double *G;                 /* m x bundle, ld = m (grown on demand) */
    size_t G_cap;
    double f;
    /* cuts */
    orc_tri *Cuts, *List;
    orc_pent *Pent_Cuts, *Pent_List;
    orc_hept *Hepta_Cuts, *Hepta_List;
    /* node-local incumbent (Bab_LBGet / updateSolution) */
    const double *SP_L;
    double LB;
    int *xbest;
    int improved;
    /* stats */
    int ipm_calls, ipm_iters, status, max_cuts, max_bundle_seen;
    /* optional step trace (rows of ORC_TRACE_W doubles), for tests */
    int count;                 /* current outer iteration (trace only) */
    double *trace;
    int trace_cap, trace_n;
};

#define ORC_TRACE_W 8
/* Append one trace row (ignored when no trace buffer is attached). */
static inline void orc_trace(orc_ws *w, double tag, double a, double b, double c, double d,
                             double e, double f, double g)
{
    if (!w->trace || w->trace_n >= w->trace_cap) return;
    double *r = w->trace + (size_t)ORC_TRACE_W * w->trace_n++;
    r[0] = tag; r[1] = a; r[2] = b; r[3] = c; r[4] = d; r[5] = e; r[6] = f; r[7] = g;
}

int orc_ws_ensure_G(orc_ws *w, size_t need);
/* bundle.c:9-239 on the workspace */
int orc_bundle_method(orc_ws *w, double *t, int bdl_iter);
/* bundle.c:242-275 on the workspace cut lists */
double orc_ws_fct_eval(orc_ws *w, const double *gamma, double *X, double *g);
/* heuristic.c:19-115; returns the value of the last GW round, x[n0-1] out */
double orc_run_heuristic(orc_ws *w, const int *xfixed, const int *solX, int *x, uint64_t key);
#endif
