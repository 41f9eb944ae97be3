/*
 * oracle.h -- CPU FP64 restatement of BiqBin's node-bounding path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, as the checker that the HIP path is compared against.
 *
 * It restates, with plain loops and its own small dense linear algebra (no
 * BLAS/LAPACK), the algorithms of the reference files:
 *   src/ipm_mc_pk.c            (orc_ipm_mc_pk)
 *   src/operators.c            (orc_op_B, orc_op_Bt)
 *   src/cutting_planes.c       (triangle / pentagonal / heptagonal separation)
 *   src/qap_simulated_annealing.c
 *   src/bundle.c               (bundle_method, fct_eval, lambda_eta, solve_lambda)
 *   src/heuristic.c            (runHeuristic_unpacked, GW_heuristic, mc_1opt)
 *   src/bounding.c             (SDPbound)
 *   src/evaluate.c             (createSubproblem, getFixedValue)
 *   src/bab_functions.c:96-134, 387-436, src/heap.c (B&B driver)
 *   src/process_input.c:259-344 (Laplacian with homogenising row)
 * Each function cites the file:line it follows.
 *
 * Deliberate, documented deviations (shared with the GPU path so that the two
 * can be compared call by call):
 *   - Randomness comes from a counter-based SplitMix64 stream keyed per node,
 *     per heuristic call and per SA chain (orc_rng.h) instead of glibc rand().
 *   - The simplex QP's (k+1)x(k+1) KKT system is solved through its Schur
 *     complement with a Cholesky factor instead of dsysv (same solution in
 *     exact arithmetic).
 *   - Every loop that the reference leaves unbounded has a cap far above what
 *     is ever reached; hitting a cap returns an error status.
 *   - The incumbent that SDPbound reads (Bab_LBGet in the reference) is the
 *     node's own copy: LB at node start raised by the node's own heuristic
 *     solutions.  With one node at a time this is exactly the reference.
 */
#ifndef BQB_ORACLE_H
#define BQB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Parameter block, same fields/order/defaults as the reference's PARAM_FIELDS
 * X-macro (src/biqbin.h:49-78). */
typedef struct orc_params {
    int init_bundle_iter;   /* 3  */
    int max_bundle_iter;    /* 15 */
    int triag_iter;         /* 5  */
    int pent_iter;          /* 5  */
    int hept_iter;          /* 5  */
    int max_outer_iter;     /* 20 */
    int extra_iter;         /* 10 */
    double violated_TriIneq;/* 1e-3 */
    int TriIneq;            /* 5000 */
    int adjust_TriIneq;     /* 1 */
    int PentIneq;           /* 5000 */
    int HeptaIneq;          /* 5000 */
    int Pent_Trials;        /* 60 */
    int Hepta_Trials;       /* 50 */
    int include_Pent;       /* 1 */
    int include_Hepta;      /* 1 */
    int root;               /* 0 */
    int use_diff;           /* 1 */
    int time_limit;         /* 0 */
    int branchingStrategy;  /* 1 = MOST_FRACTIONAL */
} orc_params;

/* Cut records, same layout as src/biqbin.h:81-107. */
typedef struct orc_tri  { int i, j, k, type; double value, y; } orc_tri;
typedef struct orc_pent { int type; int perm[5]; double value, y; } orc_pent;
typedef struct orc_hept { int type; int perm[7]; double value, y; } orc_hept;

/* Status codes (the reference calls MPI_Abort(...,10) in these places). */
enum {
    ORC_OK = 0,
    ORC_ERR_CHOL_Z = 1,       /* ipm_mc_pk.c:132-136 */
    ORC_ERR_INV_Z = 2,        /* ipm_mc_pk.c:139-143 */
    ORC_ERR_SCHUR = 3,        /* ipm_mc_pk.c:169-173, 226-230 */
    ORC_ERR_IPM_ITERS = 4,    /* cap on the unbounded IPM loop */
    ORC_ERR_LINESEARCH = 5,   /* cap on the unbounded backtracking loop */
    ORC_ERR_BUNDLE_FULL = 6,  /* bounding.c:312-315, bundle.c:140-143 */
    ORC_ERR_QP = 7,           /* bundle.c:552-556 */
    ORC_ERR_HEUR_CHOL = 8,    /* heuristic.c:70-74 */
    ORC_ERR_SA_STAGES = 9,
    ORC_ERR_ALLOC = 10
};

/* Statistics of one IPM call. */
typedef struct orc_ipm_stats {
    int iters;           /* main-loop iterations */
    int chol_trials;     /* line-search Cholesky trials, primal + dual */
    int status;
} orc_ipm_stats;

/* One node's bounding workspace: everything the reference keeps in process
 * globals (src/global_var.h) lives here instead. */
typedef struct orc_ws orc_ws;

orc_ws *orc_ws_create(int n0, int max_bundle, int max_tri, int max_pent, int max_hept,
                      const orc_params *p);
void orc_ws_destroy(orc_ws *w);

/* ---- counter RNG (shared definition with the GPU path) ---- */
uint64_t orc_mix64(uint64_t z);
uint64_t orc_key(uint64_t a, uint64_t b);
uint64_t orc_draw(uint64_t key, uint64_t ctr);

/* ---- dense helpers ---- */
/* Cholesky test/factor, upper (A = U'U), column-major, returns LAPACK-style info. */
int orc_chol_upper(double *A, int n);
/* Cholesky lower (A = LL'), zeroes the strictly upper triangle like heuristic.c:77-79. */
int orc_chol_lower(double *A, int n);

/* ---- the hot path ---- */
/* src/ipm_mc_pk.c:28-359 */
void orc_ipm_mc_pk(const double *L, int n, double *X, double *phi, orc_ipm_stats *st);
/* Extended form that also returns y and the alpha trace (2*iters entries). */
void orc_ipm_mc_pk_ex(const double *L, int n, double *X, double *y_out, double *phi,
                      orc_ipm_stats *st, double *alpha_trace, int trace_cap);

/* src/operators.c:51-132: y -= B(X) over the cut lists. */
void orc_op_B(int n, const orc_tri *tri, int ntri, const orc_pent *pent, int npent,
              const orc_hept *hept, int nhept, double *y, const double *X);
/* src/operators.c:140-481: X -= B'(t). */
void orc_op_Bt(int n, const orc_tri *tri, int ntri, const orc_pent *pent, int npent,
               const orc_hept *hept, int nhept, double *X, const double *t);

/* src/cutting_planes.c:20-39 */
double orc_eval_triangle(const double *X, int n, int type, int ii, int jj, int kk);
/* src/cutting_planes.c:48-127; returns max violation, fills list (cap = list_cap). */
double orc_get_violated_triangles(const double *X, int n, double threshold, int list_cap,
                                  orc_tri *list, int *list_size);
/* src/qap_simulated_annealing.c:12-190 with a keyed counter RNG. */
double orc_qap_sa(const int *H, int k, const double *X, int n, int *perm_out,
                  uint64_t key, int *status);
/* Pentagonal/heptagonal H tables (cutting_planes.c:242-244, 435-438). */
const int *orc_H_pent(int type);
const int *orc_H_hept(int type);
/* src/cutting_planes.c:224-309 / 416-503: trial loops over SA chains. */
double orc_get_violated_pent(const double *X, int n, int trials, int list_cap,
                             orc_pent *list, int *list_size, uint64_t key, int *status);
double orc_get_violated_hept(const double *X, int n, int trials, int list_cap,
                             orc_hept *list, int *list_size, uint64_t key, int *status);

/* src/heuristic.c:195-296 (works on x in {-1,1}); returns final cost. */
double orc_mc_1opt(int *x, const double *L, int n);
/* src/bab_functions.c:96-107: x'Lx over the top-left (n0-1)x(n0-1) block of SP_L. */
double orc_evaluate_solution(const double *SP_L, int n0, const int *sol);

/* src/evaluate.c:38-115: build the 1/4-scaled subproblem matrix, returns n. */
int orc_create_subproblem(const double *SP_L, int n0, const int *xfixed, const int *solX,
                          double *PP_L);
/* src/evaluate.c:123-138 */
double orc_get_fixed_value(const double *SP_L, int n0, const int *xfixed, const int *solX);

/* src/process_input.c:259-344: SP->L from an adjacency matrix (n0 x n0). */
void orc_build_L_from_adj(const double *Adj, int n0, double *SP_L);

/* src/bundle.c:390-653 */
int orc_solve_lambda(int k, const double *Q, const double *c, double *lambda);
/* src/bundle.c:292-373 */
int orc_lambda_eta(int m, int k, const double *zeta, const double *G, int ldG,
                   const double *gamma, double *dgamma, double *lambda, double *eta, double *t);

/* One node's bound (src/evaluate.c:11-20 -> src/bounding.c:24-342). */
typedef struct orc_node_in {
    const int *xfixed;   /* [n0-1] */
    const int *solX;     /* [n0-1] */
    double LB;           /* incumbent value at node start */
    double diff;         /* root gap used by the give-up test */
    int is_root;         /* rank == 0 semantics (sets diff) */
    uint64_t seed;       /* node RNG seed */
} orc_node_in;

typedef struct orc_node_out {
    double bound;        /* returned upper bound (f + fixedvalue) */
    double basic_bound;  /* basic SDP bound + fixedvalue */
    double LB;           /* LB after node (>= LB in) */
    int improved;        /* 1 if xbest beats the LB in */
    double diff;         /* diff out (only changed if is_root and not pruned early) */
    int prune, giveup;
    int outer_iters;     /* SDPbound main-loop count */
    int ipm_calls, ipm_iters;
    int max_cuts, max_bundle;
    int status;
} orc_node_out;

/* fracsol[n0-1] and xbest[n0-1] are written.  xbest is the best heuristic
 * solution seen by this node (valid when out->improved). */
void orc_sdpbound(orc_ws *w, const double *SP_L, int n0, const orc_node_in *in,
                  orc_node_out *out, double *fracsol, int *xbest);

/* Plain cutting-plane pieces exposed for call-by-call parity tests. */
/* fct_eval (bundle.c:242-275) on explicit cut lists. X[n*n], g[m]. */
double orc_fct_eval(const double *L, int n, const orc_tri *tri, int ntri,
                    const orc_pent *pent, int npent, const orc_hept *hept, int nhept,
                    const double *gamma, double *X, double *g, orc_ipm_stats *st);

/* ---- end-to-end branch and bound (single process) ---- */
typedef struct orc_bab_out {
    double best;          /* maximum cut value */
    double root_bound;
    int nodes;            /* evaluated nodes */
    int status;
    double seconds;
} orc_bab_out;
/* sol[n0-1] receives the best solution.  batch = nodes popped per round
 * (bounded against the same LB snapshot). */
void orc_bab_solve(const double *SP_L, int n0, const orc_params *p, uint64_t seed, int batch,
                   int *sol, orc_bab_out *out);

void orc_params_default(orc_params *p);

/* Attach a step trace to a workspace (rows of 8 doubles; cap rows) and read
 * how many rows the last orc_sdpbound wrote.  Row tags:
 *   1 bundle step   (count, f_test, f, t, bundle, m, serious)
 *   2 QP            (count, t, lambda_0, sum dgamma^2, bundle, m, 0)
 *   3 outer round   (count, NIneq, NPentIneq, NHeptaIneq, viol3, LB, bound)
 *   4 heuristic     (count, LB, value, 0, 0, 0, 0)
 *   5 step decision (count, del, f before the step, f_test, max lambda, k, serious)
 *   6 bundle element (count, i, lambda_i, max lambda, k, serious, 0), one
 *     row per element i < k of the step, before the purge                */
void orc_ws_set_trace(orc_ws *w, double *buf, int cap);
int orc_ws_trace_n(const orc_ws *w);

/* Test helper: the triangle update on a given X and cut list (orc_bound.c). */
double orc_update_triangles_on(const double *X, int n, const orc_params *p, orc_tri *cuts, int *ncuts,
                               int max_tri, const double *y, orc_tri *list, int *list_size);

/* Test helper: the heuristic of one node on a given X (see orc_heur.c). */
double orc_heuristic_on(const double *SP_L, int n0, const int *xfixed, const int *solX,
                        const double *PP_L, const double *X, int n, uint64_t key, int *x_out,
                        int *status);

/* Bound many independent nodes with up to nthreads threads (cpu baseline). */
void orc_sdpbound_many(const double *SP_L, int n0, const orc_params *p, int count,
                       const int *xfixed, const int *solX, const double *LB,
                       const double *diff, const uint64_t *seeds, int nthreads, double *bounds,
                       int *status);

#ifdef __cplusplus
}
#endif
#endif
