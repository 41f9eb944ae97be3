/*
 * biqbin_gpu.h -- C ABI of the MI355X node-bounding library
 * (libbiqbin_gpu.so).  Plain C types and pointers only.
 *
 * Two layers:
 *   1. Drop-in entry points with the reference's signatures (src/biqbin.h);
 *      each comment cites the reference declaration it replaces.
 *   2. Batch entry points for the single-process node queue: many B&B
 *      subproblems of different size are bounded per launch.
 *
 * Error convention: batch entry points return 0 on success and -1 on an API
 * or HIP failure (message via bqb_last_error()); per-subproblem numerical
 * failures are reported in a status array (codes below).  Drop-in entry
 * points keep the reference behaviour of terminating the process with exit
 * code 10 (the reference calls MPI_Abort(MPI_COMM_WORLD, 10)).
 */
#ifndef BIQBIN_GPU_H
#define BIQBIN_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes per subproblem ------------------------------------- */
#define BQB_OK 0
#define BQB_ERR_CHOL_Z 1       /* ipm_mc_pk.c:132-136 */
#define BQB_ERR_SCHUR 3        /* ipm_mc_pk.c:169-173, 226-230 */
#define BQB_ERR_IPM_ITERS 4
#define BQB_ERR_LINESEARCH 5
#define BQB_ERR_BUNDLE_FULL 6  /* bounding.c:312-315, bundle.c:140-143 */
#define BQB_ERR_QP 7           /* bundle.c:552-556 */
#define BQB_ERR_HEUR_CHOL 8    /* heuristic.c:70-74 */
#define BQB_ERR_SA_STAGES 9
#define BQB_ERR_ALLOC 10
#define BQB_ERR_SIZE 11
#define BQB_ERR_CUTS_FULL 12

/* ---- library management ---------------------------------------------- */
int bqb_init(int device);                 /* select the HIP device (default 0) */
const char *bqb_last_error(void);

/* ---- drop-in: reference signatures ------------------------------------ */
/* replaces src/ipm_mc_pk.c:28 (declared src/biqbin.h:176) */
void ipm_mc_pk(double *L, int n, double *X, double *phi, int print);

/* ---- batch: interior-point solve of the basic SDP ---------------------- */
/* Solves max <L_b, X_b> s.t. diag(X_b) = e, X_b psd for `nodes` independent
 * matrices of sizes n[b] (<= 160).  L: the matrices concatenated, each
 * column-major n[b] x n[b]; X (same layout) receives the primal matrices;
 * phi[b] = e'y (the dual value, a valid upper bound).  iters/trials/status/
 * y_out (concatenated, length sum n[b]) / alpha_trace (nodes x 2*trace_cap)
 * may be NULL.  Follows src/ipm_mc_pk.c:28-359 per matrix. */
int bqb_ipm_batch(int nodes, const int *n, const double *L, double *X, double *phi, int *iters,
                  int *trials, int *status, double *y_out, double *alpha_trace, int trace_cap);

#ifdef __cplusplus
}
#endif
#endif
