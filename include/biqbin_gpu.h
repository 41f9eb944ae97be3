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
/* Padded leading dimension of the register tile chosen for subproblems up to
 * nmax vertices (0 if nmax is too large). */
int bqb_ipm_ld_for(int nmax);

int bqb_ipm_batch(int nodes, const int *n, const double *L, double *X, double *phi, int *iters,
                  int *trials, int *status, double *y_out, double *alpha_trace, int trace_cap);

/* Diagnostic: one matrix, stop after max_iters iterations and return the
 * kernel's internal inv(Z) (Zi), inv(Zi o X) (Mi) and search direction dX of
 * the last iteration (each n x n column-major; any pointer may be NULL), and
 * in dbg_out (4*ld doubles, ld = padded size) the initial y, diag(Z),
 * diag(L) and (phi, psi, mu). */
int bqb_ipm_debug(int n, const double *L, int max_iters, double *X, double *Zi, double *Mi,
                  double *dX, double *y, double *phi, int *iters, double *alpha_trace,
                  double *dbg_out);

/* Diagnostic: apply one workgroup-distributed primitive to an n x n matrix
 * (column-major) with register tile `tile` (1604, 1607 or 3204 = P*100 + R):
 * op 0 inverse by symmetric sweep, 1 lower Cholesky factor, 2 Cholesky
 * positive-definiteness test, 3 symmetrisation (A + A')/2.  *ok receives
 * the primitive's success flag. */
int bqb_dist_selftest(int tile, int op, int n, const double *in, double *out, int *ok);

#ifdef __cplusplus
}
#endif
#endif
