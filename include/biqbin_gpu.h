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

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- types shared with the reference (binary compatible) --------------- */
/* BiqBin parameters: same fields, order, types and meaning as the
 * reference's PARAM_FIELDS X-macro (src/biqbin.h:49-78). */
typedef struct BiqBinParameters {
    int init_bundle_iter;
    int max_bundle_iter;
    int triag_iter;
    int pent_iter;
    int hept_iter;
    int max_outer_iter;
    int extra_iter;
    double violated_TriIneq;
    int TriIneq;
    int adjust_TriIneq;
    int PentIneq;
    int HeptaIneq;
    int Pent_Trials;
    int Hepta_Trials;
    int include_Pent;
    int include_Hepta;
    int root;
    int use_diff;
    int time_limit;
    int branchingStrategy;
} BiqBinParameters;

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

/* ---- per-node event flags (stats[8] of the batch API) ------------------ */
/* The bundle reached the context's capacity: the node ends as a give-up
 * with its bound at that moment (f + fixedvalue, a valid bound), where the
 * reference aborts the run (src/bounding.c:312-315). */
#define BQB_FLAG_BUNDLE_FULL 1
/* A separation produced more triangle / pentagonal / heptagonal cuts than
 * the context holds; the surplus was dropped (the bound stays valid, but
 * its trajectory departs from the reference's). */
#define BQB_FLAG_TRI_TRUNC 2
#define BQB_FLAG_PENT_TRUNC 4
#define BQB_FLAG_HEPT_TRUNC 8
/* An interior-point solve of the node with products on the bf16 matrix
 * cores failed and was repeated with FP64 products. */
#define BQB_FLAG_FP64_RETRY 16
/* An interior-point solve of the node by the blocked kernel (large
 * subproblems) failed and was repeated with the register-tile kernel. */
#define BQB_FLAG_REG_RETRY 32
/* The evaluation of the multipliers a node inherited from its parent failed
 * (interior-point status); the node went on from gamma = 0 as the reference
 * does. */
#define BQB_FLAG_WARM_FAILED 64
/* ints per node in the stats arrays of the batch API */
#define BQB_NSTATS 9

/* ---- library management ---------------------------------------------- */
int bqb_init(int device);                 /* select the HIP device (default 0) */
const char *bqb_last_error(void);
/* Sizes in bytes of the structures a binding mirrors, in this order:
 * BiqBinParameters, bqb_solve_options, bqb_solve_stats, bqb_trajectory and
 * the node state that bqb_ctx_download copies out (buffer id 0).  Writes
 * min(count, 5) entries and returns 5. */
int bqb_abi_sizes(int *sizes, int count);

/* Binary layout of the reference's Problem and BabNode
 * (src/biqbin_cpp_api.h:27-63, NMAX = 1024). */
#define BQB_NMAX 1024
typedef struct BqbProblem {
    double *L;          /* objective matrix, n x n column-major */
    int n;
    int NIneq;
    int NPentIneq;
    int NHeptaIneq;
    int bundle;
} BqbProblem;
typedef struct BqbBabNode {
    int xfixed[BQB_NMAX];       /* 1 where the variable is fixed */
    int solX[BQB_NMAX];         /* BabSolution sol: fixed values */
    double fracsol[BQB_NMAX];   /* written by the bound (branching) */
    int level;
    double upper_bound;
} BqbBabNode;

/* ---- drop-in: reference signatures ------------------------------------ */
/* A host built from the reference's main.c, bab_functions.c, heap.c,
 * process_input.c and allocate_free.c links this library in place of the
 * files of the bounding path.  The incumbent and root gap are read and
 * raised through that host's Bab_LBGet()/updateSolution() and its globals
 * params/diff (weak references; a host without them gets library-internal
 * ones, see bqb_dropin_incumbent).  The declarations are skipped when the
 * reference's own biqbin.h (include guard BIQBIN_H) is already included, so
 * its declarations, which have the same ABI, are the ones in force. */
#ifndef BIQBIN_H
/* replaces src/ipm_mc_pk.c:28 (declared src/biqbin.h:176) */
void ipm_mc_pk(double *L, int n, double *X, double *phi, int print);
/* replaces src/evaluate.c:11 (src/biqbin.h:155): createSubproblem + SDPbound */
double Evaluate(BqbBabNode *node, BqbProblem *SP, BqbProblem *PP, int rank);
/* replaces src/bounding.c:24 (src/biqbin.h:137) */
double SDPbound(BqbBabNode *node, BqbProblem *SP, BqbProblem *PP, int rank);
/* replaces src/evaluate.c:38 (src/biqbin.h:156) */
void createSubproblem(BqbBabNode *node, BqbProblem *SP, BqbProblem *PP);
/* replaces src/evaluate.c:123 (src/biqbin.h:157) */
double getFixedValue(BqbBabNode *node, BqbProblem *SP);
/* replaces src/bundle.c:242 (src/biqbin.h:140); reads the host's global cut
 * lists Cuts / Pent_Cuts / Hepta_Cuts (src/global_var.h:42-51) with the
 * counts in PP */
double fct_eval(const BqbProblem *PP, double *dual_gamma, double *X, double *g);
#endif
/* The library-internal incumbent of the drop-in entry points (value; sol
 * receives n0-1 entries if not NULL) and release of their device context. */
double bqb_dropin_incumbent(int *sol);
void bqb_dropin_release(void);

/* ---- host layer: input, objective, node queue ---------------------------- */
/* Defaults of the reference's PARAM_FIELDS (src/biqbin.h:49-78). */
void bqb_params_default(BiqBinParameters *p);
/* replaces readParameters (src/process_input.c:153-180): defaults, then
 * "name = value" lines.  0 on success. */
int bqb_read_params(const char *path, BiqBinParameters *p);
/* replaces readData (src/process_input.c:206-250): rudy format, 1-based
 * vertices, integer weights.  Returns a malloc'd nv x nv adjacency matrix
 * (free with free()) or NULL. */
double *bqb_read_rudy(const char *path, int *nv);
/* replaces process_adj_matrix (src/process_input.c:259-344): SP->L of an
 * nv-vertex graph (nv x nv, column-major); the last vertex is the one fixed
 * to 0, so the B&B has nv-1 variables. */
int bqb_build_L(int nv, const double *adj, double *L);
/* replaces evaluateSolution (src/bab_functions.c:96-107) for a given SP->L. */
double bqb_evaluate_solution(int n0, const double *SP_L, const int *sol);

typedef struct bqb_solve_stats {
    double best;          /* best cut value found (proven optimal unless stopped) */
    double root_bound;
    long long nodes;      /* evaluated nodes, root included */
    long long batches;    /* bound launches */
    int status;           /* first non-zero node status */
    int stopped;          /* params.time_limit reached */
    int max_queue;
    double seconds;       /* wall time */
    double bound_ms;      /* time inside the batched bound (device events) */
    long long bundle_full;     /* nodes given up because their bundle was full */
    long long cuts_truncated;  /* nodes whose separation overflowed a cut capacity */
} bqb_solve_stats;

/* Single-process branch and bound on one GPU; replaces the MPI
 * master/worker loop (src/main.c, src/bab_functions.c:39-69, 224-297) with
 * one node queue: up to `batch` best nodes are bounded per launch against
 * the incumbent at that moment.  sol receives n0-1 entries (the cut side of
 * every vertex but the last, which is on side 0).  Returns 0 on success. */
int bqb_solve(int n0, const double *SP_L, const BiqBinParameters *params, uint64_t seed, int batch,
              int *sol, bqb_solve_stats *stats);
/* The same with a bundle capacity per node (<= 0: the default 64); a node
 * whose bundle fills ends as a give-up (BQB_FLAG_BUNDLE_FULL) and is
 * branched. */
int bqb_solve_cap(int n0, const double *SP_L, const BiqBinParameters *params, uint64_t seed, int batch,
                  int max_bundle, int *sol, bqb_solve_stats *stats);
/* The same with the precision of the bounds (bqb_ctx_set_precision:
 * slices 2..4 bound on the bf16 matrix cores with certified bounds,
 * certify = 1 certifies FP64 bounds) and a learned cut model
 * (bqb_ctx_set_model; NULL: none). */
typedef struct bqb_mlp bqb_mlp;
int bqb_solve_ex(int n0, const double *SP_L, const BiqBinParameters *params, uint64_t seed, int batch,
                 int max_bundle, int slices, int certify, const bqb_mlp *model, float drop, int *sol,
                 bqb_solve_stats *stats);
/* All options of the node queue in one block (no counterpart in the
 * reference).  warm = 1: every child node inherits its parent's final cuts
 * whose multiplier is >= 1e-5 and that do not contain the newly fixed
 * variable, with their multipliers (bqb_ctx_set_warm).  graphs: 1 launch
 * trains as HIP graphs, 0 eager, -1 the library default.  ipm_warm = 1:
 * every bundle-step IPM starts from the slot's previous solve
 * (bqb_ctx_set_ipm_warm). */
typedef struct bqb_solve_options {
    int batch;            /* node slots (>= 1) */
    int max_bundle;       /* bundle capacity per node (<= 0: 64) */
    int slices, certify;  /* bqb_ctx_set_precision */
    const bqb_mlp *model; /* learned cut model or NULL */
    float drop;
    int warm;
    int graphs;
    int ipm_warm;
} bqb_solve_options;
int bqb_solve_opt(int n0, const double *SP_L, const BiqBinParameters *params, uint64_t seed,
                  const bqb_solve_options *opt, int *sol, bqb_solve_stats *stats);

/* ---- batch: interior-point solve of the basic SDP ---------------------- */
/* Solves max <L_b, X_b> s.t. diag(X_b) = e, X_b psd for `nodes` independent
 * matrices of sizes n[b] (<= 256).  L: the matrices concatenated, each
 * column-major n[b] x n[b]; X (same layout) receives the primal matrices;
 * phi[b] = e'y (the dual value, a valid upper bound).  iters/trials/status/
 * y_out (concatenated, length sum n[b]) / alpha_trace (nodes x 2*trace_cap)
 * may be NULL.  Follows src/ipm_mc_pk.c:28-359 per matrix. */
/* Padded leading dimension of the register tile chosen for subproblems up to
 * nmax vertices (0 if nmax is too large). */
int bqb_ipm_ld_for(int nmax);

int bqb_ipm_batch(int nodes, const int *n, const double *L, double *X, double *phi, int *iters,
                  int *trials, int *status, double *y_out, double *alpha_trace, int trace_cap);

/* Mixed-precision form of bqb_ipm_batch (no counterpart in the reference;
 * follows src/ipm_mc_pk.c:28-359 with its two n x n products,
 * ipm_mc_pk.c:186-189 and 253-257, on the bf16 matrix cores): every FP64
 * operand is split into `slices` (2..4) bf16 slices whose products the
 * matrix cores sum exactly in FP32; the slice products are combined in FP64
 * (error <= (4s+1) n 2^(-8s) times the operands' row/column maxima).  The
 * inversions, the Cholesky line searches and everything else stay FP64.
 * slices = 0 runs the FP64 products.  phi_cert[b] (may be NULL) receives a
 * certified upper bound of the basic SDP value: e'y plus n times a rigorous
 * bound on -lambda_min(Diag(y) - L) from an FP64 Cholesky factor and a
 * residual enclosed with directed rounding.  *ms (may be NULL): device time
 * of the IPM launch. */
int bqb_ipm_batch_mixed(int nodes, const int *n, const double *L, double *X, double *phi, double *phi_cert,
                        int *iters, int *trials, int *status, double *y_out, int slices, double *ms);

/* No counterpart in the reference.  The batched IPM of bqb_ipm_batch for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): every pointer is
 * device memory owned by the caller -- n[nodes] the sizes (each <= nmax),
 * L the matrices concatenated as in bqb_ipm_batch, X (same layout) the
 * primal matrices, phi/iters/status[nodes] -- and the work memory comes
 * from one library-owned device array of BQB_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends.  Calls needing more (bqb_ipm_noalloc_bytes)
 * are rejected with -1.  Work is enqueued on `stream` (a hipStream_t, NULL
 * for the default stream) and not waited for.  The function calls no
 * hipMalloc/hipFree.  All calls on one GPU share its array: calls that
 * overlap on one GPU race. */
#define BQB_NOALLOC_SCRATCH_BYTES (64u << 20)
size_t bqb_ipm_noalloc_bytes(int nodes, int nmax);
int bqb_ipm_batch_noalloc(int nodes, int nmax, const int *n, const double *L, double *X, double *phi,
                          int *iters, int *status, void *stream);

/* IPM tile selection (no counterpart in the reference).  Subproblem sizes
 * fall into BQB_TILE_CLASSES classes by their slab leading dimension
 * (64, 80, 96, 112, 128, 160, 192, 224, 256); each class runs either its
 * register tile (P*100 + R: 1604-1607, 3204-3208) or the blocked solve
 * with the same slabs (ld*100 + 1), which keeps the matrices in their slabs
 * and streams them through register micro-tiles (no spilled register
 * tile; faster above n ~ 190).  Failed blocked solves are repeated with
 * the register tile (BQB_FLAG_REG_RETRY).
 * bqb_set_ipm_tile forces one tile for contexts and batches created
 * afterwards whose size it covers (0 restores the table); -1 for an
 * unknown tile. */
#define BQB_TILE_CLASSES 9
int bqb_set_ipm_tile(int tile);
/* The per-class table (BQB_TILE_CLASSES entries): read, or install (each
 * entry the class's register tile or ld*100 + 1; -1 otherwise). */
int bqb_ipm_tile_table(int *tiles);
int bqb_set_ipm_tile_table(const int *tiles);
/* The tile autotuner: for every class whose slab size lies in
 * [nmin, nmax] (rounded up to its class), times `nodes` synthetic
 * subproblems of the class's largest size (Max-Cut L of G(n, 0.3) with
 * weights +-1) with the register tile and with the blocked solve, `reps`
 * launches each after one warm-up, and installs the faster in the table
 * (it takes another only if that is at least 5 % faster, so near-ties keep
 * the default).  ms_out (optional, 2 per class) receives the medians
 * (register, blocked); tiles (optional) the resulting table. */
int bqb_autotune_ipm(int nmin, int nmax, int nodes, int reps, double *ms_out, int *tiles);
/* Tile and product precision per class for contexts whose precision is
 * automatic (bqb_ctx_set_precision with slices = -1): tiles as above,
 * slices 0 (FP64 products) or 2..4 (bf16 matrix-core products).  Defaults:
 * FP64 products in every class but 193-224, where the register tile with
 * 4 bf16 slices is the fastest measured solve. */
int bqb_ipm_auto_table(int *tiles, int *slices);
int bqb_set_ipm_auto_table(const int *tiles, const int *slices);
int bqb_ipm_auto_for(int nmax, int *slices);
/* The autotuner for automatic precision: per class in [nmin, nmax], times
 * the register and blocked solves with FP64 products and with 4 bf16
 * slices (ms_out, optional: 4 per class in that order; a variant a class
 * cannot run reads 0) and installs the fastest in the automatic table (a
 * variant replaces the current entry only if at least 5 % faster). */
int bqb_autotune_ipm_auto(int nmin, int nmax, int nodes, int reps, double *ms_out, int *tiles, int *slices);

/* ---- batch: the full node bound (SDPbound) ----------------------------- */
/* A context holds the device arenas of up to max_nodes B&B nodes of one
 * instance: SP_L is the reference's SP->L (n0 x n0, column-major, built as
 * in src/process_input.c:259-344), params the BiqBin parameters.  Capacities
 * <= 0 take the defaults (bundle 64, triangle cuts 8192, pentagonal and
 * heptagonal cuts 2048 each). */
typedef struct bqb_ctx bqb_ctx;
bqb_ctx *bqb_ctx_create(int n0, const double *SP_L, const BiqBinParameters *params, int max_nodes,
                        int max_bundle, int max_tri, int max_pent, int max_hept);
void bqb_ctx_destroy(bqb_ctx *c);
size_t bqb_ctx_bytes(const bqb_ctx *c);
int bqb_ctx_ld(const bqb_ctx *c);
double bqb_ctx_last_ms(const bqb_ctx *c);

/* Bounds `nodes` (<= max_nodes) B&B nodes with the algorithm of
 * src/bounding.c:24-342 (Evaluate = createSubproblem + SDPbound,
 * src/evaluate.c:11-20).  Per node b: xfixed/solX rows of length n0-1
 * (src/biqbin_cpp_api.h:54-63), LB[b] the incumbent value at its start,
 * diff[b] the root gap used by the give-up test, is_root[b] the reference's
 * rank == 0 semantics, seeds[b] the node's random stream.  Outputs per node:
 * bound (f + fixedvalue), fracsol (n0-1), xbest (n0-1, best heuristic cut;
 * valid if improved), LB_out (LB raised by the node's own heuristic),
 * diff_out, status, and BQB_NSTATS stats ints (prune, giveup, outer
 * iterations, IPM calls, IPM iterations, max cuts, max bundle size, n,
 * BQB_FLAG_* events).  Any output pointer
 * may be NULL.  A node's results depend only on its own inputs. */
int bqb_bound_batch(bqb_ctx *c, int nodes, const int *xfixed, const int *solX, const double *LB,
                    const double *diff, const int *is_root, const uint64_t *seeds, double *bound,
                    double *fracsol, int *xbest, double *LB_out, int *improved, double *diff_out,
                    int *status, int *stats);
/* Same for any number of nodes: they stream through the max_nodes slots,
 * a finished node's slot is refilled after every outer round. */
int bqb_bound_stream(bqb_ctx *c, int count, const int *xfixed, const int *solX, const double *LB,
                     const double *diff, const int *is_root, const uint64_t *seeds, double *bound,
                     double *fracsol, int *xbest, double *LB_out, int *improved, double *diff_out,
                     int *status, int *stats);
/* Same through nctx (1..BQB_MAX_GROUPS) contexts of the same instance and
 * parameters, each driven by its own host thread on its own HIP stream: a
 * context takes the next node whenever one of its slots is free, so the
 * kernels of one slot group run in the idle tail of the other's launches.
 * Each launch still covers one context's slots.  Results are the same as
 * bqb_bound_stream's (a node's results depend only on its inputs).  The
 * contexts must not be used by other threads meanwhile. */
#define BQB_MAX_GROUPS 16
int bqb_bound_stream_multi(bqb_ctx *const *ctx, int nctx, int count, const int *xfixed, const int *solX,
                           const double *LB, const double *diff, const int *is_root, const uint64_t *seeds,
                           double *bound, double *fracsol, int *xbest, double *LB_out, int *improved,
                           double *diff_out, int *status, int *stats);

/* Incremental use (the node queue): push a node into a free slot (returns
 * the slot, -1 if none is free; tag >= 0 identifies it), advance every slot
 * by one outer round (nodes pushed since the last call first run their
 * basic bound), pop a finished node (returns 1, or 0 if none is waiting;
 * outputs as bqb_bound_batch for one node). */
int bqb_ctx_push(bqb_ctx *c, long long tag, const int *xfixed, const int *solX, double LB,
                 double diff, int is_root, uint64_t seed);
int bqb_ctx_advance(bqb_ctx *c);
int bqb_ctx_pop(bqb_ctx *c, long long *tag, double *bound, double *fracsol, int *xbest, double *LB_out,
                int *improved, double *diff_out, int *status, int *stats);
int bqb_ctx_busy(const bqb_ctx *c);       /* occupied slots */
int bqb_ctx_capacity(const bqb_ctx *c);   /* max_nodes */

/* Trajectory capture (the training data of a learned warm start; no
 * counterpart in the reference).  While capture is on, every node that
 * finishes leaves a record: its outer rounds (after each: f, bundle step t,
 * bound and the triangle / pentagonal / heptagonal cut counts; the rounds of
 * bounding.c:170-336), and its final cut set with the multipliers gamma
 * (laid out [triangles | pentagonals | heptagonals] as bounding.c:268-275),
 * for which f = fct_eval(gamma).  Cut records: tri 4 ints (i, j, k, type),
 * pentagonal and heptagonal 8 ints (type, then 5 or 7 vertices, padded);
 * vertices 0-based in the node's subproblem.  Records persist until
 * cleared; their pointers are owned by the context. */
typedef struct bqb_trajectory {
    long long tag;
    int n, rounds, ntri, npent, nhept, status;
    double f, t, bound, fixedvalue;
    const double *round_f, *round_t, *round_bound;   /* [rounds] */
    const int *round_cuts;                           /* [rounds*3] */
    const int *tri, *pent, *hept;                    /* [ntri*4], [npent*8], [nhept*8] */
    const double *gamma;                             /* [ntri+npent+nhept] */
} bqb_trajectory;
int bqb_ctx_capture(bqb_ctx *c, int on);             /* start (no node in flight) / stop */
int bqb_ctx_trajectory_count(const bqb_ctx *c);
int bqb_ctx_trajectory(const bqb_ctx *c, int i, bqb_trajectory *out);
void bqb_ctx_trajectory_clear(bqb_ctx *c);

/* Stepwise use of a context (tests and diagnostics): load a batch into
 * slots 0..nodes-1, run one phase (0-16 control phases; 100 basic IPM,
 * 101 IPM on L0, 102 heuristic,
 * 103 triangle update, 104/105 pentagonal/heptagonal update, 106 op_Bt
 * gather lists, 107 L0 = L - B'(gamma or gamma_test if arg), 108 bundle QP
 * (arg: an upper bound on every bundle size, 0 if unknown; a bound <= 16
 * selects the small-LDS QP kernel), 109 bundle step, 110/111 certificate of
 * the basic bound / of the bundle center of the nodes whose center moved
 * since their last certificate (mask phase 16), 112/113 learned model on the
 * cuts of the first / of a round's separation, 114 certificate of a
 * warm-started center right after its rebuild (the whole bound leaves it to
 * the round end, 111), 115 upload the nodes pushed since the last advance
 * and run only their basic sequence), run the whole bound,
 * read results, and copy one node's
 * device buffer (ids documented in bqb_sdpbound.c) to or from the host. */
int bqb_ctx_load(bqb_ctx *c, int nodes, const int *xfixed, const int *solX, const double *LB,
                 const double *diff, const int *is_root, const uint64_t *seeds);
int bqb_ctx_run(bqb_ctx *c, int phase, int arg);
int bqb_ctx_bound(bqb_ctx *c);
int bqb_ctx_results(bqb_ctx *c, double *bound, double *fracsol, int *xbest, double *LB_out,
                    int *improved, double *diff_out, int *status, int *stats);
int bqb_ctx_download(bqb_ctx *c, int id, int node, void *host, size_t bytes);
int bqb_ctx_upload(bqb_ctx *c, int id, int node, const void *host, size_t bytes);
int bqb_ctx_sync(bqb_ctx *c);
/* Heuristic plug point (src/heuristic.c:11-17 dispatches to
 * runHeuristic_unpacked, src/biqbin_cpp_api.h:65, or to a host override,
 * src/wrapper.cpp:210-226): a host function with runHeuristic_unpacked's
 * arguments (plus a user pointer) replaces the device GW heuristic of the
 * context at every point the reference calls runHeuristic; it receives
 * SP->L, SP->n, the node's PP->L and PP->n, its xfixed and sol.X, and x (n0-1
 * entries, fixed variables preset) to fill; the node's incumbent is raised
 * with x as updateSolution does.  NULL restores the device heuristic. */
typedef double (*bqb_heuristic_fn)(double *P0_L, int P0_N, double *P_L, int P_N, int *xfixed, int *solX,
                                   int *x, void *user);
int bqb_ctx_set_heuristic(bqb_ctx *c, bqb_heuristic_fn fn, void *user);

/* Precision of a context's bounds (no counterpart in the reference; call
 * while no node is in flight): slices = 0 runs the IPM products in FP64 (the
 * reference's arithmetic), 2..4 on the bf16 matrix cores as in
 * bqb_ipm_batch_mixed, -1 takes the tile and products of the context's size
 * class from the automatic table (bqb_ipm_auto_table) and certifies.  certify = 1 (implied by slices > 0) makes every
 * value the node prunes on and returns a certified upper bound: the basic
 * bound, and at the end of every outer round the bundle center if it moved
 * (serious steps, bundle.c:93-135) since its last certificate, get a
 * certificate  sum(gamma) + e'y + n * bound(-lambda_min(Diag(y) - L + B'gamma))
 * computed with directed rounding (src/bounding.c:109, 180, 197 prune on it). */
int bqb_ctx_set_precision(bqb_ctx *c, int slices, int certify);
/* bf16 matrix-core flops counted since the profile started (2 n^3 per slice
 * pair and product). */
int bqb_ctx_ipm_flops_mx(bqb_ctx *c, double *flops_mx);

/* ---- learned cut model (no counterpart in the reference) ----------------
 * A 3 x 256 MLP (32 -> 256 -> 256 -> 256 -> 2, ReLU) on the bf16 matrix
 * cores with FP32 accumulation, trained by the library's own forward and
 * backward kernels (Adam).  Input: the features of a cut (16 used, see
 * parallel_biqbin_maxcut_amd/learn.py); outputs: log1p(gamma / 0.05) of the
 * cut's final multiplier and the logit of the cut being active (gamma > 1e-5)
 * at the end of the node.  Parameters: BQB_MLP_NPARAM floats, layout
 * W1[256][32] b1 W2[256][256] b2 W3[256][256] b3 W4[16][256] b4[16] (rows 2..15
 * of W4 and b4 unused, zero). */
#define BQB_MLP_NPARAM 144144
bqb_mlp *bqb_mlp_create(const float *params, uint64_t seed);   /* params NULL: random init */
void bqb_mlp_destroy(bqb_mlp *m);
int bqb_mlp_get_params(bqb_mlp *m, float *out);
int bqb_mlp_set_params(bqb_mlp *m, const float *in);
/* feat [M][32] FP32 (rounded to bf16 on upload), out [M][2] */
int bqb_mlp_forward(bqb_mlp *m, int M, const float *feat, float *out);
/* one forward + backward over all M rows: grad [BQB_MLP_NPARAM], mean loss */
int bqb_mlp_gradient(bqb_mlp *m, int M, const float *feat, const float *target, float *grad, double *loss);
/* Adam, `epochs` passes over the rows in shuffled batches; loss_out[epochs] */
int bqb_mlp_train(bqb_mlp *m, int M, const float *feat, const float *target, int epochs, int batch, float lr,
                  uint64_t seed, double *loss_out);
/* Diagnostic: the features the bound computes for explicit cuts of one node
 * (X, L n x n column-major; cut records as bqb_fct_eval); out [m][32]. */
int bqb_cut_features(int n, const double *X, const double *L, int ntri, const int *tri, int npent,
                     const int *pent, int nhept, const int *hept, float round, float gap, float *out);
/* Use the model in a context's bounds: after every separation the new cuts
 * (src/cutting_planes.c:201, 392, 588 start them at 0) start from the
 * predicted multiplier times the predicted probability of being active, new
 * cuts predicted active with probability below `drop` are left out (0:
 * none), and the round's rebuild evaluation becomes the value f of the
 * warm-started center.  NULL detaches.  The model must outlive its use by
 * the context. */
int bqb_ctx_set_model(bqb_ctx *c, const bqb_mlp *model, float drop);
/* Training data: log the features of a quarter of the new cuts of every
 * separation (up to capacity records; 0 stops); records: feat [k][16],
 * key [k][12] = cut size, type, 7 vertices (padded), round, node seed (2 ints). */
int bqb_ctx_learn_record(bqb_ctx *c, int capacity);
int bqb_ctx_learn_records(bqb_ctx *c, int *count, float *feat, int *key);

/* IPM launch profile (no counterpart in the reference; for benchmarks):
 * start with max_launches > 0 (event pair per IPM launch, device flop
 * counter reset), stop with 0; read launches, summed ms and algorithmic
 * flops (6 n^3 per iteration + n^3/3 per line-search Cholesky test). */
int bqb_ctx_profile_ipm(bqb_ctx *c, int max_launches);
int bqb_ctx_ipm_stats(bqb_ctx *c, int *launches, double *ms_total, double *flops);

/* Launch trains as HIP graphs (no counterpart in the reference, whose loop
 * runs on the host, src/bounding.c:167-330).  on = 1: the basic sequence of
 * newly loaded nodes and each outer-round slice (bundle iterations, heuristic,
 * separation, rebuild) are captured once per shape (occupied slots, bundle
 * iterations in the slice, bundle-size class) and replayed with one
 * hipGraphLaunch; results are bitwise those of the eager launches.  Graphs
 * are not used while the IPM profile is on or a host heuristic is set.
 * Off by default. */
int bqb_ctx_set_graphs(bqb_ctx *c, int on);

/* Inherited cuts (no counterpart in the reference, whose every node starts
 * from no cuts, src/bounding.c:55-71).  With warm on (call while no node is
 * in flight), a node pushed with bqb_ctx_push may be given a cut set right
 * after the push (before the next advance): records as in bqb_fct_eval
 * (tri 4 ints i > j > k, type 1-4; pentagonal / heptagonal 8 ints: type,
 * then 5 or 7 distinct vertices, padded), vertices 0-based in the node's
 * subproblem (its free variables in order, then the last vertex), with
 * multipliers gamma >= 0 laid out [triangles | pentagonal | heptagonal].
 * The node runs the reference's basic sequence (basic bound, heuristic,
 * prune and give-up tests, first triangle separation, which keeps the
 * inherited cuts with gamma >= 1e-5); its bundle then starts from the
 * inherited multipliers, evaluated by fct_eval, when their value is below
 * the basic bound (any gamma >= 0 gives a valid bound), else from gamma = 0
 * as the reference's.  After every bqb_ctx_pop, bqb_ctx_last_cuts gives the
 * popped node's final cut set and multipliers (empty if it was pruned or
 * had none); the pointers stay valid until the next pop. */
int bqb_ctx_set_warm(bqb_ctx *c, int on);
/* Warm-started bundle IPMs (no counterpart in the reference, whose
 * ipm_mc_pk.c:65-99 starts every solve at X = I, y = 1.1 + column sums of
 * |L|).  With on = 1, each IPM of a bundle step (L - B'(gamma) -> X) starts
 * from the previous one of the same node: X0 = 0.95 X + 0.05 I and
 * y0_i = y_i + dL_ii + sum_{j != i} |dL_ij| + 0.01, dL the change of the
 * objective, which keeps the dual slack positive definite; if rounding
 * defeats that, the solve restarts at the reference's point.  Bounds stay
 * valid (every iterate's y is dual feasible) and agree with the cold solves'
 * within the IPM's 1e-2 gap; the trajectory, and so the node count, may
 * differ.  Register-tile solves only (n up to the largest register tile);
 * larger nodes start cold.  Allocates max_nodes * (ld^2 doubles + 1 int). */
int bqb_ctx_set_ipm_warm(bqb_ctx *c, int on);
int bqb_ctx_push_cuts(bqb_ctx *c, int slot, int ntri, const int *tri, int npent, const int *pent, int nhept,
                      const int *hept, const double *gamma);
int bqb_ctx_last_cuts(const bqb_ctx *c, int *ntri, const int **tri, int *npent, const int **pent, int *nhept,
                      const int **hept, const double **gamma);
/* The graph setting of contexts created afterwards (bqb_solve included);
 * returns the previous one. */
int bqb_set_default_graphs(int on);
/* Host-side counters since creation (or the last reset with reset = 1):
 * out[0] advance calls, out[1] outer-round slices, out[2] phases launched
 * eagerly, out[3] graph launches, out[4] phases replayed inside graphs,
 * out[5] blocking host synchronisations, out[6] graph instantiations. */
#define BQB_NCOUNTERS 7
int bqb_ctx_counters(bqb_ctx *c, long long *out, int reset);

/* fct_eval (src/bundle.c:242-275) on explicit cut lists, using node slot 0
 * of a context whose tile fits n: X = IPM solution for L - B'(gamma)
 * (n x n column-major in and out; X may be NULL), g = 1 - B(X) (m entries,
 * may be NULL), *f = phi + sum(gamma).  Cut records: tri 4 ints (i, j, k,
 * type 1-4), pent 6 ints (type 1-3, 5 vertices), hept 8 ints (type 1-4, 7
 * vertices), vertices 0-based.  *status receives the IPM status. */
int bqb_fct_eval(bqb_ctx *c, int n, const double *L, int ntri, const int *tri, int npent,
                 const int *pent, int nhept, const int *hept, const double *gamma, double *X,
                 double *g, double *f, int *status);

/* Diagnostic: one matrix, stop after max_iters iterations and return the
 * kernel's internal inv(Z) (Zi), inv(Zi o X) (Mi) and search direction dX of
 * the last iteration (each n x n column-major; any pointer may be NULL), and
 * in dbg_out (4*ld doubles, ld = padded size) the initial y, diag(Z),
 * diag(L) and (phi, psi, mu). */
int bqb_ipm_debug(int n, const double *L, int max_iters, double *X, double *Zi, double *Mi,
                  double *dX, double *y, double *phi, int *iters, double *alpha_trace,
                  double *dbg_out);

/* Diagnostic (no counterpart in the reference): copy `bytes` (multiple of
 * 16) between two device buffers `reps` times with a plain streaming kernel
 * (copy_probe_kernel); *ms = average time per copy.  Calibrates HBM byte
 * counters: each copy moves exactly 2 x bytes. */
int bqb_copy_probe(size_t bytes, int reps, double *ms);

/* Diagnostic (no counterpart in the reference): C = A B for n x n
 * column-major matrices (n <= 256) through the sliced bf16 matrix-core
 * product of the mixed-precision IPM with `slices` (1..4) slices. */
int bqb_mx_selftest(int n, int slices, const double *A, const double *B, double *C);

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
