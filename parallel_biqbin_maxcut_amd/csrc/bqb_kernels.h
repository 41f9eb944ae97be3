// bqb_kernels.h -- argument blocks and launchers shared between the HIP
// translation units and the C host code.  Plain C layout, no C++ types.
#ifndef BQB_KERNELS_H
#define BQB_KERNELS_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (same numbering as the CPU oracle).  The reference calls
 * MPI_Abort(MPI_COMM_WORLD, 10) at the matching places. */
enum {
    ERR_OK = 0,
    ERR_CHOL_Z = 1,        /* ipm_mc_pk.c:132-136 */
    ERR_INV_Z = 2,
    ERR_SCHUR = 3,         /* ipm_mc_pk.c:169-173, 226-230 */
    ERR_IPM_ITERS = 4,     /* iteration cap of the otherwise unbounded IPM loop */
    ERR_LINESEARCH = 5,    /* cap of the otherwise unbounded backtracking */
    ERR_BUNDLE_FULL = 6,   /* bounding.c:312-315, bundle.c:140-143 */
    ERR_QP = 7,            /* bundle.c:552-556 */
    ERR_HEUR_CHOL = 8,     /* heuristic.c:70-74 */
    ERR_SA_STAGES = 9,
    ERR_ALLOC = 10,
    ERR_SIZE = 11,         /* subproblem larger than the compiled tile */
    ERR_CUTS_FULL = 12     /* device cut arena full */
};

/* Batched IPM: node b has size n[b] <= ld; its L and X are column-major
 * ld x ld slabs at offset b*ld*ld (doubles), zero outside the leading n x n
 * block (the kernel keeps X zero there).  work holds BQB_IPM_WORK_SLABS(s)
 * ld x ld slabs per node: inv(Z), inv(Z o X), dX, and with mixed products
 * (mix_slices = s > 0) three more for the bf16 slices of the operands. */
#define BQB_IPM_WORK_SLABS(s) ((s) > 0 ? 6 : 3)
#define BQB_MIX_MAX_SLICES 4
typedef struct IpmBatchArgs {
    const int *n;
    const double *L;
    double *X;
    double *work;
    int ld;
    double *phi;
    int *iters;
    int *trials;
    int *status;
    const int *active;     /* optional mask (NULL = all) */
    double *y_out;         /* optional [nodes*ystride] */
    long ystride;
    double *alpha_trace;   /* optional [nodes * 2 * trace_cap] */
    int trace_cap;
    int max_iters;         /* 0: default cap; k > 0: stop after k iterations */
    double *dbg;           /* optional [nodes * 4 * ld]: initial y, zd, diag(L), and phi/psi/mu */
    unsigned long long *flops;  /* optional: += algorithmic FP64 flops of each node's solve */
    int mix_slices;        /* 0: FP64 products; 2..4: products on bf16 matrix cores with this many slices */
    unsigned long long *flops_mx;   /* optional (mixed): += bf16 matrix-core flops, 2 n^3 per slice pair */
    int wslabs;            /* work slabs per node in `work` (the blocked solve needs 4) */
    /* warm start from the slot's last solve (optional; needs y_out):
     * warm_mode 0 clears warm_ok; 1 starts from X, y_out and lprev where
     * warm_ok is set, keeps L in lprev and sets warm_ok to the solve's
     * success.  The blocked solve always starts cold and clears warm_ok. */
    double *lprev;         /* [nodes*ld*ld] the L of the slot's last solve */
    int *warm_ok;          /* [nodes] */
    int warm_mode;
} IpmBatchArgs;

/* Algorithmic FP64 flops of one IPM solve of size n: per iteration two SPD
 * inversions (n^3 each) and two n x n products (2 n^3 each), plus n^3/3 per
 * Cholesky test of the line searches. */
#define BQB_IPM_FLOPS(n, iters, trials) \
    ((double)(n) * (n) * (n) * (6.0 * (iters) + (trials) / 3.0))
/* FP64 part with the products on the matrix cores: the two inversions. */
#define BQB_IPM_FLOPS_MIXED(n, iters, trials) \
    ((double)(n) * (n) * (n) * (2.0 * (iters) + (trials) / 3.0))
/* FP64 flops of the two products of `iters` iterations (4 n^3 each). */
#define BQB_IPM_FLOPS_PRODUCTS(n, iters) ((double)(n) * (n) * (n) * 4.0 * (iters))
/* bf16 matrix-core flops of one mixed solve: two products of s(s+1)/2
 * slice pairs, 2 n^3 each, per iteration. */
#define BQB_IPM_FLOPS_MX(n, iters, s) ((double)(n) * (n) * (n) * 2.0 * (s) * ((s) + 1) * (iters))

/* The blocked solve for large subproblems (tile codes ld*100 + 1). */
hipError_t bqb_launch_ipm_blk(const IpmBatchArgs *a, int nodes, hipStream_t s);
/* The register tile whose slabs have leading dimension ld (a blocked tile
 * code's register counterpart), 0 if none. */
int bqb_ipm_reg_tile_for_ld(int ld);
/* The register-tile size rule alone (no override, no tuned table). */
int bqb_ipm_tile_reg_for(int nmax);
/* After a mixed-precision or blocked launch: solve again, with FP64
 * products and the given tile, the nodes whose solve failed (mask2:
 * [nodes] scratch; flags: optional per-node flag words at the given int
 * stride, which receive `flag`). */
hipError_t bqb_launch_ipm_retry(const IpmBatchArgs *a, int nodes, int tile, int *mask2, int *flags, int stride,
                                int flag, hipStream_t s);
hipError_t bqb_launch_cert_ipm(const int *nvec, const double *L, const double *y, long ystride, double *work,
                               int ld, int wslabs, const int *status, double *out, int nodes, hipStream_t s);

int bqb_ipm_tile_for(int nmax);
int bqb_ipm_class_ld(int c);
hipError_t bqb_launch_ipm_batch(const IpmBatchArgs *a, int nodes, int tile, hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif
