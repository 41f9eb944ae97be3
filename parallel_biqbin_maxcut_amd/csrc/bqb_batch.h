/* bqb_batch.h -- device layout of a batch of B&B nodes being bounded
 * together (shared by the HIP kernels and the C host driver).
 *
 * Everything the reference keeps in process globals for the one node it
 * bounds at a time (src/global_var.h, sized by src/allocate_free.c:8-42)
 * exists once per node slot here, in HBM, at fixed strides.  Matrices are
 * zero-padded column-major ld x ld slabs (ld = padded size of the IPM tile).
 */
#ifndef BQB_BATCH_H
#define BQB_BATCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#include "../../include/biqbin_gpu.h"

/* Parameter block (the reference's PARAM_FIELDS layout). */
typedef BiqBinParameters BqbParams;

/* Scalar state of one node slot (bounding.c:24-342 locals and the globals
 * f, diff and the incumbent it reads). */
typedef struct NodeState {
    int n;              /* SDP size n = n0 - #fixed */
    int status;
    int active;         /* still inside SDPbound */
    int prune, giveup, done;
    int count;          /* outer iterations (bounding.c:170) */
    int bdl_iter;
    int NIneq, NPentIneq, NHeptaIneq, bundle;
    int sep_tri, sep_pent, sep_hept;    /* this round's separation decisions */
    int heur_idx;       /* heuristic call index (random stream) */
    int improved;       /* node found a solution better than its LB in */
    int ipm_calls, ipm_iters;
    int max_cuts, max_bundle_seen;
    int is_root;
    int serious;        /* last bundle step was serious */
    int cuts_changed;   /* op_Bt gather lists must be rebuilt */
    int off_pent, off_hept;   /* gamma offsets of the pentagonal / heptagonal
                                 multipliers before this round's purge */
    int stage;          /* slot life cycle: BQB_STAGE_* */
    int bdl_left;       /* bundle iterations left in the current outer round */
    int rphase;         /* outer round: 0 to begin, 1 bundle iterations, 2 round end */
    int flags;          /* BQB_FLAG_* events of the node (reported in stats[8]) */
    int warm;           /* loaded with inherited cuts and multipliers: its first
                           bundle center is evaluated at them (until done) */
    int cert_stale;     /* the center moved (serious step) since f_cert was
                           computed; its y is in ycen */
    int nsurv_tri, nsurv_pent, nsurv_hept;   /* with a learned model: cuts of
                           each list before the last separation's new ones */
    double fixedvalue, f, f_test, t, bound, basic_bound, oldf;
    double LB, diff, viol3, viol5, gap, del, lmax;
    double f_cert;      /* certified upper bound of f (when certifying) */
    uint64_t seed;
} NodeState;

/* Slot life cycle: a node is loaded (NEW), runs the basic bound and the
 * first separation, enters the outer loop (LOOP) and leaves it when it is
 * pruned, given up or finished (active = 0); the host then frees the slot. */
enum { BQB_STAGE_FREE = 0, BQB_STAGE_NEW = 1, BQB_STAGE_LOOP = 2 };

/* Cut records (src/biqbin.h:81-107 hold the same information). */
typedef struct TriCut { int i, j, k, type; } TriCut;
typedef struct KgonCut { int type; int p[7]; } KgonCut;     /* pentagonal uses p[0..4] */

typedef struct BatchDims {
    int nodes;          /* node slots in use */
    int n0;             /* vertices of the original problem */
    int ld;             /* slab leading dimension */
    int max_bundle;
    int max_tri, max_pent, max_hept, max_m;
    int tri_list;       /* params.TriIneq (triangle list capacity) */
    int kg_list;        /* pentagonal/heptagonal list capacity */
    int max_contrib;    /* op_Bt gather capacity */
    int npairs;         /* ld*(ld-1)/2 */
    int sa_chains5, sa_chains7;   /* SA chains per pentagonal / heptagonal round */
    int tmp_ld;         /* per-node length of tmpm */
    int mix_slices;     /* IPM products: 0 FP64, 2..4 bf16 slices on the matrix cores */
    int certify;        /* prune on certified bounds (f_cert) instead of f */
    int work_slabs;     /* ld x ld slabs of IPM work per node */
    int learned;        /* new cuts start from the cut model's multipliers */
} BatchDims;

/* Device pointers of a batch (all arrays are per node at the stated stride). */
typedef struct BatchPtrs {
    NodeState *st;                  /* [nodes] */
    BqbParams *params;              /* [1] */
    const double *SP_L;             /* [n0*n0] original objective (process_input.c:259-344) */
    const int *xfixed, *solX;       /* [nodes*(n0-1)] */
    int *nvec;                      /* [nodes] n of each node (for the IPM kernel) */
    int *mask;                      /* [nodes] phase mask */
    int *mask2;                     /* [nodes] nodes re-solved with FP64 products */
    double *L, *L0, *X, *Xt, *Zh;   /* [nodes*ld*ld] each */
    double *ipm_work;               /* [nodes*3*ld*ld] */
    double *phi;                    /* [nodes] */
    double *ytest;                  /* [nodes*ld] dual vector y of the last IPM */
    double *ycen;                   /* [nodes*ld] y of the bundle center's evaluation (certification) */
    int *ipm_iters, *ipm_trials, *ipm_status;   /* [nodes] */
    double *Xb;                     /* [nodes*max_bundle*ld*ld] bundle matrices */
    double *G;                      /* [nodes*max_bundle*max_m], column i at i*m (ld = m) */
    double *F, *lambda, *zeta;      /* [nodes*max_bundle] */
    double *gamma, *dgamma, *gtest, *eta, *g;   /* [nodes*max_m] */
    double *tmpm;                   /* [nodes*tmp_ld] bundle QP partial sums */
    TriCut *tri;                    /* [nodes*max_tri] */
    KgonCut *pent, *hept;           /* [nodes*max_pent], [nodes*max_hept] */
    double *tri_y, *pent_y, *hept_y;/* dual multipliers stored with the cuts */
    TriCut *tri_lst;                /* [nodes*tri_list] */
    double *tri_lval;               /* [nodes*tri_list] */
    KgonCut *kg_lst;                /* [nodes*kg_list] */
    double *kg_lval;                /* [nodes*kg_list] */
    int *csr_ptr;                   /* [nodes*(npairs+1)] op_Bt gather offsets per matrix pair */
    int *csr_cut;                   /* [nodes*max_contrib] (cut index << 1) | (sign < 0) */
    int *csr_fill;                  /* [nodes*npairs] scratch counters */
    double *fracsol;                /* [nodes*(n0-1)] */
    int *xbest;                     /* [nodes*(n0-1)] best heuristic solution of the node */
    int *xheur;                     /* [nodes*(n0-1)] scratch solution of the heuristic */
    int *perm;                      /* [nodes*chains*ld] SA permutations */
    double *sa_val;                 /* [nodes*chains] */
    int *sa_perm;                   /* [nodes*chains*7] */
    int *sa_status;                 /* [nodes*chains] */
    /* staging of newly loaded nodes (compact, [nload]) and the slot map */
    NodeState *ld_st;               /* [nodes] */
    int *ld_xf, *ld_sx;             /* [nodes*(n0-1)] */
    int *ld_of_slot;                /* [nodes]: index into the staging, or -1 */
} BatchPtrs;

/* bundle QP: threads per node, and bundle elements whose G'eta partial sums
   stay in registers (the others use tmpm, BQB_QP_THREADS doubles each) */
#define BQB_QP_THREADS 256
#define BQB_QP_REG_PARTIALS 16

#define BQB_SA_CHAINS 256           /* >= max(Pent_Trials*3, Hepta_Trials*4) supported */
#define BQB_TRI_LIST_MAX 4096       /* largest TriIneq (triangle list capacity) supported */

#ifdef __cplusplus
}
#endif
#endif
