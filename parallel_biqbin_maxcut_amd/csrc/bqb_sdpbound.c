/*
 * bqb_sdpbound.c -- host driver of the batched node bound (C).
 *
 * A context owns the device arenas of up to max_nodes node slots for one
 * instance (SP->L, params).  bqb_bound_batch() bounds a batch of B&B nodes
 * with the reference's SDPbound algorithm (src/bounding.c:24-342): every
 * per-node decision is taken on the device by the control kernel; the host
 * only sequences the phases, which are the same for all nodes of the batch
 * (nodes that have finished simply mask themselves out), and polls how many
 * nodes are still active once per outer iteration.
 *
 * Each node's bound is a function of (node, LB at batch start, diff, seed):
 * the incumbent a node reads is the batch-start LB raised by its own
 * heuristic solutions (with one node per batch this is exactly the
 * reference's Bab_LBGet()).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "bqb_batch.h"
#include "bqb_kernels.h"
#include "bqb_phase.h"
#include "bqb_runtime.h"

hipError_t bqb_launch_control(const BatchDims *d, const BatchPtrs *p, int phase, int arg, hipStream_t s);
hipError_t bqb_launch_csr_build(const BatchDims *d, const BatchPtrs *p, hipStream_t s);
hipError_t bqb_launch_opbt(const BatchDims *d, const BatchPtrs *p, int use_gtest, hipStream_t s);
hipError_t bqb_launch_trisep(const BatchDims *d, const BatchPtrs *p, int tri_ineq, hipStream_t s);
hipError_t bqb_launch_sa(const BatchDims *d, const BatchPtrs *p, int k, hipStream_t s);
hipError_t bqb_launch_heur(const BatchDims *d, const BatchPtrs *p, int tile, int spl_integer, hipStream_t s);
hipError_t bqb_launch_bundle_qp(const BatchDims *d, const BatchPtrs *p, int kbound, hipStream_t s);
hipError_t bqb_launch_bundle_step(const BatchDims *d, const BatchPtrs *p, hipStream_t s);
hipError_t bqb_launch_cert_node(const BatchDims *d, const BatchPtrs *p, int mode, hipStream_t s);
hipError_t bqb_launch_learn(const BatchDims *d, const BatchPtrs *p, const bqb_mlp *m, float drop, int *rec_count,
                            int rec_cap, float *rec_feat, int *rec_key, int *idx, unsigned char *keep,
                            int after_round, hipStream_t s);

/* A captured launch train: kind 0 the basic sequence of newly loaded nodes
 * (key: occupied slots), kind 1 an outer-round slice (key: occupied slots,
 * bundle iterations, largest bundle clamped to 17, the last size where the
 * QP variant choice still changes). */
#define BQB_GRAPH_CACHE 16
struct train_graph {
    int kind, hi, nb, kcls;
    int phases;
    unsigned long long stamp;  /* 0: empty */
    hipGraphExec_t exec;
};

struct bqb_ctx {
    BatchDims d;
    BatchPtrs p;
    BqbParams params;
    int max_nodes;
    int tile;           /* IPM register tile (P*100 + R) */
    int tile0;          /* the tile chosen at creation (FP64 / explicit precision) */
    int heur_tile;      /* heuristic tile (P*100 + R) */
    int spl_int;        /* SP_L integer-valued (|entries| <= 2^20) */
    void *arena;
    size_t arena_bytes;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    NodeState *h_st;
    double last_ms;
    /* IPM launch profile: an event pair around every IPM launch while
     * profiling is on, and a device counter of algorithmic IPM flops */
    int profile_ipm;
    hipEvent_t *pev;
    int pev_cap, pev_n;
    unsigned long long *d_flops;     /* [2]: FP64 and bf16 matrix-core flops */
    /* slot engine: occupied slots carry a caller tag (-1 = free); nodes
     * pushed since the last advance wait in the host staging arrays */
    long long *tag;            /* [max_nodes] */
    int *free_stack, nfree;    /* free slots, lowest on top */
    int *of_slot;              /* [max_nodes] staging index or -1 */
    NodeState *ld_st;          /* [max_nodes] staged states */
    int *ld_xf, *ld_sx;        /* [max_nodes*(n0-1)] staged rows */
    int nload;
    int loaded;                /* nodes loaded by bqb_ctx_load still need their basic bound */
    int hi;                    /* 1 + highest occupied slot */
    double *h_frac;            /* [max_nodes*(n0-1)] */
    int *h_xb;                 /* [max_nodes*(n0-1)] */
    /* finished nodes not yet popped (FIFO) */
    struct done_node *done;
    int ndone, done_head, done_cap;
    /* trajectory capture: per-slot log of completed outer rounds, and the
     * records of the nodes that finished while capture was on */
    int capture;
    int tr_rmax;               /* rounds per slot log */
    double *tr_f, *tr_t, *tr_bound;   /* [max_nodes*tr_rmax] */
    int *tr_cuts;              /* [max_nodes*tr_rmax*3] */
    int *tr_n;                 /* [max_nodes] rounds logged */
    bqb_trajectory *traj;      /* each record owns one block (its round_f) */
    int ntraj, traj_cap;
    /* learned cut model: prediction (model) or recording of new cuts */
    const bqb_mlp *model;
    float drop;
    int rec_cap;
    int *d_rec_count;          /* [1] */
    float *d_rec_feat;         /* [rec_cap][16] */
    int *d_rec_key;            /* [rec_cap][12] */
    int *d_lidx;               /* [max_nodes][max_m] */
    double *d_lprev;           /* [max_nodes][ld*ld] the L of each slot's last bundle IPM (warm IPM) */
    int *d_warm_ok;            /* [max_nodes] */
    int ipm_warm;
    unsigned char *d_lkeep;    /* [max_nodes][max_m] */
    /* host heuristic replacing the device GW (bqb_ctx_set_heuristic) */
    bqb_heuristic_fn heur_fn;
    void *heur_user;
    double *h_slab;            /* [ld*ld] one node's L slab */
    double *h_PL;              /* [n0*n0] its n x n subproblem matrix */
    int *h_mask, *h_xf, *h_sx, *h_x;
    const double *h_SPL;       /* SP_L as given at creation (host copy) */
    /* launch trains captured as HIP graphs, keyed by their shape */
    int graphs;
    struct train_graph gcache[BQB_GRAPH_CACHE];
    unsigned long long gstamp;
    int capturing;             /* phases go into a capture, not counted as eager */
    int cap_phases;
    long long cnt[BQB_NCOUNTERS];
    /* inherited cuts (bqb_ctx_set_warm): the cut sets attached to staged
     * nodes until they are uploaded, and the final cut set of the node
     * popped last */
    int warm;
    struct cut_pack **wstage;  /* [max_nodes] by slot, or NULL */
    struct cut_pack *pop_cuts;
};

/* A node's cut set: records in the device layout and the multipliers
 * [triangles | pentagonal | heptagonal]; one malloc'd block. */
struct cut_pack {
    int ntri, npent, nhept;
    TriCut *tri;
    KgonCut *pent, *hept;
    double *y;
};

struct done_node {
    long long tag;
    NodeState st;
    double *frac;    /* n0-1 */
    int *xb;         /* n0-1 */
    struct cut_pack *cuts;   /* final cut set (warm contexts, branched nodes) */
};

static struct cut_pack *pack_alloc(int ntri, int npent, int nhept)
{
    const size_t bytes = sizeof(struct cut_pack) + sizeof(TriCut) * (size_t)ntri
                         + sizeof(KgonCut) * (size_t)(npent + nhept) + sizeof(double) * (size_t)(ntri + npent + nhept);
    struct cut_pack *k = malloc(bytes);
    if (!k) { bqb_set_error("cut set: out of memory"); return NULL; }
    k->ntri = ntri;
    k->npent = npent;
    k->nhept = nhept;
    k->y = (double *)(k + 1);
    k->tri = (TriCut *)(k->y + ntri + npent + nhept);
    k->pent = (KgonCut *)(k->tri + ntri);
    k->hept = k->pent + npent;
    return k;
}

#define CK(call, what) do { if (bqb_check((call), (what))) goto fail; } while (0)

/* graph setting of new contexts (bqb_set_default_graphs) */
static int g_default_graphs = 0;

/* A blocking wait for the context's stream (counted). */
static int ctx_sync(bqb_ctx *c)
{
    c->cnt[5]++;
    return bqb_check(hipStreamSynchronize(c->stream), "sync");
}

static void graphs_clear(bqb_ctx *c)
{
    for (int i = 0; i < BQB_GRAPH_CACHE; ++i)
        if (c->gcache[i].stamp) {
            hipGraphExecDestroy(c->gcache[i].exec);
            c->gcache[i].stamp = 0;
        }
}

static void traj_free(bqb_ctx *c)
{
    for (int i = 0; i < c->ntraj; ++i) free((void *)c->traj[i].round_f);
    free(c->traj);
    c->traj = NULL;
    c->ntraj = c->traj_cap = 0;
}

static void free_host(bqb_ctx *c)
{
    traj_free(c);
    free(c->h_slab); free(c->h_PL); free(c->h_mask); free(c->h_xf); free(c->h_sx); free(c->h_x);
    free((void *)c->h_SPL);
    free(c->tr_f); free(c->tr_t); free(c->tr_bound); free(c->tr_cuts); free(c->tr_n);
    free(c->h_st); free(c->tag); free(c->free_stack); free(c->of_slot); free(c->ld_st);
    free(c->ld_xf); free(c->ld_sx); free(c->h_frac); free(c->h_xb);
    for (int i = 0; i < c->done_cap; ++i) { free(c->done[i].frac); free(c->done[i].xb); free(c->done[i].cuts); }
    free(c->done);
    if (c->wstage)
        for (int i = 0; i < c->max_nodes; ++i) free(c->wstage[i]);
    free(c->wstage);
    free(c->pop_cuts);
}

/* All slots free, nothing staged, no finished nodes pending. */
static void engine_reset(bqb_ctx *c)
{
    for (int i = 0; i < c->max_nodes; ++i) {
        c->tag[i] = -1;
        c->of_slot[i] = -1;
        c->free_stack[i] = c->max_nodes - 1 - i;
    }
    c->nfree = c->max_nodes;
    c->nload = 0;
    c->loaded = 0;
    c->hi = 0;
    for (int i = 0; i < c->done_cap; ++i) { free(c->done[i].cuts); c->done[i].cuts = NULL; }
    c->ndone = c->done_head = 0;
    if (c->wstage)
        for (int i = 0; i < c->max_nodes; ++i) { free(c->wstage[i]); c->wstage[i] = NULL; }
    if (c->tr_n) memset(c->tr_n, 0, sizeof(int) * c->max_nodes);
}

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

bqb_ctx *bqb_ctx_create(int n0, const double *SP_L, const BqbParams *params, int max_nodes,
                        int max_bundle, int max_tri, int max_pent, int max_hept)
{
    if (bqb_ensure_init()) return NULL;
    if (n0 < 2 || max_nodes < 1) { bqb_set_error("bad context size"); return NULL; }
    if (max_bundle <= 0) max_bundle = 64;
    if (max_bundle > 64) { bqb_set_error("max_bundle > 64 is not supported"); return NULL; }
    if (max_tri <= 0) max_tri = 8192;
    if (max_pent <= 0) max_pent = 2048;
    if (max_hept <= 0) max_hept = 2048;
    int tile = bqb_ipm_tile_for(n0);
    if (!tile) { bqb_set_error("n0 = %d exceeds the largest compiled tile", n0); return NULL; }
    /* the heuristic's 1-opt runs one hyperplane per wave: 16 waves (32 x 32
     * threads) where the 32 x R tile covers n0 */
    /* the heuristic takes 256-thread tiles up to n0 = 112 (so several other
     * kernels share its CU), 1024 threads above */
    const int heur_tile = (n0 <= 64) ? 1604 : (n0 <= 80) ? 1605 : (n0 <= 96) ? 1606 : (n0 <= 112) ? 1607
                        : (n0 <= 128) ? 3204 : (n0 <= 160) ? 3205 : (n0 <= 192) ? 3206 : (n0 <= 224) ? 3207 : 3208;
    const int tile_ld = (tile / 100) * (tile % 100), heur_ld = (heur_tile / 100) * (heur_tile % 100);
    bqb_ctx *c = calloc(1, sizeof(bqb_ctx));
    if (!c) { bqb_set_error("calloc"); return NULL; }
    c->params = *params;
    if (c->params.adjust_TriIneq) c->params.TriIneq = n0 * 10;   /* process_input.c:135-136 */
    if (c->params.TriIneq > BQB_TRI_LIST_MAX || c->params.TriIneq < 1) {
        bqb_set_error("TriIneq = %d outside 1..%d", c->params.TriIneq, BQB_TRI_LIST_MAX);
        free(c);
        return NULL;
    }
    if (c->params.Pent_Trials * 3 > BQB_SA_CHAINS || c->params.Hepta_Trials * 4 > BQB_SA_CHAINS) {
        bqb_set_error("Pent_Trials*3 and Hepta_Trials*4 must not exceed %d", BQB_SA_CHAINS);
        free(c);
        return NULL;
    }
    c->max_nodes = max_nodes;
    c->tile = c->tile0 = tile;
    c->heur_tile = heur_tile;
    c->spl_int = 1;
    for (size_t e = 0; e < (size_t)n0 * n0 && c->spl_int; ++e)
        if (SP_L[e] != rint(SP_L[e]) || fabs(SP_L[e]) > 1048576.0) c->spl_int = 0;
    BatchDims *d = &c->d;
    d->nodes = max_nodes;
    d->n0 = n0;
    d->ld = tile_ld > heur_ld ? tile_ld : heur_ld;   /* slabs fit both tiles */
    d->max_bundle = max_bundle;
    d->max_tri = max_tri;
    d->max_pent = max_pent;
    d->max_hept = max_hept;
    d->max_m = max_tri + max_pent + max_hept;
    d->tmp_ld = (max_bundle - BQB_QP_REG_PARTIALS) * BQB_QP_THREADS;
    if (d->tmp_ld < d->max_m) d->tmp_ld = d->max_m;
    d->tri_list = c->params.TriIneq;
    d->kg_list = BQB_SA_CHAINS;
    d->max_contrib = 3 * max_tri + 10 * max_pent + 21 * max_hept;
    d->npairs = d->ld * (d->ld - 1) / 2;
    d->sa_chains5 = 3 * c->params.Pent_Trials;
    d->sa_chains7 = 4 * c->params.Hepta_Trials;
    d->mix_slices = 0;
    d->certify = 0;
    d->work_slabs = BQB_IPM_WORK_SLABS(0);

    const size_t N = (size_t)max_nodes, slab = (size_t)d->ld * d->ld, B = (size_t)(n0 - 1);
    struct { void **ptr; size_t bytes; } items[] = {
        {(void **)&c->p.st, sizeof(NodeState) * N},
        {(void **)&c->p.params, sizeof(BqbParams)},
        {(void **)&c->p.SP_L, sizeof(double) * (size_t)n0 * n0},
        {(void **)&c->p.xfixed, sizeof(int) * N * B},
        {(void **)&c->p.solX, sizeof(int) * N * B},
        {(void **)&c->p.nvec, sizeof(int) * N},
        {(void **)&c->p.mask, sizeof(int) * N},
        {(void **)&c->p.mask2, sizeof(int) * N},
        {(void **)&c->p.L, sizeof(double) * N * slab},
        {(void **)&c->p.L0, sizeof(double) * N * slab},
        {(void **)&c->p.X, sizeof(double) * N * slab},
        {(void **)&c->p.Xt, sizeof(double) * N * slab},
        {(void **)&c->p.Zh, sizeof(double) * N * slab},
        /* room for the mixed-precision layout (bqb_ctx_set_precision) */
        {(void **)&c->p.ipm_work, sizeof(double) * N * BQB_IPM_WORK_SLABS(1) * slab},
        {(void **)&c->p.phi, sizeof(double) * N},
        {(void **)&c->p.ytest, sizeof(double) * N * d->ld},
        {(void **)&c->p.ycen, sizeof(double) * N * d->ld},
        {(void **)&c->p.ipm_iters, sizeof(int) * N},
        {(void **)&c->p.ipm_trials, sizeof(int) * N},
        {(void **)&c->p.ipm_status, sizeof(int) * N},
        {(void **)&c->p.Xb, sizeof(double) * N * max_bundle * slab},
        {(void **)&c->p.G, sizeof(double) * N * max_bundle * d->max_m},
        {(void **)&c->p.F, sizeof(double) * N * max_bundle},
        {(void **)&c->p.lambda, sizeof(double) * N * max_bundle},
        {(void **)&c->p.zeta, sizeof(double) * N * max_bundle},
        {(void **)&c->p.gamma, sizeof(double) * N * d->max_m},
        {(void **)&c->p.dgamma, sizeof(double) * N * d->max_m},
        {(void **)&c->p.gtest, sizeof(double) * N * d->max_m},
        {(void **)&c->p.eta, sizeof(double) * N * d->max_m},
        {(void **)&c->p.g, sizeof(double) * N * d->max_m},
        {(void **)&c->p.tmpm, sizeof(double) * N * d->tmp_ld},
        {(void **)&c->p.tri, sizeof(TriCut) * N * max_tri},
        {(void **)&c->p.pent, sizeof(KgonCut) * N * max_pent},
        {(void **)&c->p.hept, sizeof(KgonCut) * N * max_hept},
        {(void **)&c->p.tri_y, sizeof(double) * N * max_tri},
        {(void **)&c->p.pent_y, sizeof(double) * N * max_pent},
        {(void **)&c->p.hept_y, sizeof(double) * N * max_hept},
        {(void **)&c->p.tri_lst, sizeof(TriCut) * N * d->tri_list},
        {(void **)&c->p.tri_lval, sizeof(double) * N * d->tri_list},
        {(void **)&c->p.kg_lst, sizeof(KgonCut) * N * d->kg_list},
        {(void **)&c->p.kg_lval, sizeof(double) * N * d->kg_list},
        {(void **)&c->p.csr_ptr, sizeof(int) * N * (d->npairs + 1)},
        {(void **)&c->p.csr_cut, sizeof(int) * N * d->max_contrib},
        {(void **)&c->p.csr_fill, sizeof(int) * N * d->npairs},
        {(void **)&c->p.fracsol, sizeof(double) * N * B},
        {(void **)&c->p.xbest, sizeof(int) * N * B},
        {(void **)&c->p.xheur, sizeof(int) * N * B},
        {(void **)&c->p.perm, sizeof(int) * N * BQB_SA_CHAINS * d->ld},
        {(void **)&c->p.sa_val, sizeof(double) * N * BQB_SA_CHAINS},
        {(void **)&c->p.sa_perm, sizeof(int) * N * BQB_SA_CHAINS * 7},
        {(void **)&c->p.sa_status, sizeof(int) * N * BQB_SA_CHAINS},
        {(void **)&c->d_flops, 2 * sizeof(unsigned long long)},
        {(void **)&c->p.ld_st, sizeof(NodeState) * N},
        {(void **)&c->p.ld_xf, sizeof(int) * N * B},
        {(void **)&c->p.ld_sx, sizeof(int) * N * B},
        {(void **)&c->p.ld_of_slot, sizeof(int) * N},
    };
    const int nitems = (int)(sizeof items / sizeof items[0]);
    size_t total = 0;
    for (int i = 0; i < nitems; ++i) total += align_up(items[i].bytes);
    CK(hipMalloc(&c->arena, total), "hipMalloc arena");
    c->arena_bytes = total;
    CK(hipMemset(c->arena, 0, total), "hipMemset arena");
    {
        char *base = (char *)c->arena;
        for (int i = 0; i < nitems; ++i) {
            *items[i].ptr = base;
            base += align_up(items[i].bytes);
        }
    }
    CK(hipMemcpy((void *)c->p.SP_L, SP_L, sizeof(double) * (size_t)n0 * n0, hipMemcpyHostToDevice), "H2D SP_L");
    {
        double *h = malloc(sizeof(double) * (size_t)n0 * n0);
        if (!h) { bqb_set_error("calloc"); goto fail; }
        memcpy(h, SP_L, sizeof(double) * (size_t)n0 * n0);
        c->h_SPL = h;
    }
    CK(hipMemcpy(c->p.params, &c->params, sizeof(BqbParams), hipMemcpyHostToDevice), "H2D params");
    CK(hipStreamCreate(&c->stream), "hipStreamCreate");
    CK(hipEventCreate(&c->ev0), "hipEventCreate");
    CK(hipEventCreate(&c->ev1), "hipEventCreate");
    c->h_st = calloc(N, sizeof(NodeState));
    c->tag = malloc(sizeof(long long) * N);
    c->free_stack = malloc(sizeof(int) * N);
    c->of_slot = malloc(sizeof(int) * N);
    c->ld_st = calloc(N, sizeof(NodeState));
    c->ld_xf = malloc(sizeof(int) * N * B);
    c->ld_sx = malloc(sizeof(int) * N * B);
    c->h_frac = malloc(sizeof(double) * N * B);
    c->h_xb = malloc(sizeof(int) * N * B);
    if (!c->h_st || !c->tag || !c->free_stack || !c->of_slot || !c->ld_st || !c->ld_xf || !c->ld_sx ||
        !c->h_frac || !c->h_xb) {
        bqb_set_error("calloc");
        goto fail;
    }
    engine_reset(c);
    c->graphs = g_default_graphs;
    return c;
fail:
    if (c->arena) hipFree(c->arena);
    free_host(c);
    free(c);
    return NULL;
}

static void free_events(bqb_ctx *c)
{
    for (int i = 0; i < c->pev_cap; ++i) hipEventDestroy(c->pev[i]);
    free(c->pev);
    c->pev = NULL;
    c->pev_cap = c->pev_n = 0;
}

void bqb_ctx_destroy(bqb_ctx *c)
{
    if (!c) return;
    hipStreamSynchronize(c->stream);
    graphs_clear(c);
    if (c->d_rec_count) hipFree(c->d_rec_count);
    if (c->d_lidx) hipFree(c->d_lidx);
    if (c->d_lprev) hipFree(c->d_lprev);
    free_events(c);
    free_host(c);
    hipFree(c->arena);
    hipStreamDestroy(c->stream);
    hipEventDestroy(c->ev0);
    hipEventDestroy(c->ev1);
    free(c);
}

size_t bqb_ctx_bytes(const bqb_ctx *c) { return c->arena_bytes; }
double bqb_ctx_last_ms(const bqb_ctx *c) { return c->last_ms; }
int bqb_ctx_ld(const bqb_ctx *c) { return c->d.ld; }

int bqb_abi_sizes(int *sizes, int count)
{
    const int z[5] = {(int)sizeof(BiqBinParameters), (int)sizeof(bqb_solve_options), (int)sizeof(bqb_solve_stats),
                      (int)sizeof(bqb_trajectory), (int)sizeof(NodeState)};
    for (int i = 0; i < count && i < 5; ++i) sizes[i] = z[i];
    return 5;
}

static int ipm_launch(bqb_ctx *c, const double *L, double *X)
{
    IpmBatchArgs a;
    memset(&a, 0, sizeof a);
    a.n = c->p.nvec;
    a.L = L;
    a.X = X;
    a.work = c->p.ipm_work;
    a.ld = c->d.ld;
    a.phi = c->p.phi;
    a.iters = c->p.ipm_iters;
    a.trials = c->p.ipm_trials;
    a.status = c->p.ipm_status;
    a.active = c->p.mask;
    a.mix_slices = c->d.mix_slices;
    a.wslabs = BQB_IPM_WORK_SLABS(1);     /* the arena's stride per slot */
    if (c->d.certify || c->ipm_warm) {   /* the certificate and the warm start need y */
        a.y_out = c->p.ytest;
        a.ystride = c->d.ld;
    }
    if (c->ipm_warm) {             /* bundle solves (L0 -> Xt) start from the slot's last */
        a.lprev = c->d_lprev;
        a.warm_ok = c->d_warm_ok;
        a.warm_mode = (X == c->p.Xt) ? 1 : 0;
    }
    const int prof = c->profile_ipm && c->pev_n + 2 <= c->pev_cap;
    if (prof) {
        a.flops = c->d_flops;
        a.flops_mx = c->d_flops + 1;
        hipEventRecord(c->pev[c->pev_n], c->stream);
    }
    if (bqb_check(bqb_launch_ipm_batch(&a, c->d.nodes, c->tile, c->stream), "ipm launch")) return -1;
    /* failed solves again: mixed -> FP64 products, blocked -> register tile */
    const int stride = (int)(sizeof(NodeState) / sizeof(int));
    if (a.mix_slices &&
        bqb_check(bqb_launch_ipm_retry(&a, c->d.nodes, c->tile, c->p.mask2, &c->p.st[0].flags, stride,
                                       BQB_FLAG_FP64_RETRY, c->stream), "ipm retry launch"))
        return -1;
    if (!a.mix_slices && c->tile % 100 == 1 &&
        bqb_check(bqb_launch_ipm_retry(&a, c->d.nodes, bqb_ipm_reg_tile_for_ld(c->tile / 100), c->p.mask2,
                                       &c->p.st[0].flags, stride, BQB_FLAG_REG_RETRY, c->stream), "ipm retry launch"))
        return -1;
    if (prof) {
        hipEventRecord(c->pev[c->pev_n + 1], c->stream);
        c->pev_n += 2;
    }
    return 0;
}

/* The heuristic phase with a host heuristic (src/heuristic.c:11-17 calls
 * runHeuristic_unpacked or a host override): for every node of the phase
 * mask, the host function gets the original objective SP->L, the node's
 * subproblem matrix PP->L (as createSubproblem builds it), its fixings, and x
 * (fixed variables preset, src/bounding.c:186-194); the node's incumbent is
 * then raised as updateSolution does (src/bab_functions.c:114-134). */
static int host_heuristic(bqb_ctx *c)
{
    const int B = c->d.n0 - 1, ld = c->d.ld, hi = c->d.nodes;
    hipStream_t s = c->stream;
    if (bqb_check(hipMemcpyAsync(c->h_mask, c->p.mask, sizeof(int) * hi, hipMemcpyDeviceToHost, s), "D2H mask") ||
        bqb_check(hipMemcpyAsync(c->h_st, c->p.st, sizeof(NodeState) * hi, hipMemcpyDeviceToHost, s), "D2H state") ||
        bqb_check(hipMemcpyAsync(c->h_xf, c->p.xfixed, sizeof(int) * B * hi, hipMemcpyDeviceToHost, s), "D2H xfixed") ||
        bqb_check(hipMemcpyAsync(c->h_sx, c->p.solX, sizeof(int) * B * hi, hipMemcpyDeviceToHost, s), "D2H solX") ||
        ctx_sync(c))
        return -1;
    int changed = 0;
    for (int i = 0; i < hi; ++i) {
        if (!c->h_mask[i]) continue;
        NodeState *st = &c->h_st[i];
        const int n = st->n;
        if (bqb_check(hipMemcpy(c->h_slab, c->p.L + (size_t)i * ld * ld, sizeof(double) * ld * ld,
                                hipMemcpyDeviceToHost), "D2H L"))
            return -1;
        for (int j = 0; j < n; ++j) memcpy(c->h_PL + (size_t)j * n, c->h_slab + (size_t)j * ld, sizeof(double) * n);
        const int *xf = c->h_xf + (size_t)i * B, *sx = c->h_sx + (size_t)i * B;
        for (int v = 0; v < B; ++v) c->h_x[v] = xf[v] ? sx[v] : 0;
        c->heur_fn((double *)c->h_SPL, c->d.n0, c->h_PL, n, (int *)xf, (int *)sx, c->h_x, c->heur_user);
        const double val = bqb_evaluate_solution(c->d.n0, c->h_SPL, c->h_x);
        if (val > st->LB) {
            st->LB = val;
            st->improved = 1;
            changed = 1;
            if (bqb_check(hipMemcpy(c->p.xbest + (size_t)i * B, c->h_x, sizeof(int) * B, hipMemcpyHostToDevice),
                          "H2D xbest"))
                return -1;
        }
    }
    if (changed && bqb_check(hipMemcpy(c->p.st, c->h_st, sizeof(NodeState) * hi, hipMemcpyHostToDevice), "H2D state"))
        return -1;
    return 0;
}

int bqb_ctx_set_heuristic(bqb_ctx *c, bqb_heuristic_fn fn, void *user)
{
    if (bqb_ctx_busy(c) > 0) { bqb_set_error("set_heuristic: nodes are in flight"); return -1; }
    const size_t B = (size_t)(c->d.n0 - 1), N = (size_t)c->max_nodes, ld = (size_t)c->d.ld;
    if (fn && !c->h_slab) {
        c->h_slab = malloc(sizeof(double) * ld * ld);
        c->h_PL = malloc(sizeof(double) * (size_t)c->d.n0 * c->d.n0);
        c->h_mask = malloc(sizeof(int) * N);
        c->h_xf = malloc(sizeof(int) * N * B);
        c->h_sx = malloc(sizeof(int) * N * B);
        c->h_x = malloc(sizeof(int) * (B + 1));
        if (!c->h_slab || !c->h_PL || !c->h_mask || !c->h_xf || !c->h_sx || !c->h_x) {
            bqb_set_error("calloc");
            return -1;
        }
    }
    c->heur_fn = fn;
    c->heur_user = user;
    return 0;
}

/* Run one phase kernel on the current batch (also used by tests). */
static int upload_staged(bqb_ctx *c);
static int basic_train(bqb_ctx *c);

int bqb_ctx_run(bqb_ctx *c, int what, int arg)
{
    const BatchDims *d = &c->d;
    const BatchPtrs *p = &c->p;
    hipStream_t s = c->stream;
    if (c->capturing) c->cap_phases++;
    else c->cnt[2]++;
    switch (what) {
    case 100: return ipm_launch(c, p->L, p->X);
    case 101: return ipm_launch(c, p->L0, p->Xt);
    case 102:
        if (c->heur_fn) return host_heuristic(c);
        return bqb_check(bqb_launch_heur(d, p, c->heur_tile, c->spl_int, s), "heur launch");
    case 103: return bqb_check(bqb_launch_trisep(d, p, c->params.TriIneq, s), "trisep launch");
    case 104: return bqb_check(bqb_launch_sa(d, p, 5, s), "sa5 launch");
    case 105: return bqb_check(bqb_launch_sa(d, p, 7, s), "sa7 launch");
    case 106: return bqb_check(bqb_launch_csr_build(d, p, s), "csr launch");
    case 107: return bqb_check(bqb_launch_opbt(d, p, arg, s), "opbt launch");
    case 108: return bqb_check(bqb_launch_bundle_qp(d, p, arg, s), "qp launch");
    case 109: return bqb_check(bqb_launch_bundle_step(d, p, s), "step launch");
    case 110: return bqb_check(bqb_launch_cert_node(d, p, 0, s), "certificate launch");
    case 111: return bqb_check(bqb_launch_cert_node(d, p, 1, s), "certificate launch");
    case 114: return bqb_check(bqb_launch_cert_node(d, p, 2, s), "certificate launch");
    case 115: {
        /* the nodes pushed since the last advance: upload and run their
         * basic sequence only (stepwise inspection of a node's start) */
        if (c->nload && upload_staged(c)) return -1;
        if (c->hi == 0) return 0;
        c->d.nodes = c->hi;
        c->loaded = 0;
        return basic_train(c);
    }
    case 112: case 113:
        if (!c->model && !c->rec_cap) return 0;
        return bqb_check(bqb_launch_learn(d, p, c->model, c->drop, c->rec_cap ? c->d_rec_count : NULL, c->rec_cap,
                                          c->d_rec_feat, c->d_rec_key, c->d_lidx, c->d_lkeep, what == 113, s),
                         "learn launch");
    default:
        if (what >= 0 && what <= PH_CERT_MASK)
            return bqb_check(bqb_launch_control(d, p, what, arg, s), "control launch");
        bqb_set_error("unknown phase %d", what);
        return -1;
    }
}

#define RUN(what, arg) do { if (bqb_ctx_run(c, (what), (arg))) return -1; } while (0)

/* ---- slot engine ---------------------------------------------------------
 * Nodes are pushed into free slots, advanced one outer round at a time and
 * popped when they have finished.  bqb_ctx_advance() runs, in order:
 *   - for the nodes pushed since the last call: load, createSubproblem,
 *     basic IPM, heuristic, prune / give-up test, first triangle
 *     separation, bundle start (bounding.c:24-163);
 *   - one outer round (bounding.c:170-326) for every node in the loop,
 *     including the ones just started: each node runs its own number of
 *     bundle iterations;
 * and then moves the nodes that have left SDPbound to the finished list.
 * A node's bound depends only on its own inputs, so it does not matter which
 * slot it occupies or which nodes share the launches with it. */

int bqb_ctx_push(bqb_ctx *c, long long tag, const int *xfixed, const int *solX, double LB,
                 double diff, int is_root, uint64_t seed)
{
    if (tag < 0) { bqb_set_error("push: tag must be >= 0"); return -1; }
    if (c->nfree == 0) { bqb_set_error("push: all %d slots are busy", c->max_nodes); return -1; }
    const int slot = c->free_stack[--c->nfree];
    const int k = c->nload++;
    const size_t B = (size_t)(c->d.n0 - 1);
    NodeState *st = &c->ld_st[k];
    memset(st, 0, sizeof *st);
    st->LB = LB;
    st->diff = diff;
    st->is_root = is_root;
    st->seed = seed;
    st->stage = BQB_STAGE_NEW;
    memcpy(c->ld_xf + k * B, xfixed, sizeof(int) * B);
    memcpy(c->ld_sx + k * B, solX, sizeof(int) * B);
    c->of_slot[slot] = k;
    c->tag[slot] = tag;
    if (c->tr_n) c->tr_n[slot] = 0;
    if (slot + 1 > c->hi) c->hi = slot + 1;
    return slot;
}

int bqb_ctx_busy(const bqb_ctx *c) { return c->max_nodes - c->nfree; }
int bqb_ctx_capacity(const bqb_ctx *c) { return c->max_nodes; }

/* Copy the staged nodes into their slots (PH_LOAD) and set d.nodes. */
static int upload_staged(bqb_ctx *c)
{
    const size_t B = (size_t)(c->d.n0 - 1);
    hipStream_t s = c->stream;
    c->d.nodes = c->hi;
    if (bqb_check(hipMemcpyAsync(c->p.ld_of_slot, c->of_slot, sizeof(int) * c->hi, hipMemcpyHostToDevice, s), "H2D slots")) return -1;
    if (bqb_check(hipMemcpyAsync(c->p.ld_st, c->ld_st, sizeof(NodeState) * c->nload, hipMemcpyHostToDevice, s), "H2D staged state")) return -1;
    if (bqb_check(hipMemcpyAsync(c->p.ld_xf, c->ld_xf, sizeof(int) * B * c->nload, hipMemcpyHostToDevice, s), "H2D staged xfixed")) return -1;
    if (bqb_check(hipMemcpyAsync(c->p.ld_sx, c->ld_sx, sizeof(int) * B * c->nload, hipMemcpyHostToDevice, s), "H2D staged solX")) return -1;
    /* inherited cut sets go straight into their slots' cut arrays */
    if (c->wstage)
        for (int i = 0; i < c->hi; ++i) {
            const struct cut_pack *k = c->wstage[i];
            if (!k) continue;
            const size_t ti = (size_t)i * c->d.max_tri, pi = (size_t)i * c->d.max_pent, hi = (size_t)i * c->d.max_hept;
            if ((k->ntri && (bqb_check(hipMemcpyAsync(c->p.tri + ti, k->tri, sizeof(TriCut) * k->ntri, hipMemcpyHostToDevice, s), "H2D cuts") ||
                             bqb_check(hipMemcpyAsync(c->p.tri_y + ti, k->y, sizeof(double) * k->ntri, hipMemcpyHostToDevice, s), "H2D cuts"))) ||
                (k->npent && (bqb_check(hipMemcpyAsync(c->p.pent + pi, k->pent, sizeof(KgonCut) * k->npent, hipMemcpyHostToDevice, s), "H2D cuts") ||
                              bqb_check(hipMemcpyAsync(c->p.pent_y + pi, k->y + k->ntri, sizeof(double) * k->npent, hipMemcpyHostToDevice, s), "H2D cuts"))) ||
                (k->nhept && (bqb_check(hipMemcpyAsync(c->p.hept + hi, k->hept, sizeof(KgonCut) * k->nhept, hipMemcpyHostToDevice, s), "H2D cuts") ||
                              bqb_check(hipMemcpyAsync(c->p.hept_y + hi, k->y + k->ntri + k->npent, sizeof(double) * k->nhept, hipMemcpyHostToDevice, s), "H2D cuts"))))
                return -1;
        }
    if (bqb_ctx_run(c, PH_LOAD, 0)) return -1;
    for (int i = 0; i < c->hi; ++i) c->of_slot[i] = -1;
    c->nload = 0;
    /* the host copies of the staging must stay valid until the copies ran */
    if (ctx_sync(c)) return -1;
    if (c->wstage)
        for (int i = 0; i < c->hi; ++i) { free(c->wstage[i]); c->wstage[i] = NULL; }
    return 0;
}

int bqb_ctx_set_warm(bqb_ctx *c, int on)
{
    if (bqb_ctx_busy(c) || c->nload) { bqb_set_error("set_warm: nodes in flight"); return -1; }
    if (on && !c->wstage) {
        c->wstage = calloc(c->max_nodes, sizeof(struct cut_pack *));
        if (!c->wstage) { bqb_set_error("calloc"); return -1; }
    }
    graphs_clear(c);           /* the basic train changes */
    c->warm = on ? 1 : 0;
    return 0;
}

int bqb_ctx_set_ipm_warm(bqb_ctx *c, int on)
{
    if (bqb_ctx_busy(c) || c->nload) { bqb_set_error("set_ipm_warm: nodes in flight"); return -1; }
    if (on && !c->d_lprev) {
        const size_t slab = (size_t)c->d.ld * c->d.ld, N = (size_t)c->max_nodes;
        if (bqb_check(hipMalloc((void **)&c->d_lprev, sizeof(double) * N * slab + sizeof(int) * N),
                      "hipMalloc warm IPM"))
            return -1;
        c->d_warm_ok = (int *)(c->d_lprev + N * slab);
    }
    /* no slot holds a usable last solve yet */
    if (on && bqb_check(hipMemsetAsync(c->d_warm_ok, 0, sizeof(int) * c->max_nodes, c->stream), "hipMemset warm"))
        return -1;
    graphs_clear(c);           /* the IPM launches change */
    c->ipm_warm = on ? 1 : 0;
    return 0;
}

int bqb_ctx_push_cuts(bqb_ctx *c, int slot, int ntri, const int *tri, int npent, const int *pent, int nhept,
                      const int *hept, const double *gamma)
{
    if (!c->warm) { bqb_set_error("push_cuts: inherited cuts are off (bqb_ctx_set_warm)"); return -1; }
    if (slot < 0 || slot >= c->max_nodes || c->of_slot[slot] < 0) {
        bqb_set_error("push_cuts: slot %d holds no node pushed since the last advance", slot);
        return -1;
    }
    if (ntri < 0 || npent < 0 || nhept < 0 || ntri > c->d.max_tri || npent > c->d.max_pent || nhept > c->d.max_hept) {
        bqb_set_error("push_cuts: %d/%d/%d cuts exceed the capacities %d/%d/%d", ntri, npent, nhept, c->d.max_tri,
                      c->d.max_pent, c->d.max_hept);
        return -1;
    }
    NodeState *st = &c->ld_st[c->of_slot[slot]];
    const int B = c->d.n0 - 1;
    const int *xf = c->ld_xf + (size_t)c->of_slot[slot] * B;
    int n = 1;
    for (int i = 0; i < B; ++i) n += !xf[i];
    /* records must be valid cuts of the node's n x n subproblem: the
     * separations' vertex order (triangles i > j > k as
     * cutting_planes.c:58-62, k-gonal vertices distinct) and type ranges */
    for (int q = 0; q < ntri; ++q) {
        const int *r = tri + 4 * q;
        if (!(r[0] < n && r[0] > r[1] && r[1] > r[2] && r[2] >= 0 && r[3] >= 1 && r[3] <= 4)) {
            bqb_set_error("push_cuts: triangle cut %d (%d %d %d type %d) is not a cut of an %d-vertex subproblem",
                          q, r[0], r[1], r[2], r[3], n);
            return -1;
        }
    }
    for (int q = 0; q < npent + nhept; ++q) {
        const int k = q < npent ? 5 : 7, *r = q < npent ? pent + 8 * q : hept + 8 * (q - npent);
        int ok = r[0] >= 1 && r[0] <= (k == 5 ? 3 : 4);
        for (int a = 0; a < k && ok; ++a) {
            ok = r[1 + a] >= 0 && r[1 + a] < n;
            for (int b = 0; b < a && ok; ++b) ok = r[1 + a] != r[1 + b];
        }
        if (!ok) { bqb_set_error("push_cuts: %d-gonal cut %d is not a cut of an %d-vertex subproblem", k, q, n); return -1; }
    }
    for (int q = 0; q < ntri + npent + nhept; ++q)
        if (!(gamma[q] >= 0.0) || !isfinite(gamma[q])) {
            bqb_set_error("push_cuts: multiplier %d is not a finite value >= 0", q);
            return -1;
        }
    struct cut_pack *k = pack_alloc(ntri, npent, nhept);
    if (!k) return -1;
    memcpy(k->tri, tri, sizeof(TriCut) * ntri);
    memcpy(k->pent, pent, sizeof(KgonCut) * npent);
    memcpy(k->hept, hept, sizeof(KgonCut) * nhept);
    memcpy(k->y, gamma, sizeof(double) * (ntri + npent + nhept));
    free(c->wstage[slot]);
    c->wstage[slot] = k;
    st->warm = ntri + npent + nhept > 0;
    st->NIneq = ntri;
    st->NPentIneq = npent;
    st->NHeptaIneq = nhept;
    return 0;
}

int bqb_ctx_last_cuts(const bqb_ctx *c, int *ntri, const int **tri, int *npent, const int **pent, int *nhept,
                      const int **hept, const double **gamma)
{
    const struct cut_pack *k = c->pop_cuts;
    if (ntri) *ntri = k ? k->ntri : 0;
    if (npent) *npent = k ? k->npent : 0;
    if (nhept) *nhept = k ? k->nhept : 0;
    if (tri) *tri = k ? (const int *)k->tri : NULL;
    if (pent) *pent = k ? (const int *)k->pent : NULL;
    if (hept) *hept = k ? (const int *)k->hept : NULL;
    if (gamma) *gamma = k ? k->y : NULL;
    return 0;
}

static int fetch_states(bqb_ctx *c)
{
    if (bqb_check(hipMemcpyAsync(c->h_st, c->p.st, sizeof(NodeState) * c->hi, hipMemcpyDeviceToHost, c->stream), "D2H state"))
        return -1;
    return ctx_sync(c);
}

static int done_push(bqb_ctx *c, long long tag, const NodeState *st, const double *frac, const int *xb,
                     struct cut_pack *cuts)
{
    const int B = c->d.n0 - 1;
    if (c->ndone == c->done_cap) {
        const int cap = c->done_cap ? 2 * c->done_cap : c->max_nodes;
        struct done_node *d = malloc(sizeof(struct done_node) * cap);
        if (!d) { bqb_set_error("calloc"); return -1; }
        for (int i = 0; i < c->ndone; ++i) d[i] = c->done[(c->done_head + i) % c->done_cap];
        for (int i = c->ndone; i < cap; ++i) {
            d[i].frac = malloc(sizeof(double) * B);
            d[i].xb = malloc(sizeof(int) * B);
            d[i].cuts = NULL;
            if (!d[i].frac || !d[i].xb) {
                for (int j = c->ndone; j <= i; ++j) { free(d[j].frac); free(d[j].xb); }
                free(d);
                bqb_set_error("calloc");
                return -1;
            }
        }
        free(c->done);
        c->done = d;
        c->done_cap = cap;
        c->done_head = 0;
    }
    struct done_node *e = &c->done[(c->done_head + c->ndone) % c->done_cap];
    e->tag = tag;
    e->st = *st;
    memcpy(e->frac, frac, sizeof(double) * B);
    memcpy(e->xb, xb, sizeof(int) * B);
    e->cuts = cuts;
    ++c->ndone;
    return 0;
}

/* Move the nodes that have left SDPbound to the finished list, free slots. */
/* Append the rounds completed since the last call to each slot's log: a
 * round is complete once its round-end phases ran (rphase back to 0) or the
 * node left the loop in them. */
static void log_rounds(bqb_ctx *c)
{
    for (int i = 0; i < c->hi; ++i) {
        const NodeState *s = &c->h_st[i];
        if (c->tag[i] < 0 || c->tr_n[i] >= s->count || c->tr_n[i] >= c->tr_rmax) continue;
        if (s->rphase != 0 && s->active) continue;
        const int r = i * c->tr_rmax + c->tr_n[i]++;
        c->tr_f[r] = s->f;
        c->tr_t[r] = s->t;
        c->tr_bound[r] = s->bound;
        c->tr_cuts[3 * r] = s->NIneq;
        c->tr_cuts[3 * r + 1] = s->NPentIneq;
        c->tr_cuts[3 * r + 2] = s->NHeptaIneq;
    }
}

/* Start the record of a finished slot: one block for the round log, the
 * cut records and gamma; the cut records and gamma are copied from the
 * device on the context's stream (the caller synchronizes). */
static int traj_start(bqb_ctx *c, int i)
{
    const NodeState *s = &c->h_st[i];
    if (c->ntraj == c->traj_cap) {
        const int cap = c->traj_cap ? 2 * c->traj_cap : 64;
        bqb_trajectory *t = realloc(c->traj, sizeof(bqb_trajectory) * cap);
        if (!t) { bqb_set_error("calloc"); return -1; }
        c->traj = t;
        c->traj_cap = cap;
    }
    const int R = c->tr_n[i], m = s->NIneq + s->NPentIneq + s->NHeptaIneq;
    const size_t bytes = sizeof(double) * (3 * (size_t)R + m) + sizeof(int) * (3 * (size_t)R + 4 * (size_t)s->NIneq
                         + 8 * (size_t)(s->NPentIneq + s->NHeptaIneq)) + 8;
    char *blk = malloc(bytes);
    if (!blk) { bqb_set_error("calloc"); return -1; }
    bqb_trajectory *t = &c->traj[c->ntraj++];
    memset(t, 0, sizeof *t);
    t->tag = c->tag[i];
    t->n = s->n;
    t->rounds = R;
    t->ntri = s->NIneq;
    t->npent = s->NPentIneq;
    t->nhept = s->NHeptaIneq;
    t->f = s->f;
    t->t = s->t;
    t->bound = s->bound;
    t->fixedvalue = s->fixedvalue;
    t->status = s->status;
    double *rf = (double *)blk, *rt = rf + R, *rb = rt + R, *g = rb + R;
    int *rc = (int *)(g + m), *tri = rc + 3 * R, *pent = tri + 4 * s->NIneq, *hept = pent + 8 * s->NPentIneq;
    const size_t off = (size_t)i * c->tr_rmax;
    memcpy(rf, c->tr_f + off, sizeof(double) * R);
    memcpy(rt, c->tr_t + off, sizeof(double) * R);
    memcpy(rb, c->tr_bound + off, sizeof(double) * R);
    memcpy(rc, c->tr_cuts + 3 * off, sizeof(int) * 3 * R);
    t->round_f = rf; t->round_t = rt; t->round_bound = rb; t->round_cuts = rc;
    t->tri = tri; t->pent = pent; t->hept = hept; t->gamma = g;
    hipStream_t st = c->stream;
    if (s->NIneq && bqb_check(hipMemcpyAsync(tri, c->p.tri + (size_t)i * c->d.max_tri, sizeof(TriCut) * s->NIneq, hipMemcpyDeviceToHost, st), "D2H cuts")) return -1;
    if (s->NPentIneq && bqb_check(hipMemcpyAsync(pent, c->p.pent + (size_t)i * c->d.max_pent, sizeof(KgonCut) * s->NPentIneq, hipMemcpyDeviceToHost, st), "D2H cuts")) return -1;
    if (s->NHeptaIneq && bqb_check(hipMemcpyAsync(hept, c->p.hept + (size_t)i * c->d.max_hept, sizeof(KgonCut) * s->NHeptaIneq, hipMemcpyDeviceToHost, st), "D2H cuts")) return -1;
    if (m && bqb_check(hipMemcpyAsync(g, c->p.gamma + (size_t)i * c->d.max_m, sizeof(double) * m, hipMemcpyDeviceToHost, st), "D2H gamma")) return -1;
    return 0;
}

static int collect(bqb_ctx *c)
{
    const size_t B = (size_t)(c->d.n0 - 1);
    int any = 0;
    for (int i = 0; i < c->hi; ++i)
        if (c->tag[i] >= 0 && !c->h_st[i].active) { any = 1; break; }
    if (!any) return 0;
    if (bqb_check(hipMemcpyAsync(c->h_frac, c->p.fracsol, sizeof(double) * B * c->hi, hipMemcpyDeviceToHost, c->stream), "D2H fracsol")) return -1;
    if (bqb_check(hipMemcpyAsync(c->h_xb, c->p.xbest, sizeof(int) * B * c->hi, hipMemcpyDeviceToHost, c->stream), "D2H xbest")) return -1;
    if (c->capture)
        for (int i = 0; i < c->hi; ++i)
            if (c->tag[i] >= 0 && !c->h_st[i].active && traj_start(c, i)) return -1;
    /* warm contexts: the final cut set and multipliers of every node that
     * may be branched (not pruned, no error) */
    struct cut_pack **fin = NULL;
    if (c->warm) {
        fin = calloc(c->hi, sizeof *fin);
        if (!fin) { bqb_set_error("calloc"); return -1; }
        for (int i = 0; i < c->hi; ++i) {
            const NodeState *t = &c->h_st[i];
            const int m = t->NIneq + t->NPentIneq + t->NHeptaIneq;
            if (c->tag[i] < 0 || t->active || t->prune || t->status || m == 0) continue;
            struct cut_pack *k = fin[i] = pack_alloc(t->NIneq, t->NPentIneq, t->NHeptaIneq);
            hipStream_t st = c->stream;
            if (!k ||
                (k->ntri && bqb_check(hipMemcpyAsync(k->tri, c->p.tri + (size_t)i * c->d.max_tri, sizeof(TriCut) * k->ntri, hipMemcpyDeviceToHost, st), "D2H cuts")) ||
                (k->npent && bqb_check(hipMemcpyAsync(k->pent, c->p.pent + (size_t)i * c->d.max_pent, sizeof(KgonCut) * k->npent, hipMemcpyDeviceToHost, st), "D2H cuts")) ||
                (k->nhept && bqb_check(hipMemcpyAsync(k->hept, c->p.hept + (size_t)i * c->d.max_hept, sizeof(KgonCut) * k->nhept, hipMemcpyDeviceToHost, st), "D2H cuts")) ||
                bqb_check(hipMemcpyAsync(k->y, c->p.gamma + (size_t)i * c->d.max_m, sizeof(double) * m, hipMemcpyDeviceToHost, st), "D2H gamma")) {
                ctx_sync(c);
                for (int j = 0; j <= i; ++j) free(fin[j]);
                free(fin);
                return -1;
            }
        }
    }
    if (ctx_sync(c)) {
        if (fin) for (int i = 0; i < c->hi; ++i) free(fin[i]);
        free(fin);
        return -1;
    }
    for (int i = 0; i < c->hi; ++i) {
        if (c->tag[i] < 0 || c->h_st[i].active) continue;
        if (done_push(c, c->tag[i], &c->h_st[i], c->h_frac + i * B, c->h_xb + i * B, fin ? fin[i] : NULL)) {
            for (int j = i; j < c->hi; ++j) free(fin ? fin[j] : NULL);
            free(fin);
            return -1;
        }
        if (fin) fin[i] = NULL;
        c->tag[i] = -1;
    }
    free(fin);
    /* free slots, lowest index on top of the stack */
    c->nfree = 0;
    for (int i = c->max_nodes - 1; i >= 0; --i)
        if (c->tag[i] < 0) c->free_stack[c->nfree++] = i;
    c->hi = 0;
    for (int i = 0; i < c->max_nodes; ++i)
        if (c->tag[i] >= 0) c->hi = i + 1;
    return 0;
}

/* The basic sequence of newly loaded nodes (bounding.c:24-163). */
static int basic_train(bqb_ctx *c)
{
    RUN(PH_SETUP, 0);
    RUN(100, 0);             /* basic IPM: L -> X */
    RUN(PH_POST_BASIC, 0);
    if (c->d.certify) RUN(110, 0);   /* certified basic bound */
    RUN(102, 0);             /* heuristic */
    RUN(PH_CHECK_BASIC, 0);
    RUN(103, 0);             /* first triangle separation */
    RUN(112, 0);             /* learned warm start / ranking of the new cuts */
    RUN(PH_BUNDLE_INIT, 0);
    RUN(106, 0);             /* op_Bt gather lists */
    if (c->warm) {           /* inherited multipliers: f(gamma) as the first center */
        RUN(PH_WARM, 0);
        RUN(107, 0);         /* L0 = L - B'(gamma) */
        RUN(101, 0);         /* IPM: L0 -> X_test */
        RUN(PH_REBUILD, 1);  /* a kept center is certified at its first round end */
    }
    return 0;
}

/* One outer-round slice: nb bundle iterations, then the round end
 * (bounding.c:170-326).  kmax bounds the bundle sizes at the slice start. */
static int round_train(bqb_ctx *c, int nb, int kmax)
{
    RUN(PH_OUTER_BEGIN, 0);
    for (int b = 0; b < nb; ++b) {
        RUN(PH_BUNDLE_NEXT, b);
        /* QP -> gamma_test; each bundle iteration appends at most one
         * element, so no bundle exceeds kmax + b (the QP kernel sizes its
         * LDS from this bound) */
        RUN(108, kmax + b);
        RUN(107, 1);         /* L0 = L - B'(gamma_test) */
        RUN(101, 0);         /* IPM: L0 -> X_test */
        RUN(109, 0);         /* serious / null step */
    }
    /* the certificate of the center the round ends with, where it moved:
     * the value the round-end prune test reads (bounding.c:180, 197) */
    if (c->d.certify) {
        RUN(PH_CERT_MASK, 0);
        RUN(111, 0);
    }
    RUN(PH_OUTER_A, 0);
    RUN(102, 0);             /* heuristic */
    RUN(PH_OUTER_B, 0);
    RUN(103, 0);             /* triangles */
    RUN(PH_SEP_PENT, 0);
    RUN(104, 0);             /* pentagonal */
    RUN(PH_SEP_HEPT, 0);
    RUN(105, 0);             /* heptagonal */
    RUN(113, 0);             /* learned warm start / ranking of the new cuts */
    RUN(PH_OUTER_C, 0);
    RUN(106, 0);             /* new gather lists */
    RUN(107, 0);             /* L0 = L - B'(gamma) */
    RUN(101, 0);             /* IPM: L0 -> X_test */
    RUN(PH_REBUILD, 0);      /* a moved center is certified at the round end */
    return 0;
}

/* Runs a train through its captured graph (capturing it on first use):
 * kind 0 basic_train, kind 1 round_train(nb, kmax).  The kernels and their
 * arguments are exactly the eager launches', in the same order on the same
 * stream. */
static int train_graph(bqb_ctx *c, int kind, int nb, int kmax)
{
    /* the QP launch depends on kmax + b only through its class (<= 0,
     * <= 16, > 16): above 16 every kmax gives the same train */
    const int kcls = kind ? (kmax > 17 ? 17 : kmax) : 0;
    if (!kind) nb = 0;
    struct train_graph *g = NULL, *victim = &c->gcache[0];
    for (int i = 0; i < BQB_GRAPH_CACHE; ++i) {
        struct train_graph *e = &c->gcache[i];
        if (e->stamp && e->kind == kind && e->hi == c->hi && e->nb == nb && e->kcls == kcls) { g = e; break; }
        if (e->stamp < victim->stamp) victim = e;
    }
    if (!g) {
        if (victim->stamp) {
            hipGraphExecDestroy(victim->exec);
            victim->stamp = 0;
        }
        hipGraph_t graph = NULL;
        if (bqb_check(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture")) return -1;
        c->capturing = 1;
        c->cap_phases = 0;
        const int rc = kind ? round_train(c, nb, kmax) : basic_train(c);
        c->capturing = 0;
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc || bqb_check(e, "end capture")) {
            if (graph) hipGraphDestroy(graph);
            return -1;
        }
        const hipError_t ei = hipGraphInstantiate(&victim->exec, graph, NULL, NULL, 0);
        hipGraphDestroy(graph);
        if (bqb_check(ei, "graph instantiate")) return -1;
        victim->kind = kind;
        victim->hi = c->hi;
        victim->nb = nb;
        victim->kcls = kcls;
        victim->phases = c->cap_phases;
        c->cnt[6]++;
        g = victim;
    }
    g->stamp = ++c->gstamp;
    c->cnt[3]++;
    c->cnt[4] += g->phases;
    return bqb_check(hipGraphLaunch(g->exec, c->stream), "graph launch");
}

/* Graphs replace the eager trains unless a host step or an event-timed
 * launch sits inside them. */
static int use_graphs(const bqb_ctx *c) { return c->graphs && !c->heur_fn && !c->profile_ipm; }

int bqb_set_default_graphs(int on)
{
    const int prev = g_default_graphs;
    g_default_graphs = on ? 1 : 0;
    return prev;
}

int bqb_ctx_set_graphs(bqb_ctx *c, int on)
{
    if (!on) graphs_clear(c);
    c->graphs = on ? 1 : 0;
    return 0;
}

int bqb_ctx_counters(bqb_ctx *c, long long *out, int reset)
{
    if (out)
        for (int i = 0; i < BQB_NCOUNTERS; ++i) out[i] = c->cnt[i];
    if (reset) memset(c->cnt, 0, sizeof c->cnt);
    return 0;
}

int bqb_ctx_advance(bqb_ctx *c)
{
    const BqbParams *P = &c->params;
    c->cnt[0]++;
    if (c->nload) {
        if (upload_staged(c)) return -1;
        c->loaded = 1;
    }
    if (c->hi == 0) return 0;
    c->d.nodes = c->hi;
    if (c->loaded) {
        if (use_graphs(c) ? train_graph(c, 0, 0, 0) : basic_train(c)) return -1;
        c->loaded = 0;
    }
    if (fetch_states(c)) return -1;
    /* one slice of the outer loop: every node in the loop begins a round if
     * it is between rounds, runs up to init_bundle_iter of its bundle
     * iterations, and goes through the round end once its bundle iterations
     * are done (a round longer than the slice continues in the next call) */
    int in_loop = 0, bmax = 0, kmax = 0;
    const int slice = P->init_bundle_iter > 0 ? P->init_bundle_iter : 1;
    for (int i = 0; i < c->hi; ++i) {
        const NodeState *s = &c->h_st[i];
        if (c->tag[i] >= 0 && s->bundle > kmax) kmax = s->bundle;
        if (c->tag[i] < 0 || !s->active || s->stage != BQB_STAGE_LOOP) continue;
        in_loop = 1;
        int left = s->bdl_left;
        if (s->rphase == 0) {
            /* the bundle length of the round it begins (PH_OUTER_BEGIN) */
            left = P->init_bundle_iter + (s->count + 1) / 2;
            if (s->count == 0) left = P->init_bundle_iter;
            else if (left > P->max_bundle_iter) left = P->max_bundle_iter;
        } else if (s->rphase == 2) {
            left = 0;
        }
        if (left > bmax) bmax = left;
    }
    if (in_loop) {
        const int nb = bmax < slice ? bmax : slice;
        c->cnt[1]++;
        if (use_graphs(c) ? train_graph(c, 1, nb, kmax) : round_train(c, nb, kmax)) return -1;
        if (fetch_states(c)) return -1;
    }
    if (c->capture) log_rounds(c);
    return collect(c);
}

int bqb_ctx_capture(bqb_ctx *c, int on)
{
    if (on && !c->tr_n) {
        c->tr_rmax = c->params.max_outer_iter + c->params.extra_iter + 1;
        if (c->tr_rmax < 1) c->tr_rmax = 1;
        const size_t R = (size_t)c->max_nodes * c->tr_rmax;
        c->tr_f = malloc(sizeof(double) * R);
        c->tr_t = malloc(sizeof(double) * R);
        c->tr_bound = malloc(sizeof(double) * R);
        c->tr_cuts = malloc(sizeof(int) * 3 * R);
        c->tr_n = calloc(c->max_nodes, sizeof(int));
        if (!c->tr_f || !c->tr_t || !c->tr_bound || !c->tr_cuts || !c->tr_n) {
            free(c->tr_f); free(c->tr_t); free(c->tr_bound); free(c->tr_cuts); free(c->tr_n);
            c->tr_f = c->tr_t = c->tr_bound = NULL; c->tr_cuts = c->tr_n = NULL;
            bqb_set_error("calloc");
            return -1;
        }
    }
    if (on && !c->capture && bqb_ctx_busy(c) > 0) {
        bqb_set_error("capture: start it while no node is in flight");
        return -1;
    }
    c->capture = on ? 1 : 0;
    return 0;
}

int bqb_ctx_trajectory_count(const bqb_ctx *c) { return c->ntraj; }

int bqb_ctx_trajectory(const bqb_ctx *c, int i, bqb_trajectory *out)
{
    if (i < 0 || i >= c->ntraj) { bqb_set_error("trajectory %d of %d", i, c->ntraj); return -1; }
    *out = c->traj[i];
    return 0;
}

void bqb_ctx_trajectory_clear(bqb_ctx *c) { traj_free(c); }

int bqb_ctx_pop(bqb_ctx *c, long long *tag, double *bound, double *fracsol, int *xbest, double *LB_out,
                int *improved, double *diff_out, int *status, int *stats)
{
    if (c->ndone == 0) return 0;
    const struct done_node *e = &c->done[c->done_head];
    const NodeState *s = &e->st;
    const size_t B = (size_t)(c->d.n0 - 1);
    if (tag) *tag = e->tag;
    if (bound) *bound = s->bound;
    if (fracsol) memcpy(fracsol, e->frac, sizeof(double) * B);
    if (xbest) memcpy(xbest, e->xb, sizeof(int) * B);
    if (LB_out) *LB_out = s->LB;
    if (improved) *improved = s->improved;
    if (diff_out) *diff_out = s->diff;
    if (status) *status = s->status;
    if (stats) {
        stats[0] = s->prune; stats[1] = s->giveup; stats[2] = s->count; stats[3] = s->ipm_calls;
        stats[4] = s->ipm_iters; stats[5] = s->max_cuts; stats[6] = s->max_bundle_seen; stats[7] = s->n;
        stats[8] = s->flags;
    }
    free(c->pop_cuts);
    c->pop_cuts = e->cuts;
    c->done[c->done_head].cuts = NULL;
    c->done_head = (c->done_head + 1) % c->done_cap;
    --c->ndone;
    return 1;
}

/* Stage a batch into slots 0..nodes-1 and load it on the device (the basic
 * bound has not run yet).  Discards anything in flight. */
int bqb_ctx_load(bqb_ctx *c, int nodes, const int *xfixed, const int *solX, const double *LB,
                 const double *diff, const int *is_root, const uint64_t *seeds)
{
    if (nodes < 1 || nodes > c->max_nodes) { bqb_set_error("batch of %d nodes (capacity %d)", nodes, c->max_nodes); return -1; }
    engine_reset(c);
    const size_t B = (size_t)(c->d.n0 - 1);
    for (int i = 0; i < nodes; ++i)
        if (bqb_ctx_push(c, i, xfixed + i * B, solX + i * B, LB ? LB[i] : 0.0, diff ? diff[i] : 0.0,
                         is_root ? is_root[i] : 0, seeds ? seeds[i] : (uint64_t)i) < 0)
            return -1;
    if (upload_staged(c)) return -1;
    c->loaded = 1;
    return 0;
}

/* Bound the loaded batch to the end (bounding.c:24-342 for every node). */
int bqb_ctx_bound(bqb_ctx *c)
{
    while (bqb_ctx_busy(c) > 0)
        if (bqb_ctx_advance(c)) return -1;
    return 0;
}

/* Results of slots 0..nodes-1 as left on the device by the last batch. */
int bqb_ctx_results(bqb_ctx *c, double *bound, double *fracsol, int *xbest, double *LB_out,
                    int *improved, double *diff_out, int *status, int *stats)
{
    const int nodes = c->d.nodes;
    const size_t B = (size_t)(c->d.n0 - 1);
    if (bqb_check(hipMemcpyAsync(c->h_st, c->p.st, sizeof(NodeState) * nodes, hipMemcpyDeviceToHost, c->stream), "D2H state")) return -1;
    if (fracsol && bqb_check(hipMemcpyAsync(fracsol, c->p.fracsol, sizeof(double) * nodes * B, hipMemcpyDeviceToHost, c->stream), "D2H fracsol")) return -1;
    if (xbest && bqb_check(hipMemcpyAsync(xbest, c->p.xbest, sizeof(int) * nodes * B, hipMemcpyDeviceToHost, c->stream), "D2H xbest")) return -1;
    if (ctx_sync(c)) return -1;
    for (int i = 0; i < nodes; ++i) {
        const NodeState *s = &c->h_st[i];
        if (bound) bound[i] = s->bound;
        if (LB_out) LB_out[i] = s->LB;
        if (improved) improved[i] = s->improved;
        if (diff_out) diff_out[i] = s->diff;
        if (status) status[i] = s->status;
        if (stats) {
            int *o = stats + BQB_NSTATS * i;
            o[0] = s->prune; o[1] = s->giveup; o[2] = s->count; o[3] = s->ipm_calls;
            o[4] = s->ipm_iters; o[5] = s->max_cuts; o[6] = s->max_bundle_seen; o[7] = s->n;
            o[8] = s->flags;
        }
    }
    return 0;
}

/* Bound `count` nodes (any number) through the context's slots, refilling
 * free slots after every outer round; outputs in input order. */
/* A stream of nodes shared by one or more contexts: each context takes the
 * next node index whenever it has a free slot, and writes each result at the
 * node's index. */
struct stream_job {
    pthread_mutex_t mu;
    int count, next, failed;
    const int *xfixed, *solX;
    const double *LB, *diff;
    const int *is_root;
    const uint64_t *seeds;
    double *bound, *fracsol;
    int *xbest;
    double *LB_out;
    int *improved;
    double *diff_out;
    int *status, *stats;
};

/* error of a context that stopped because another context failed */
#define STREAM_STOPPED "stopped: another context failed"

static int stream_take(struct stream_job *j)
{
    pthread_mutex_lock(&j->mu);
    const int i = (j->failed || j->next >= j->count) ? -1 : j->next++;
    pthread_mutex_unlock(&j->mu);
    return i;
}

static int stream_failed(struct stream_job *j)
{
    pthread_mutex_lock(&j->mu);
    const int f = j->failed;
    pthread_mutex_unlock(&j->mu);
    return f;
}

/* Run the context until the shared pool is empty and its slots drained. */
static int stream_run(bqb_ctx *c, struct stream_job *j)
{
    engine_reset(c);
    const size_t B = (size_t)(c->d.n0 - 1);
    double *frac = malloc(sizeof(double) * (B ? B : 1));
    int *xb = malloc(sizeof(int) * (B ? B : 1));
    if (!frac || !xb) { free(frac); free(xb); bqb_set_error("calloc"); return -1; }
    int rc = -1, pool = 1;
    for (;;) {
        while (pool && c->nfree > 0) {
            const int i = stream_take(j);
            if (i < 0) { pool = 0; break; }
            if (bqb_ctx_push(c, i, j->xfixed + i * B, j->solX + i * B, j->LB ? j->LB[i] : 0.0,
                             j->diff ? j->diff[i] : 0.0, j->is_root ? j->is_root[i] : 0,
                             j->seeds ? j->seeds[i] : (uint64_t)i) < 0)
                goto out;
        }
        if (bqb_ctx_busy(c) == 0) break;
        if (bqb_ctx_advance(c)) goto out;
        long long t;
        double bd, lbo, dfo;
        int imp, st, sts[BQB_NSTATS];
        while (bqb_ctx_pop(c, &t, &bd, frac, xb, &lbo, &imp, &dfo, &st, sts)) {
            if (j->bound) j->bound[t] = bd;
            if (j->fracsol) memcpy(j->fracsol + t * B, frac, sizeof(double) * B);
            if (j->xbest) memcpy(j->xbest + t * B, xb, sizeof(int) * B);
            if (j->LB_out) j->LB_out[t] = lbo;
            if (j->improved) j->improved[t] = imp;
            if (j->diff_out) j->diff_out[t] = dfo;
            if (j->status) j->status[t] = st;
            if (j->stats) memcpy(j->stats + BQB_NSTATS * t, sts, sizeof sts);
        }
        if (stream_failed(j)) { bqb_set_error(STREAM_STOPPED); goto out; }
    }
    rc = 0;
out:
    free(frac);
    free(xb);
    return rc;
}

struct stream_worker {
    struct stream_job *job;
    bqb_ctx *c;
    int rc;
    char err[512];
};

static void *stream_thread(void *arg)
{
    struct stream_worker *w = arg;
    w->rc = bqb_ensure_init();
    if (w->rc == 0) w->rc = stream_run(w->c, w->job);
    if (w->rc) {
        snprintf(w->err, sizeof w->err, "%s", bqb_last_error());
        pthread_mutex_lock(&w->job->mu);
        w->job->failed = 1;
        pthread_mutex_unlock(&w->job->mu);
    }
    return NULL;
}

int bqb_bound_stream_multi(bqb_ctx *const *ctx, int nctx, int count, const int *xfixed, const int *solX,
                           const double *LB, const double *diff, const int *is_root, const uint64_t *seeds,
                           double *bound, double *fracsol, int *xbest, double *LB_out, int *improved,
                           double *diff_out, int *status, int *stats)
{
    if (count < 0) { bqb_set_error("negative node count"); return -1; }
    if (nctx < 1 || nctx > BQB_MAX_GROUPS || !ctx) { bqb_set_error("1..%d contexts", BQB_MAX_GROUPS); return -1; }
    for (int i = 0; i < nctx; ++i) {
        if (!ctx[i] || ctx[i]->d.n0 != ctx[0]->d.n0) { bqb_set_error("contexts of different instances"); return -1; }
        for (int k = 0; k < i; ++k)
            if (ctx[k] == ctx[i]) { bqb_set_error("a context appears twice"); return -1; }
    }
    struct stream_job j = {.count = count, .xfixed = xfixed, .solX = solX, .LB = LB, .diff = diff,
                           .is_root = is_root, .seeds = seeds, .bound = bound, .fracsol = fracsol,
                           .xbest = xbest, .LB_out = LB_out, .improved = improved, .diff_out = diff_out,
                           .status = status, .stats = stats};
    pthread_mutex_init(&j.mu, NULL);
    struct stream_worker w[BQB_MAX_GROUPS];
    pthread_t th[BQB_MAX_GROUPS];
    int started = 0;
    for (int i = 1; i < nctx; ++i) {
        w[i] = (struct stream_worker){.job = &j, .c = ctx[i]};
        if (pthread_create(&th[i], NULL, stream_thread, &w[i])) {
            pthread_mutex_lock(&j.mu);
            j.failed = 1;
            pthread_mutex_unlock(&j.mu);
            break;
        }
        started = i;
    }
    hipEventRecord(ctx[0]->ev0, ctx[0]->stream);
    int rc = stream_run(ctx[0], &j);
    /* the first real error wins: a context that only stopped because
     * another one failed reports the sentinel, not the cause */
    char err0[512];
    snprintf(err0, sizeof err0, "%s", rc ? bqb_last_error() : "");
    int real = rc && strcmp(err0, STREAM_STOPPED) != 0;
    if (rc) {
        pthread_mutex_lock(&j.mu);
        j.failed = 1;
        pthread_mutex_unlock(&j.mu);
    }
    for (int i = 1; i <= started; ++i) {
        pthread_join(th[i], NULL);
        if (!w[i].rc) continue;
        if (!real && strcmp(w[i].err, STREAM_STOPPED) != 0) { snprintf(err0, sizeof err0, "%s", w[i].err); real = 1; }
        else if (!rc) snprintf(err0, sizeof err0, "%s", w[i].err);
        rc = -1;
    }
    if (rc) bqb_set_error("%s", err0);
    if (!rc && started < nctx - 1) { rc = -1; bqb_set_error("pthread_create failed"); }
    pthread_mutex_destroy(&j.mu);
    if (rc) return rc;
    hipEventRecord(ctx[0]->ev1, ctx[0]->stream);
    if (bqb_check(hipEventSynchronize(ctx[0]->ev1), "sync")) return -1;
    float ms = 0.f;
    hipEventElapsedTime(&ms, ctx[0]->ev0, ctx[0]->ev1);
    ctx[0]->last_ms = ms;
    return 0;
}

int bqb_bound_stream(bqb_ctx *c, int count, const int *xfixed, const int *solX, const double *LB,
                     const double *diff, const int *is_root, const uint64_t *seeds, double *bound,
                     double *fracsol, int *xbest, double *LB_out, int *improved, double *diff_out,
                     int *status, int *stats)
{
    return bqb_bound_stream_multi(&c, 1, count, xfixed, solX, LB, diff, is_root, seeds, bound, fracsol,
                                  xbest, LB_out, improved, diff_out, status, stats);
}

int bqb_bound_batch(bqb_ctx *c, int nodes, const int *xfixed, const int *solX, const double *LB,
                    const double *diff, const int *is_root, const uint64_t *seeds, double *bound,
                    double *fracsol, int *xbest, double *LB_out, int *improved, double *diff_out,
                    int *status, int *stats)
{
    if (nodes < 1 || nodes > c->max_nodes) { bqb_set_error("batch of %d nodes (capacity %d)", nodes, c->max_nodes); return -1; }
    return bqb_bound_stream(c, nodes, xfixed, solX, LB, diff, is_root, seeds, bound, fracsol, xbest,
                            LB_out, improved, diff_out, status, stats);
}

/* fct_eval (src/bundle.c:242-275) of one matrix on node slot 0:
 * X = argmax of the basic SDP for L - B'(gamma), f = phi + sum(gamma),
 * g = 1 - B(X).  Cut records: tri 4 ints (i, j, k, type), pent 6 ints
 * (type, 5 vertices), hept 8 ints (type, 7 vertices); vertices 0-based < n.
 * L and X are n x n column-major (X may be NULL), g has m entries (may be
 * NULL).  Returns 0, or -1 with bqb_last_error(); *status receives the IPM
 * status. */
int bqb_fct_eval(bqb_ctx *c, int n, const double *L, int ntri, const int *tri, int npent,
                 const int *pent, int nhept, const int *hept, const double *gamma, double *X,
                 double *g, double *f, int *status)
{
    if (bqb_ctx_busy(c) > 0) { bqb_set_error("fct_eval: the context has nodes in flight"); return -1; }
    const BatchDims d0 = c->d;
    const int ld = d0.ld, m = ntri + npent + nhept;
    if (n < 1 || n > ld) { bqb_set_error("fct_eval: n = %d outside 1..%d", n, ld); return -1; }
    if (ntri < 0 || npent < 0 || nhept < 0 || ntri > d0.max_tri || npent > d0.max_pent || nhept > d0.max_hept) {
        bqb_set_error("fct_eval: cut counts exceed the context capacity");
        return -1;
    }
    const int *lists[3] = {tri, pent, hept};
    const int counts[3] = {ntri, npent, nhept}, widths[3] = {3, 5, 7}, stride[3] = {4, 6, 8};
    for (int t = 0; t < 3; ++t)
        for (int q = 0; q < counts[t]; ++q) {
            const int *r = lists[t] + (size_t)q * stride[t];
            const int *v = (t == 0) ? r : r + 1;
            const int type = (t == 0) ? r[3] : r[0];
            if (type < 1 || type > (t == 1 ? 3 : 4)) { bqb_set_error("fct_eval: bad cut type"); return -1; }
            for (int a = 0; a < widths[t]; ++a) {
                if (v[a] < 0 || v[a] >= n) { bqb_set_error("fct_eval: cut vertex out of range"); return -1; }
                for (int b = 0; b < a; ++b)
                    if (v[a] == v[b]) { bqb_set_error("fct_eval: repeated cut vertex"); return -1; }
            }
        }
    const size_t slab = (size_t)ld * ld;
    double *hL = calloc(slab, sizeof(double));
    TriCut *ht = malloc(sizeof(TriCut) * (ntri ? ntri : 1));
    KgonCut *hp = calloc(npent ? npent : 1, sizeof(KgonCut));
    KgonCut *hh = calloc(nhept ? nhept : 1, sizeof(KgonCut));
    int rc = -1;
    if (!hL || !ht || !hp || !hh) { bqb_set_error("calloc"); goto out; }
    for (int j = 0; j < n; ++j) memcpy(hL + (size_t)j * ld, L + (size_t)j * n, sizeof(double) * n);
    for (int q = 0; q < ntri; ++q) {
        ht[q].i = tri[4 * q]; ht[q].j = tri[4 * q + 1]; ht[q].k = tri[4 * q + 2]; ht[q].type = tri[4 * q + 3];
    }
    for (int q = 0; q < npent; ++q) { hp[q].type = pent[6 * q]; memcpy(hp[q].p, pent + 6 * q + 1, 5 * sizeof(int)); }
    for (int q = 0; q < nhept; ++q) { hh[q].type = hept[8 * q]; memcpy(hh[q].p, hept + 8 * q + 1, 7 * sizeof(int)); }
    NodeState st;
    memset(&st, 0, sizeof st);
    st.n = n; st.active = 1; st.NIneq = ntri; st.NPentIneq = npent; st.NHeptaIneq = nhept;
    const int one = 1;
    c->d.nodes = 1;
    hipStream_t s = c->stream;
#define UP(dst, src, bytes, what) if (bqb_check(hipMemcpyAsync((void *)(dst), (src), (bytes), hipMemcpyHostToDevice, s), what)) goto out
    UP(c->p.L, hL, sizeof(double) * slab, "H2D L");
    UP(c->p.st, &st, sizeof st, "H2D state");
    UP(c->p.nvec, &n, sizeof(int), "H2D n");
    UP(c->p.mask, &one, sizeof(int), "H2D mask");
    if (ntri) UP(c->p.tri, ht, sizeof(TriCut) * ntri, "H2D cuts");
    if (npent) UP(c->p.pent, hp, sizeof(KgonCut) * npent, "H2D cuts");
    if (nhept) UP(c->p.hept, hh, sizeof(KgonCut) * nhept, "H2D cuts");
    if (m) UP(c->p.gamma, gamma, sizeof(double) * m, "H2D gamma");
#undef UP
    if (bqb_ctx_run(c, 106, 0) || bqb_ctx_run(c, 107, 0) || bqb_ctx_run(c, 101, 0) ||
        bqb_ctx_run(c, PH_FCT_FINISH, 0))
        goto out;
    if (bqb_check(hipMemcpyAsync(&st, c->p.st, sizeof st, hipMemcpyDeviceToHost, s), "D2H state")) goto out;
    if (X && bqb_check(hipMemcpyAsync(hL, c->p.Xt, sizeof(double) * slab, hipMemcpyDeviceToHost, s), "D2H X")) goto out;
    if (g && m && bqb_check(hipMemcpyAsync(g, c->p.g, sizeof(double) * m, hipMemcpyDeviceToHost, s), "D2H g")) goto out;
    if (ctx_sync(c)) goto out;
    if (X)
        for (int j = 0; j < n; ++j) memcpy(X + (size_t)j * n, hL + (size_t)j * ld, sizeof(double) * n);
    if (f) *f = st.f_test;
    if (status) *status = st.status;
    rc = 0;
out:
    c->d.nodes = d0.nodes;
    free(hL); free(ht); free(hp); free(hh);
    return rc;
}

/* ---- raw access to one node's device buffers (tests, diagnostics) ---- */
static int buffer_of(bqb_ctx *c, int id, int node, char **ptr, size_t *bytes)
{
    const BatchDims *d = &c->d;
    const BatchPtrs *p = &c->p;
    const size_t slab = (size_t)d->ld * d->ld, B = (size_t)(d->n0 - 1);
    switch (id) {
    case 0: *ptr = (char *)(p->st + node); *bytes = sizeof(NodeState); break;
    case 1: *ptr = (char *)(p->L + node * slab); *bytes = sizeof(double) * slab; break;
    case 2: *ptr = (char *)(p->L0 + node * slab); *bytes = sizeof(double) * slab; break;
    case 3: *ptr = (char *)(p->X + node * slab); *bytes = sizeof(double) * slab; break;
    case 4: *ptr = (char *)(p->Xt + node * slab); *bytes = sizeof(double) * slab; break;
    case 5: *ptr = (char *)(p->tri + (size_t)node * d->max_tri); *bytes = sizeof(TriCut) * d->max_tri; break;
    case 6: *ptr = (char *)(p->tri_y + (size_t)node * d->max_tri); *bytes = sizeof(double) * d->max_tri; break;
    case 7: *ptr = (char *)(p->pent + (size_t)node * d->max_pent); *bytes = sizeof(KgonCut) * d->max_pent; break;
    case 8: *ptr = (char *)(p->pent_y + (size_t)node * d->max_pent); *bytes = sizeof(double) * d->max_pent; break;
    case 9: *ptr = (char *)(p->hept + (size_t)node * d->max_hept); *bytes = sizeof(KgonCut) * d->max_hept; break;
    case 10: *ptr = (char *)(p->hept_y + (size_t)node * d->max_hept); *bytes = sizeof(double) * d->max_hept; break;
    case 11: *ptr = (char *)(p->gamma + (size_t)node * d->max_m); *bytes = sizeof(double) * d->max_m; break;
    case 12: *ptr = (char *)(p->gtest + (size_t)node * d->max_m); *bytes = sizeof(double) * d->max_m; break;
    case 13: *ptr = (char *)(p->g + (size_t)node * d->max_m); *bytes = sizeof(double) * d->max_m; break;
    case 14: *ptr = (char *)(p->G + (size_t)node * d->max_bundle * d->max_m); *bytes = sizeof(double) * d->max_bundle * d->max_m; break;
    case 15: *ptr = (char *)(p->F + (size_t)node * d->max_bundle); *bytes = sizeof(double) * d->max_bundle; break;
    case 16: *ptr = (char *)(p->lambda + (size_t)node * d->max_bundle); *bytes = sizeof(double) * d->max_bundle; break;
    case 17: *ptr = (char *)(p->tri_lst + (size_t)node * d->tri_list); *bytes = sizeof(TriCut) * d->tri_list; break;
    case 18: *ptr = (char *)(p->tri_lval + (size_t)node * d->tri_list); *bytes = sizeof(double) * d->tri_list; break;
    case 19: *ptr = (char *)(p->kg_lst + (size_t)node * d->kg_list); *bytes = sizeof(KgonCut) * d->kg_list; break;
    case 20: *ptr = (char *)(p->kg_lval + (size_t)node * d->kg_list); *bytes = sizeof(double) * d->kg_list; break;
    case 21: *ptr = (char *)(p->fracsol + (size_t)node * B); *bytes = sizeof(double) * B; break;
    case 22: *ptr = (char *)(p->xbest + (size_t)node * B); *bytes = sizeof(int) * B; break;
    case 23: *ptr = (char *)(p->xheur + (size_t)node * B); *bytes = sizeof(int) * B; break;
    case 24: *ptr = (char *)(p->mask + node); *bytes = sizeof(int); break;
    case 25: *ptr = (char *)(p->phi + node); *bytes = sizeof(double); break;
    case 26: *ptr = (char *)(p->Xb + (size_t)node * d->max_bundle * slab); *bytes = sizeof(double) * d->max_bundle * slab; break;
    case 27: *ptr = (char *)(p->nvec + node); *bytes = sizeof(int); break;
    case 28: *ptr = (char *)(p->dgamma + (size_t)node * d->max_m); *bytes = sizeof(double) * d->max_m; break;
    case 29: *ptr = (char *)(p->sa_val + (size_t)node * BQB_SA_CHAINS); *bytes = sizeof(double) * BQB_SA_CHAINS; break;
    case 30: *ptr = (char *)(p->ycen + (size_t)node * d->ld); *bytes = sizeof(double) * d->ld; break;
    default: bqb_set_error("unknown buffer %d", id); return -1;
    }
    return 0;
}

int bqb_ctx_download(bqb_ctx *c, int id, int node, void *host, size_t bytes)
{
    char *ptr; size_t cap;
    if (node < 0 || node >= c->max_nodes || buffer_of(c, id, node, &ptr, &cap)) return -1;
    if (bytes > cap) bytes = cap;
    if (ctx_sync(c)) return -1;
    return bqb_check(hipMemcpy(host, ptr, bytes, hipMemcpyDeviceToHost), "D2H buffer");
}

int bqb_ctx_upload(bqb_ctx *c, int id, int node, const void *host, size_t bytes)
{
    char *ptr; size_t cap;
    if (node < 0 || node >= c->max_nodes || buffer_of(c, id, node, &ptr, &cap)) return -1;
    if (bytes > cap) { bqb_set_error("upload larger than buffer"); return -1; }
    if (ctx_sync(c)) return -1;
    return bqb_check(hipMemcpy(ptr, host, bytes, hipMemcpyHostToDevice), "H2D buffer");
}

int bqb_ctx_sync(bqb_ctx *c) { return ctx_sync(c); }

/* Start (max_launches > 0) or stop (0) timing every IPM launch with an event
 * pair and counting its algorithmic flops on the device.  Starting resets
 * the counters. */
int bqb_ctx_profile_ipm(bqb_ctx *c, int max_launches)
{
    if (ctx_sync(c)) return -1;
    if (max_launches <= 0) { c->profile_ipm = 0; return 0; }
    if (2 * max_launches > c->pev_cap) {
        free_events(c);
        c->pev = calloc((size_t)2 * max_launches, sizeof(hipEvent_t));
        if (!c->pev) { bqb_set_error("calloc"); return -1; }
        for (int i = 0; i < 2 * max_launches; ++i) {
            if (bqb_check(hipEventCreate(&c->pev[i]), "hipEventCreate")) { c->pev_cap = i; return -1; }
        }
        c->pev_cap = 2 * max_launches;
    }
    c->pev_n = 0;
    c->profile_ipm = 1;
    return bqb_check(hipMemset(c->d_flops, 0, 2 * sizeof(unsigned long long)), "hipMemset flops");
}

/* IPM launches timed since profiling started: count, summed duration (ms)
 * and algorithmic flops. */
static int learn_scratch(bqb_ctx *c)
{
    if (c->d_lidx) return 0;
    const size_t cells = (size_t)c->max_nodes * c->d.max_m;
    if (bqb_check(hipMalloc((void **)&c->d_lidx, cells * (sizeof(int) + 1)), "hipMalloc learn scratch")) return -1;
    c->d_lkeep = (unsigned char *)(c->d_lidx + cells);
    return 0;
}

int bqb_ctx_set_model(bqb_ctx *c, const bqb_mlp *model, float drop)
{
    if (bqb_ctx_busy(c) > 0) { bqb_set_error("set_model: nodes are in flight"); return -1; }
    if (model && learn_scratch(c)) return -1;
    c->model = model;
    c->drop = model ? drop : 0.f;
    c->d.learned = (model && !c->rec_cap) ? 1 : 0;
    graphs_clear(c);
    return 0;
}

int bqb_ctx_learn_record(bqb_ctx *c, int capacity)
{
    if (bqb_ctx_busy(c) > 0) { bqb_set_error("learn_record: nodes are in flight"); return -1; }
    graphs_clear(c);
    if (c->d_rec_count) { hipFree(c->d_rec_count); c->d_rec_count = NULL; }
    c->rec_cap = 0;
    if (capacity <= 0) return 0;
    if (learn_scratch(c)) return -1;
    const size_t bytes = 256 + (size_t)capacity * (16 * sizeof(float) + 12 * sizeof(int));
    if (bqb_check(hipMalloc((void **)&c->d_rec_count, bytes), "hipMalloc records")) return -1;
    c->d_rec_feat = (float *)((char *)c->d_rec_count + 256);
    c->d_rec_key = (int *)(c->d_rec_feat + (size_t)capacity * 16);
    if (bqb_check(hipMemset(c->d_rec_count, 0, sizeof(int)), "hipMemset records")) return -1;
    c->rec_cap = capacity;
    return 0;
}

int bqb_ctx_learn_records(bqb_ctx *c, int *count, float *feat, int *key)
{
    int k = 0;
    if (!c->rec_cap) { *count = 0; return 0; }
    if (ctx_sync(c)) return -1;
    if (bqb_check(hipMemcpy(&k, c->d_rec_count, sizeof(int), hipMemcpyDeviceToHost), "D2H count")) return -1;
    if (k > c->rec_cap) k = c->rec_cap;
    *count = k;
    if (k && feat && bqb_check(hipMemcpy(feat, c->d_rec_feat, sizeof(float) * 16 * (size_t)k, hipMemcpyDeviceToHost), "D2H records"))
        return -1;
    if (k && key && bqb_check(hipMemcpy(key, c->d_rec_key, sizeof(int) * 12 * (size_t)k, hipMemcpyDeviceToHost), "D2H records"))
        return -1;
    return 0;
}

int bqb_ctx_set_precision(bqb_ctx *c, int slices, int certify)
{
    if (slices < -1 || slices == 1 || slices > BQB_MIX_MAX_SLICES) {
        bqb_set_error("slices = %d: 0 (FP64 products), 2..%d, or -1 (automatic)", slices, BQB_MIX_MAX_SLICES);
        return -1;
    }
    if (bqb_ctx_busy(c) > 0) { bqb_set_error("set_precision: nodes are in flight"); return -1; }
    int tile = c->tile0;
    if (slices < 0) {
        /* the class's entry of the automatic table; its slabs are those of
         * the context's class, so only the kernel changes */
        int s = 0;
        const int t = bqb_ipm_auto_for(c->d.n0, &s);
        const int tld = (t % 100 == 1) ? t / 100 : (t / 100) * (t % 100);
        if (!t || tld > c->d.ld) { bqb_set_error("set_precision: no automatic tile for n0 = %d", c->d.n0); return -1; }
        tile = t;
        slices = s;
        certify = 1;
    }
    c->tile = tile;
    c->d.mix_slices = slices;
    c->d.certify = (certify || slices > 0) ? 1 : 0;
    c->d.work_slabs = BQB_IPM_WORK_SLABS(slices);
    graphs_clear(c);
    return 0;
}

int bqb_ctx_ipm_flops_mx(bqb_ctx *c, double *flops_mx)
{
    unsigned long long f[2] = {0, 0};
    if (ctx_sync(c)) return -1;
    if (bqb_check(hipMemcpy(f, c->d_flops, sizeof f, hipMemcpyDeviceToHost), "D2H flops")) return -1;
    *flops_mx = (double)f[1];
    return 0;
}

int bqb_ctx_ipm_stats(bqb_ctx *c, int *launches, double *ms_total, double *flops)
{
    if (ctx_sync(c)) return -1;
    double ms = 0.0;
    for (int i = 0; i + 1 < c->pev_n; i += 2) {
        float e = 0.f;
        if (bqb_check(hipEventElapsedTime(&e, c->pev[i], c->pev[i + 1]), "hipEventElapsedTime")) return -1;
        ms += e;
    }
    unsigned long long f = 0;
    if (bqb_check(hipMemcpy(&f, c->d_flops, sizeof f, hipMemcpyDeviceToHost), "D2H flops")) return -1;
    if (launches) *launches = c->pev_n / 2;
    if (ms_total) *ms_total = ms;
    if (flops) *flops = (double)f;
    return 0;
}
