/*
 * bqb_bab.c -- host side of the solver (C): instance and parameter input,
 * the objective matrix, and the branch-and-bound node queue that feeds the
 * batched device bound.
 *
 * Reference behaviour followed:
 *   src/process_input.c:153-180  readParameters ("name = value" lines over
 *                                the PARAM_FIELDS defaults)
 *   src/process_input.c:206-250  readData (rudy: "n m" then m lines "i j w",
 *                                1-based, integer weights)
 *   src/process_input.c:259-344  process_adj_matrix (SP->L)
 *   src/process_input.c:135-136  adjust_TriIneq
 *   src/bab_functions.c:25-69    initial incumbent 0, root node
 *   src/bab_functions.c:224-297  evaluate, branch on the chosen variable
 *   src/bab_functions.c:387-421  getBranchingVariable
 *   src/heap.c                   max-heap on upper_bound (ties: -1)
 *
 * The MPI master/worker loop becomes one queue on one GPU: a context with
 * `batch` node slots is kept full with the best open nodes (nodes whose
 * bound no longer beats the incumbent are dropped), every slot advances one
 * outer round of SDPbound per step, and a node that has finished raises the
 * incumbent with its heuristic solution and is branched at once.  Each node
 * reads the incumbent of the moment it entered its slot.  Children carry the
 * parent's bound until they are evaluated.  With batch = 1 this is the
 * sequential B&B of the reference with one worker.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "bqb_batch.h"
#include "bqb_rng.h"
#include "bqb_runtime.h"

/* ------------------------------------------------------------ parameters */
void bqb_params_default(BiqBinParameters *p)
{
    p->init_bundle_iter = 3;
    p->max_bundle_iter = 15;
    p->triag_iter = 5;
    p->pent_iter = 5;
    p->hept_iter = 5;
    p->max_outer_iter = 20;
    p->extra_iter = 10;
    p->violated_TriIneq = 1e-3;
    p->TriIneq = 5000;
    p->adjust_TriIneq = 1;
    p->PentIneq = 5000;
    p->HeptaIneq = 5000;
    p->Pent_Trials = 60;
    p->Hepta_Trials = 50;
    p->include_Pent = 1;
    p->include_Hepta = 1;
    p->root = 0;
    p->use_diff = 1;
    p->time_limit = 0;
    p->branchingStrategy = 1;
}

int bqb_read_params(const char *path, BiqBinParameters *p)
{
    bqb_params_default(p);
    FILE *f = fopen(path, "r");
    if (!f) { bqb_set_error("parameter file %s not found", path); return -1; }
    char line[256], name[64];
    while (fgets(line, sizeof line, f)) {
        if (sscanf(line, " %63[^= \t]", name) != 1) continue;
        const char *eq = strchr(line, '=');
        if (!eq) continue;
        const char *v = eq + 1;
#define INT_FIELD(F) if (!strcmp(name, #F)) { p->F = atoi(v); continue; }
        INT_FIELD(init_bundle_iter) INT_FIELD(max_bundle_iter) INT_FIELD(triag_iter)
        INT_FIELD(pent_iter) INT_FIELD(hept_iter) INT_FIELD(max_outer_iter) INT_FIELD(extra_iter)
        INT_FIELD(TriIneq) INT_FIELD(adjust_TriIneq) INT_FIELD(PentIneq) INT_FIELD(HeptaIneq)
        INT_FIELD(Pent_Trials) INT_FIELD(Hepta_Trials) INT_FIELD(include_Pent)
        INT_FIELD(include_Hepta) INT_FIELD(root) INT_FIELD(use_diff) INT_FIELD(time_limit)
        INT_FIELD(branchingStrategy)
#undef INT_FIELD
        if (!strcmp(name, "violated_TriIneq")) { p->violated_TriIneq = atof(v); continue; }
    }
    fclose(f);
    return 0;
}

/* ------------------------------------------------------------- instances */
double *bqb_read_rudy(const char *path, int *nv_out)
{
    FILE *f = fopen(path, "r");
    if (!f) { bqb_set_error("cannot open %s", path); return NULL; }
    int nv = 0, ne = 0;
    if (fscanf(f, "%d %d", &nv, &ne) != 2 || nv <= 0 || ne < 0) {
        fclose(f);
        bqb_set_error("%s: bad header (expected 'vertices edges')", path);
        return NULL;
    }
    double *adj = calloc((size_t)nv * nv, sizeof(double));
    if (!adj) { fclose(f); bqb_set_error("calloc"); return NULL; }
    for (int e = 0; e < ne; ++e) {
        int i, j;
        double w;
        if (fscanf(f, "%d %d %lf", &i, &j, &w) != 3) {
            bqb_set_error("%s: problem reading edge %d", path, e + 1);
            goto bad;
        }
        if (i < 1 || i > nv || j < 1 || j > nv) {
            bqb_set_error("%s: edge %d has a vertex out of range", path, e + 1);
            goto bad;
        }
        if ((double)(int)w != w) {
            bqb_set_error("%s: edge %d: all edge weights need to be integers", path, e + 1);
            goto bad;
        }
        adj[(size_t)nv * (j - 1) + (i - 1)] = w;
        adj[(size_t)nv * (i - 1) + (j - 1)] = w;
    }
    fclose(f);
    *nv_out = nv;
    return adj;
bad:
    fclose(f);
    free(adj);
    return NULL;
}

/* SP->L = [Lap, Lap e; (Lap e)', e' Lap e] over the first nv-1 vertices,
 * where Lap = Diag(Adj e) - Adj uses the full degree (the last vertex is the
 * one fixed to 0).  Column-major nv x nv. */
int bqb_build_L(int nv, const double *adj, double *L)
{
    if (nv < 2) { bqb_set_error("need at least 2 vertices"); return -1; }
    const int m = nv - 1;
    double total = 0.0;
    for (int i = 0; i < m; ++i) {
        double deg = 0.0;
        for (int j = 0; j < nv; ++j) deg += adj[j + (size_t)i * nv];
        double row = 0.0;
        for (int j = 0; j < m; ++j) {
            const double v = (i == j ? deg : 0.0) - adj[j + (size_t)i * nv];
            L[j + (size_t)i * nv] = v;
            row += v;
        }
        L[m + (size_t)i * nv] = row;     /* column i, last row */
        total += row;
    }
    for (int j = 0; j < m; ++j) L[j + (size_t)m * nv] = L[m + (size_t)j * nv];
    L[m + (size_t)m * nv] = total;
    return 0;
}

/* x' L x over the leading (n0-1) x (n0-1) block (bab_functions.c:96-107). */
double bqb_evaluate_solution(int n0, const double *SP_L, const int *sol)
{
    double v = 0.0;
    for (int i = 0; i < n0 - 1; ++i)
        for (int j = 0; j < n0 - 1; ++j) v += SP_L[j + (size_t)i * n0] * sol[i] * sol[j];
    return v;
}

/* ------------------------------------------------------------ node queue */
/* Cuts a parent hands to its children: records over the original vertices
 * (the last vertex n0-1 is the homogenising one), multipliers >= 1e-5;
 * shared by both children. */
typedef struct inherit {
    int refs;
    int ntri, npent, nhept;
    int *rec;        /* ntri x 4 (i, j, k, type), then (npent + nhept) x 8 (type, vertices) */
    double *y;       /* [ntri + npent + nhept] */
} inherit;

typedef struct qnode {
    double upper_bound;
    int level;
    uint64_t id;
    int *xfixed, *solX;
    double *fracsol;
    inherit *inh;
} qnode;

static void inh_release(inherit *h)
{
    if (h && --h->refs == 0) { free(h->rec); free(h->y); free(h); }
}

static void node_free(qnode *x)
{
    if (x) inh_release(x->inh);
    free(x);
}

/* The final cut set of the node just popped from the context (vertices in
 * its subproblem: its free variables in order, then the last vertex) as
 * records over the original vertices; only the cuts that survive the next
 * purge (multiplier >= 1e-5, cutting_planes.c:151).  NULL if none. */
static inherit *inh_from_popped(const bqb_ctx *ctx, const int *xfixed, int B, int *orig)
{
    int nt, np, nh;
    const int *tri, *pent, *hept;
    const double *y;
    bqb_ctx_last_cuts(ctx, &nt, &tri, &np, &pent, &nh, &hept, &y);
    if (nt + np + nh == 0) return NULL;
    int nf = 0;
    for (int i = 0; i < B; ++i)
        if (!xfixed[i]) orig[nf++] = i;
    orig[nf] = B;
    inherit *h = calloc(1, sizeof *h);
    if (!h) return NULL;
    h->rec = malloc(sizeof(int) * (4 * (size_t)nt + 8 * (size_t)(np + nh)));
    h->y = malloc(sizeof(double) * (size_t)(nt + np + nh));
    if (!h->rec || !h->y) { free(h->rec); free(h->y); free(h); return NULL; }
    h->refs = 1;
    int q = 0;
    for (int c = 0; c < nt; ++c) {
        if (y[c] < 1e-5) continue;
        int *r = h->rec + 4 * h->ntri;
        r[0] = orig[tri[4 * c]]; r[1] = orig[tri[4 * c + 1]]; r[2] = orig[tri[4 * c + 2]]; r[3] = tri[4 * c + 3];
        h->y[q++] = y[c];
        h->ntri++;
    }
    for (int c = 0; c < np + nh; ++c) {
        const double yc = y[nt + c];
        if (yc < 1e-5) continue;
        const int *src = c < np ? pent + 8 * c : hept + 8 * (c - np);
        const int k = c < np ? 5 : 7;
        int *r = h->rec + 4 * h->ntri + 8 * (h->npent + h->nhept);
        r[0] = src[0];
        for (int a = 1; a < 8; ++a) r[a] = a <= k ? orig[src[a]] : 0;
        h->y[q++] = yc;
        if (c < np) h->npent++; else h->nhept++;
    }
    if (q == 0) { inh_release(h); return NULL; }
    return h;
}

/* Pushes the inherited cuts of x into the slot it was just pushed to:
 * vertices renumbered into x's subproblem, cuts containing a variable that
 * x fixes (and its parent did not) left out. */
static int inh_push(bqb_ctx *ctx, int slot, const qnode *x, int B, int *pos)
{
    const inherit *h = x->inh;
    int *tri = malloc(sizeof(int) * (4 * (size_t)h->ntri + 8 * (size_t)(h->npent + h->nhept) + 1));
    double *y = malloc(sizeof(double) * (size_t)(h->ntri + h->npent + h->nhept + 1));
    if (!tri || !y) { free(tri); free(y); bqb_set_error("node queue: out of memory"); return -1; }
    int *kg = tri + 4 * h->ntri;
    int nf = 0;
    for (int i = 0; i < B; ++i) pos[i] = x->xfixed[i] ? -1 : nf++;
    pos[B] = nf;
    int nt = 0, np = 0, nh = 0, q = 0;
    for (int c = 0; c < h->ntri; ++c) {
        const int *r = h->rec + 4 * c;
        const int a = pos[r[0]], b = pos[r[1]], d = pos[r[2]];
        if (a < 0 || b < 0 || d < 0) continue;
        tri[4 * nt] = a; tri[4 * nt + 1] = b; tri[4 * nt + 2] = d; tri[4 * nt + 3] = r[3];
        y[q++] = h->y[c];
        ++nt;
    }
    for (int pass = 0; pass < 2; ++pass) {       /* pentagonal, then heptagonal */
        const int lo = pass ? h->npent : 0, hi = pass ? h->npent + h->nhept : h->npent, k = pass ? 7 : 5;
        for (int c = lo; c < hi; ++c) {
            const int *r = h->rec + 4 * h->ntri + 8 * c;
            int *o = kg + 8 * (np + nh), ok = 1;
            o[0] = r[0];
            for (int a = 1; a < 8; ++a) {
                o[a] = a <= k ? pos[r[a]] : 0;
                if (a <= k && o[a] < 0) ok = 0;
            }
            if (!ok) continue;
            y[q++] = h->y[h->ntri + c];
            if (pass) ++nh; else ++np;
        }
    }
    const int rc = (nt + np + nh == 0) ? 0 : bqb_ctx_push_cuts(ctx, slot, nt, tri, np, kg, nh, kg + 8 * np, y);
    free(tri);
    free(y);
    return rc;
}

typedef struct qheap { qnode **d; long used, cap; } qheap;

static int node_cmp(const qnode *a, const qnode *b) { return a->upper_bound > b->upper_bound ? 1 : -1; }

static int heap_push(qheap *h, qnode *x)
{
    if (h->used == h->cap) {
        const long cap = h->cap ? 2 * h->cap : 4096;
        qnode **d = realloc(h->d, sizeof(qnode *) * cap);
        if (!d) { bqb_set_error("node queue: out of memory"); return -1; }
        h->d = d;
        h->cap = cap;
    }
    long c = h->used++;
    h->d[c] = x;
    while (c > 0) {
        const long par = (c - 1) / 2;
        if (node_cmp(h->d[par], h->d[c]) >= 0) break;
        qnode *t = h->d[par]; h->d[par] = h->d[c]; h->d[c] = t;
        c = par;
    }
    return 0;
}

static qnode *heap_pop(qheap *h)
{
    qnode *top = h->d[0];
    h->d[0] = h->d[--h->used];
    long c = 0;
    for (;;) {
        long k = 2 * c + 1;
        if (k >= h->used) break;
        if (k + 1 < h->used && node_cmp(h->d[k + 1], h->d[k]) > 0) ++k;
        if (node_cmp(h->d[c], h->d[k]) >= 0) break;
        qnode *t = h->d[k]; h->d[k] = h->d[c]; h->d[c] = t;
        c = k;
    }
    return top;
}

static qnode *node_new(const qnode *parent, int B, uint64_t id)
{
    qnode *x = malloc(sizeof(qnode) + (size_t)B * (2 * sizeof(int) + sizeof(double)));
    if (!x) { bqb_set_error("node allocation failed"); return NULL; }
    x->fracsol = (double *)(x + 1);
    x->xfixed = (int *)(x->fracsol + B);
    x->solX = x->xfixed + B;
    x->id = id;
    x->inh = NULL;
    memset(x->fracsol, 0, sizeof(double) * B);
    if (parent) {
        for (int i = 0; i < B; ++i) {
            x->xfixed[i] = parent->xfixed[i];
            x->solX[i] = parent->xfixed[i] ? parent->solX[i] : 0;
        }
        x->level = parent->level + 1;
        x->upper_bound = parent->upper_bound;
    } else {
        memset(x->xfixed, 0, sizeof(int) * B);
        memset(x->solX, 0, sizeof(int) * B);
        x->level = 0;
        x->upper_bound = 0.0;
    }
    return x;
}

/* bab_functions.c:387-421 */
static int branching_variable(const qnode *x, int B, int strategy)
{
    int ic = -1;
    if (strategy == 0) {             /* LEAST_FRACTIONAL */
        double best = -1e9;
        for (int i = 0; i < B; ++i)
            if (!x->xfixed[i] && fabs(0.5 - x->fracsol[i]) > best) { ic = i; best = fabs(0.5 - x->fracsol[i]); }
    } else {                         /* MOST_FRACTIONAL */
        double best = 1e9;
        for (int i = 0; i < B; ++i)
            if (!x->xfixed[i] && fabs(0.5 - x->fracsol[i]) < best) { ic = i; best = fabs(0.5 - x->fracsol[i]); }
    }
    return ic;
}

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

int bqb_solve(int n0, const double *SP_L, const BiqBinParameters *params, uint64_t seed, int batch,
              int *sol, bqb_solve_stats *out)
{
    return bqb_solve_cap(n0, SP_L, params, seed, batch, 0, sol, out);
}

/* Counts a finished node's events in the run statistics. */
static void count_flags(bqb_solve_stats *out, int flags)
{
    if (flags & BQB_FLAG_BUNDLE_FULL) out->bundle_full += 1;
    if (flags & (BQB_FLAG_TRI_TRUNC | BQB_FLAG_PENT_TRUNC | BQB_FLAG_HEPT_TRUNC)) out->cuts_truncated += 1;
}

int bqb_solve_cap(int n0, const double *SP_L, const BiqBinParameters *params, uint64_t seed, int batch,
                  int max_bundle, int *sol, bqb_solve_stats *out)
{
    return bqb_solve_ex(n0, SP_L, params, seed, batch, max_bundle, 0, 0, NULL, 0.f, sol, out);
}

int bqb_solve_ex(int n0, const double *SP_L, const BiqBinParameters *params, uint64_t seed, int batch,
                 int max_bundle, int slices, int certify, const bqb_mlp *model, float drop, int *sol,
                 bqb_solve_stats *out)
{
    const bqb_solve_options o = {batch, max_bundle, slices, certify, model, drop, 0, -1, 0};
    return bqb_solve_opt(n0, SP_L, params, seed, &o, sol, out);
}

/* The children of x (bab_functions.c:387-421 picks the variable), pushed
 * on the heap; with inherited cuts both share x's final cut set. */
static int branch(qheap *h, qnode *x, int B, int strategy, uint64_t *next_id, inherit *inh)
{
    const int ic = branching_variable(x, B, strategy);
    for (int v = 0; v <= 1 && ic >= 0; ++v) {
        qnode *c = node_new(x, B, (*next_id)++);
        if (!c) return -1;
        c->xfixed[ic] = 1;
        c->solX[ic] = v;
        if (inh) { c->inh = inh; inh->refs++; }
        if (heap_push(h, c)) { node_free(c); return -1; }
    }
    return 0;
}

int bqb_solve_opt(int n0, const double *SP_L, const BiqBinParameters *params, uint64_t seed,
                  const bqb_solve_options *opt, int *sol, bqb_solve_stats *out)
{
    memset(out, 0, sizeof *out);
    const double t0 = now_s();
    int batch = opt->batch;
    if (batch < 1) batch = 1;
    const int B = n0 - 1;
    bqb_ctx *ctx = bqb_ctx_create(n0, SP_L, params, batch, opt->max_bundle, 0, 0, 0);
    if (!ctx) return -1;
    const int warm = opt->warm;
    if (bqb_ctx_set_precision(ctx, opt->slices, opt->certify) || bqb_ctx_set_model(ctx, opt->model, opt->drop) ||
        (opt->graphs >= 0 && bqb_ctx_set_graphs(ctx, opt->graphs)) || (warm && bqb_ctx_set_warm(ctx, 1)) ||
        (opt->ipm_warm && bqb_ctx_set_ipm_warm(ctx, 1))) {
        bqb_ctx_destroy(ctx);
        return -1;
    }
    int *vmap = malloc(sizeof(int) * (size_t)(B + 1));
    qheap h = {0};
    int rc = -1;
    uint64_t next_id = 0;
    double LB = 0.0;                       /* bab_functions.c:26-35 */
    memset(sol, 0, sizeof(int) * B);
    double diff = 0.0;
    qnode **inflight = calloc(batch, sizeof(qnode *));      /* by tag */
    int *tag_free = malloc(sizeof(int) * batch), ntag = batch;
    double *frac = malloc(sizeof(double) * B);
    int *xb = malloc(sizeof(int) * B);
    if (!inflight || !tag_free || !frac || !xb || !vmap) { bqb_set_error("calloc"); goto done; }
    for (int i = 0; i < batch; ++i) tag_free[i] = batch - 1 - i;

    /* root node (bab_functions.c:39-69) */
    {
        qnode *root = node_new(NULL, B, next_id++);
        if (!root) goto done;
        double bound, lbo, dfo;
        int imp, st, stats[BQB_NSTATS];
        const double LB0 = LB, df0 = 0.0;
        const int is_root = 1;
        const uint64_t sd = bqb_key(seed, root->id);
        if (bqb_bound_batch(ctx, 1, root->xfixed, root->solX, &LB0, &df0, &is_root, &sd, &bound, frac, xb,
                            &lbo, &imp, &dfo, &st, stats)) {
            free(root);
            goto done;
        }
        out->bound_ms += bqb_ctx_last_ms(ctx);
        out->nodes = 1;
        out->batches = 1;
        count_flags(out, stats[8]);
        if (st) {
            out->status = st;
            bqb_set_error("root node bound failed with status %d", st);
            free(root);
            goto done;
        }
        if (imp && lbo > LB) { LB = lbo; memcpy(sol, xb, sizeof(int) * B); }
        diff = dfo;
        root->upper_bound = bound;
        out->root_bound = bound;
        memcpy(root->fracsol, frac, sizeof(double) * B);
        if (LB + 1.0 < root->upper_bound && !params->root) {
            inherit *inh = warm ? inh_from_popped(ctx, root->xfixed, B, vmap) : NULL;
            const int r = branch(&h, root, B, params->branchingStrategy, &next_id, inh);
            inh_release(inh);
            if (r) { node_free(root); goto done; }
        }
        node_free(root);
    }

    /* node queue: refill free slots with the best open nodes after every
     * outer round, branch every node as soon as its bound is known */
    for (;;) {
        if (params->time_limit > 0 && now_s() - t0 > params->time_limit) { out->stopped = 1; break; }
        if (h.used > out->max_queue) out->max_queue = (int)h.used;
        while (ntag > 0 && h.used > 0) {
            qnode *x = heap_pop(&h);
            if (!(LB + 1.0 < x->upper_bound)) { node_free(x); continue; }   /* cannot beat the incumbent */
            const int t = tag_free[--ntag];
            const int slot = bqb_ctx_push(ctx, t, x->xfixed, x->solX, LB, diff, 0, bqb_key(seed, x->id));
            if (slot < 0 || (x->inh && inh_push(ctx, slot, x, B, vmap))) {
                node_free(x);
                goto done;
            }
            inh_release(x->inh);         /* the cuts are staged in the context */
            x->inh = NULL;
            inflight[t] = x;
        }
        if (bqb_ctx_busy(ctx) == 0) break;
        const double ta = now_s();
        if (bqb_ctx_advance(ctx)) goto done;
        out->bound_ms += 1e3 * (now_s() - ta);
        out->batches += 1;
        long long t;
        double bound, lbo, dfo;
        int imp, st, sts[BQB_NSTATS];
        int fail = 0;
        while (bqb_ctx_pop(ctx, &t, &bound, frac, xb, &lbo, &imp, &dfo, &st, sts)) {
            qnode *x = inflight[t];
            inflight[t] = NULL;
            tag_free[ntag++] = (int)t;
            out->nodes += 1;
            count_flags(out, sts[8]);
            if (st && !out->status) {
                out->status = st;
                bqb_set_error("node bound failed with status %d", st);
            }
            if (imp && lbo > LB) { LB = lbo; memcpy(sol, xb, sizeof(int) * B); }
            x->upper_bound = bound;
            memcpy(x->fracsol, frac, sizeof(double) * B);
            if (!out->status && !fail && LB + 1.0 < x->upper_bound) {
                inherit *inh = warm ? inh_from_popped(ctx, x->xfixed, B, vmap) : NULL;
                if (branch(&h, x, B, params->branchingStrategy, &next_id, inh)) fail = 1;
                inh_release(inh);
            }
            node_free(x);
        }
        if (fail || out->status) goto done;
    }
    rc = 0;
done:
    while (h.used > 0) node_free(heap_pop(&h));
    free(h.d);
    if (inflight)
        for (int i = 0; i < batch; ++i) node_free(inflight[i]);
    free(inflight); free(tag_free); free(frac); free(xb); free(vmap);
    bqb_ctx_destroy(ctx);
    out->best = LB;
    out->seconds = now_s() - t0;
    return rc;
}
