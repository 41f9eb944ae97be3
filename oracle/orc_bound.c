/*
 * orc_bound.c -- oracle restatement of one node's bound and of the
 * single-process branch and bound around it:
 *   src/evaluate.c:11-138 (Evaluate, createSubproblem, getFixedValue)
 *   src/bounding.c:24-342 (SDPbound)
 *   src/cutting_planes.c:132-214, 316-406, 509-602 (purge / dedupe / append)
 *   src/process_input.c:259-344 (SP->L from the adjacency matrix)
 *   src/heap.c, src/bab_functions.c:39-69, 96-134, 224-332, 387-436 (B&B)
 * TEST INFRASTRUCTURE ONLY.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "orc_internal.h"

void orc_params_default(orc_params *p)
{
    p->init_bundle_iter = 3;  p->max_bundle_iter = 15; p->triag_iter = 5;
    p->pent_iter = 5;         p->hept_iter = 5;        p->max_outer_iter = 20;
    p->extra_iter = 10;       p->violated_TriIneq = 1e-3; p->TriIneq = 5000;
    p->adjust_TriIneq = 1;    p->PentIneq = 5000;      p->HeptaIneq = 5000;
    p->Pent_Trials = 60;      p->Hepta_Trials = 50;    p->include_Pent = 1;
    p->include_Hepta = 1;     p->root = 0;             p->use_diff = 1;
    p->time_limit = 0;        p->branchingStrategy = 1;
}

orc_ws *orc_ws_create(int n0, int max_bundle, int max_tri, int max_pent, int max_hept,
                      const orc_params *p)
{
    orc_ws *w = calloc(1, sizeof(orc_ws));
    if (!w) return NULL;
    size_t nn = (size_t)n0 * n0;
    size_t mmax = (size_t)max_tri + max_pent + max_hept;
    w->p = *p;
    w->n0 = n0;
    w->max_bundle = max_bundle; w->max_tri = max_tri;
    w->max_pent = max_pent; w->max_hept = max_hept;
    w->L = calloc(nn, sizeof(double));
    w->X = calloc(nn, sizeof(double));
    w->Z = calloc(nn, sizeof(double));
    w->X_test = calloc(nn, sizeof(double));
    w->X_bundle = calloc(nn * max_bundle, sizeof(double));
    w->gamma = calloc(mmax, sizeof(double));
    w->dgamma = calloc(mmax, sizeof(double));
    w->gamma_test = calloc(mmax, sizeof(double));
    w->eta = calloc(mmax, sizeof(double));
    w->g = calloc(mmax, sizeof(double));
    w->lambda = calloc(max_bundle, sizeof(double));
    w->F = calloc(max_bundle, sizeof(double));
    w->Cuts = calloc(max_tri, sizeof(orc_tri));
    w->List = calloc(p->TriIneq > 0 ? p->TriIneq : 1, sizeof(orc_tri));
    w->Pent_Cuts = calloc(max_pent, sizeof(orc_pent));
    w->Pent_List = calloc(p->PentIneq > 0 ? p->PentIneq : 1, sizeof(orc_pent));
    w->Hepta_Cuts = calloc(max_hept, sizeof(orc_hept));
    w->Hepta_List = calloc(p->HeptaIneq > 0 ? p->HeptaIneq : 1, sizeof(orc_hept));
    w->xbest = calloc(n0, sizeof(int));
    if (!w->L || !w->X || !w->Z || !w->X_test || !w->X_bundle || !w->gamma || !w->dgamma ||
        !w->gamma_test || !w->eta || !w->g || !w->lambda || !w->F || !w->Cuts || !w->List ||
        !w->Pent_Cuts || !w->Pent_List || !w->Hepta_Cuts || !w->Hepta_List || !w->xbest) {
        orc_ws_destroy(w);
        return NULL;
    }
    return w;
}

void orc_ws_set_trace(orc_ws *w, double *buf, int cap)
{
    w->trace = buf;
    w->trace_cap = buf ? cap : 0;
    w->trace_n = 0;
}
int orc_ws_trace_n(const orc_ws *w) { return w->trace_n; }

void orc_ws_destroy(orc_ws *w)
{
    if (!w) return;
    free(w->L); free(w->X); free(w->Z); free(w->X_test); free(w->X_bundle);
    free(w->gamma); free(w->dgamma); free(w->gamma_test); free(w->eta); free(w->g);
    free(w->lambda); free(w->F); free(w->G); free(w->Cuts); free(w->List);
    free(w->Pent_Cuts); free(w->Pent_List); free(w->Hepta_Cuts); free(w->Hepta_List);
    free(w->xbest);
    free(w);
}

/* process_input.c:259-344 */
void orc_build_L_from_adj(const double *Adj, int nv, double *L)
{
    double *Adje = calloc(nv, sizeof(double));
    for (int ii = 0; ii < nv; ++ii)
        for (int jj = 0; jj < nv; ++jj)
            Adje[ii] += Adj[jj + ii * nv];
    double sum_row = 0.0, sum = 0.0;
    for (int ii = 0; ii < nv; ++ii) {
        for (int jj = 0; jj < nv; ++jj) {
            if (ii < nv - 1 && jj < nv - 1) {
                double lap = ((ii == jj) ? Adje[ii] : 0.0) - Adj[jj + ii * nv];
                L[jj + ii * nv] = lap;
                sum_row += lap;
            } else if (jj == nv - 1 && ii != nv - 1) {
                L[jj + ii * nv] = sum_row;
                sum += sum_row;
            } else if (ii == nv - 1 && jj != nv - 1) {
                L[jj + ii * nv] = L[ii + jj * nv];
            } else {
                L[jj + ii * nv] = sum;
            }
        }
        sum_row = 0.0;
    }
    free(Adje);
}

/* evaluate.c:38-115 */
int orc_create_subproblem(const double *SP_L, int n0, const int *xfixed, const int *solX,
                          double *PL)
{
    int N = n0, B = n0 - 1, nfix = 0;
    for (int i = 0; i < B; ++i) nfix += xfixed[i] ? 1 : 0;
    int n = B + 1 - nfix;
    double *fixedRow = calloc(n > 1 ? n - 1 : 1, sizeof(double));
    int index = 0, fixed = 0;
    double row_sum = 0.0, sum = 0.0;
    for (int i = 0; i < B; ++i) {
        for (int j = 0; j < B; ++j) {
            if (!xfixed[i] && !xfixed[j]) {
                PL[index] = SP_L[j + i * N];
                row_sum += PL[index];
                ++index;
            } else if ((xfixed[i] && solX[i] == 1) && !xfixed[j]) {
                fixedRow[fixed] += SP_L[j + i * N];
                ++fixed;
            }
        }
        if (!xfixed[i]) { PL[index] = row_sum; ++index; }
        row_sum = 0.0;
        fixed = 0;
    }
    for (int i = 0; i < n - 1; ++i)
        PL[i + (n - 1) * n] = PL[n - 1 + i * n];
    for (int i = 0; i < n - 1; ++i) {
        PL[i + i * n] += 2 * fixedRow[i];
        PL[n - 1 + i * n] += 2 * fixedRow[i];
        PL[i + (n - 1) * n] += 2 * fixedRow[i];
        sum += PL[i + (n - 1) * n];
    }
    PL[n - 1 + (n - 1) * n] = sum;
    for (int i = 0; i < n * n; ++i) PL[i] *= 0.25;
    free(fixedRow);
    return n;
}

/* evaluate.c:123-138 */
double orc_get_fixed_value(const double *SP_L, int n0, const int *xfixed, const int *solX)
{
    double fv = 0.0;
    for (int i = 0; i < n0 - 1; ++i)
        for (int j = 0; j < n0 - 1; ++j)
            if (xfixed[i] && xfixed[j])
                fv += SP_L[j + i * n0] * solX[i] * solX[j];
    return fv;
}

/* bab_functions.c:114-134 against the node-local incumbent. */
static void update_solution(orc_ws *w, const int *x)
{
    double v = orc_evaluate_solution(w->SP_L, w->n0, x);
    if (v > w->LB) {
        w->LB = v;
        memcpy(w->xbest, x, sizeof(int) * (w->n0 - 1));
        w->improved = 1;
    }
}

/* cutting_planes.c:132-214 */
static double update_triangles(orc_ws *w, const double *y)
{
    int subtracted = 0, next = 0;
    for (int c = 0; c < w->NIneq; ++c) {
        w->Cuts[c].y = y[c];
        if (w->Cuts[c].y < ORC_CUT_PURGE) ++subtracted;
        else {
            orc_tri keep = w->Cuts[c];
            w->Cuts[next].type = keep.type; w->Cuts[next].i = keep.i;
            w->Cuts[next].j = keep.j; w->Cuts[next].k = keep.k; w->Cuts[next].y = keep.y;
            ++next;
        }
    }
    w->NIneq -= subtracted;
    int size;
    double maxall = orc_get_violated_triangles(w->X, w->n, w->p.violated_TriIneq,
                                               w->p.TriIneq, w->List, &size);
    int added = 0;
    for (int l = 0; l < size; ++l) {
        if (next == w->max_tri) break;
        int found = 0;
        for (int c = 0; c < w->NIneq; ++c)
            if (w->Cuts[c].type == w->List[l].type && w->Cuts[c].i == w->List[l].i &&
                w->Cuts[c].j == w->List[l].j && w->Cuts[c].k == w->List[l].k)
                found = 1;
        if (!found) {
            w->Cuts[next] = w->List[l];
            w->Cuts[next].y = 0.0;
            ++next; ++added;
        }
    }
    w->NIneq += added;
    return maxall;
}

/* Test helper: updateTriangleInequalities (cutting_planes.c:132-214) on a
 * given n x n matrix X and cut list.  cuts (capacity max_tri) holds *ncuts
 * cuts with multipliers y[*ncuts]; it is purged and extended in place.  list
 * (capacity p->TriIneq) receives the separated list and *list_size its
 * length.  Returns the largest violation. */
double orc_update_triangles_on(const double *X, int n, const orc_params *p, orc_tri *cuts, int *ncuts,
                               int max_tri, const double *y, orc_tri *list, int *list_size)
{
    orc_ws *w = orc_ws_create(n, 1, max_tri, 1, 1, p);
    if (!w) return -ORC_BIG_NUMBER;
    w->n = n;
    memcpy(w->X, X, sizeof(double) * n * n);
    memcpy(w->Cuts, cuts, sizeof(orc_tri) * *ncuts);
    w->NIneq = *ncuts;
    const double mx = update_triangles(w, y);
    memcpy(cuts, w->Cuts, sizeof(orc_tri) * w->NIneq);
    *ncuts = w->NIneq;
    int size = 0;
    orc_get_violated_triangles(X, n, p->violated_TriIneq, p->TriIneq, list, &size);
    *list_size = size;
    orc_ws_destroy(w);
    return mx;
}

/* cutting_planes.c:316-406 (K=5) and 509-602 (K=7) */
#define DEFINE_KGON_UPDATE(NAME, TYPE_T, K, CUTS, LIST, COUNT, MAXC, TRIALS, LISTCAP, SEP)   \
static double NAME(orc_ws *w, const double *y, int offset, uint64_t key)                    \
{                                                                                           \
    int subtracted = 0, next = 0;                                                           \
    for (int c = 0; c < w->COUNT; ++c) {                                                    \
        w->CUTS[c].y = y[offset + c];                                                       \
        if (w->CUTS[c].y < ORC_CUT_PURGE) ++subtracted;                                     \
        else {                                                                              \
            TYPE_T keep = w->CUTS[c];                                                       \
            w->CUTS[next].type = keep.type; w->CUTS[next].y = keep.y;                       \
            for (int i = 0; i < K; ++i) w->CUTS[next].perm[i] = keep.perm[i];               \
            ++next;                                                                         \
        }                                                                                   \
    }                                                                                       \
    w->COUNT -= subtracted;                                                                 \
    int size, st = 0;                                                                       \
    double minall = SEP(w->X, w->n, w->p.TRIALS, w->p.LISTCAP, w->LIST, &size, key, &st);  \
    if (st && !w->status) w->status = st;                                                   \
    int added = 0;                                                                          \
    for (int l = 0; l < size; ++l) {                                                        \
        if (next == w->MAXC) break;                                                         \
        int found = 0;                                                                      \
        for (int c = 0; c < w->COUNT; ++c) {                                                \
            int same = (w->CUTS[c].type == w->LIST[l].type);                                \
            for (int i = 0; i < K; ++i) same = same && (w->CUTS[c].perm[i] == w->LIST[l].perm[i]); \
            if (same) found = 1;                                                            \
        }                                                                                   \
        if (!found) {                                                                       \
            w->CUTS[next] = w->LIST[l];                                                     \
            w->CUTS[next].y = 0.0;                                                          \
            ++next; ++added;                                                                \
        }                                                                                   \
    }                                                                                       \
    w->COUNT += added;                                                                      \
    return minall;                                                                          \
}
DEFINE_KGON_UPDATE(update_pent, orc_pent, 5, Pent_Cuts, Pent_List, NPentIneq, max_pent,
                   Pent_Trials, PentIneq, orc_get_violated_pent)
DEFINE_KGON_UPDATE(update_hept, orc_hept, 7, Hepta_Cuts, Hepta_List, NHeptaIneq, max_hept,
                   Hepta_Trials, HeptaIneq, orc_get_violated_hept)

static double lx_dot(const double *L, const double *X, int n)
{
    double temp = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            if (i == j) temp += L[i + i * n] * X[i + i * n];
            else temp += 2 * L[j + i * n] * X[j + i * n];
        }
    return temp;
}

static void store_fracsol(orc_ws *w, const int *xfixed, const int *solX, double *fracsol)
{
    int index = 0, n = w->n;
    for (int i = 0; i < w->n0 - 1; ++i) {
        if (xfixed[i]) fracsol[i] = (double)solX[i];
        else { fracsol[i] = 0.5 * (w->X[(n - 1) + index * n] + 1.0); ++index; }
    }
}

/* Keys of the random streams of one node. */
#define KEY_HEUR(seed, h) orc_key((seed), 1000u + (uint64_t)(h))
#define KEY_PENT(seed, c) orc_key((seed), 2000u + (uint64_t)(c))
#define KEY_HEPT(seed, c) orc_key((seed), 3000u + (uint64_t)(c))

/* bounding.c:24-342 */
void orc_sdpbound(orc_ws *w, const double *SP_L, int n0, const orc_node_in *in,
                  orc_node_out *out, double *fracsol, int *xbest_out)
{
    const orc_params *p = &w->p;
    int B = n0 - 1;
    int *x = malloc(sizeof(int) * B);
    memset(out, 0, sizeof(*out));
    w->SP_L = SP_L;
    w->LB = in->LB;
    w->improved = 0;
    w->status = 0;
    w->ipm_calls = w->ipm_iters = w->max_cuts = w->max_bundle_seen = 0;
    w->trace_n = 0;
    w->count = 0;
    double diff = in->diff;
    int prune = 0, giveup = 0, done = 0, count = 0;
    int bdl_iter = p->init_bundle_iter;
    double viol3 = 0.0, viol5 = 0.0, t, bound, oldf;

    w->n = orc_create_subproblem(SP_L, n0, in->xfixed, in->solX, w->L);
    int n = w->n, nn = n * n;
    double fixedvalue = orc_get_fixed_value(SP_L, n0, in->xfixed, in->solX);
    w->NIneq = w->NPentIneq = w->NHeptaIneq = 0;

    {
        orc_ipm_stats st;
        orc_ipm_mc_pk(w->L, n, w->X, &w->f, &st);
        w->ipm_calls++; w->ipm_iters += st.iters;
        if (st.status) { w->status = st.status; bound = w->f + fixedvalue; goto END; }
    }
    double basic_bound = w->f + fixedvalue;
    out->basic_bound = basic_bound;
    store_fracsol(w, in->xfixed, in->solX, fracsol);

    for (int i = 0; i < B; ++i) x[i] = in->xfixed[i] ? in->solX[i] : 0;
    {
        double hv = orc_run_heuristic(w, in->xfixed, in->solX, x, KEY_HEUR(in->seed, 0));
        update_solution(w, x);
        orc_trace(w, 4, 0, w->LB, hv, 0, 0, 0, 0);
    }
    bound = w->f + fixedvalue;
    if (w->status) goto END;
    if (bound < w->LB + 1.0) { prune = 1; goto END; }
    if (p->use_diff && !in->is_root && (bound > w->LB + diff + 1.0)) { giveup = 1; goto END; }

    viol3 = update_triangles(w, w->gamma);
    for (int i = 0; i < w->NIneq; ++i) w->gamma[i] = w->Cuts[i].y;
    t = 0.5 * (bound - w->LB) / (w->NIneq * viol3 * viol3);
    for (int i = 0; i < w->NIneq; ++i) w->g[i] = 1.0;
    orc_op_B(n, w->Cuts, w->NIneq, NULL, 0, NULL, 0, w->g, w->X);
    w->F[0] = lx_dot(w->L, w->X, n);
    if (orc_ws_ensure_G(w, (size_t)w->NIneq + 1)) { w->status = ORC_ERR_ALLOC; goto END; }
    memcpy(w->G, w->g, sizeof(double) * w->NIneq);
    memcpy(w->X_bundle, w->X, sizeof(double) * nn);
    w->bundle = 1;

    while (!done) {
        ++count;
        w->count = count;
        oldf = w->f;
        orc_bundle_method(w, &t, bdl_iter);
        if (w->status) { bound = w->f + fixedvalue; goto END; }
        bound = w->f + fixedvalue;
        prune = (bound < w->LB + 1.0) ? 1 : 0;
        if (!prune) {
            for (int i = 0; i < B; ++i) x[i] = in->xfixed[i] ? in->solX[i] : 0;
            double hv = orc_run_heuristic(w, in->xfixed, in->solX, x, KEY_HEUR(in->seed, count));
            update_solution(w, x);
            orc_trace(w, 4, count, w->LB, hv, 0, 0, 0, 0);
            if (w->status) goto END;
            prune = (bound < w->LB + 1.0) ? 1 : 0;
        }
        double gap = bound - w->LB;
        if (count == p->triag_iter + p->pent_iter + p->hept_iter)
            if (gap - 1.0 > (oldf - w->f) * (p->max_outer_iter - count)) giveup = 1;
        if (count == p->max_outer_iter)
            if (gap - 1.0 > (oldf - w->f) * p->extra_iter) giveup = 1;
        if (count == p->max_outer_iter + p->extra_iter) giveup = 1;

        if (!prune && !giveup) {
            int triag = w->NIneq, penta = w->NPentIneq;
            viol3 = update_triangles(w, w->gamma);
            if (p->include_Pent && (count > p->triag_iter || viol3 < 0.2))
                viol5 = update_pent(w, w->gamma, triag, KEY_PENT(in->seed, count));
            if (p->include_Hepta && ((count > p->triag_iter + p->pent_iter) ||
                                     (viol3 < 0.2 && (1 - viol5 < 0.4))))
                update_hept(w, w->gamma, triag + penta, KEY_HEPT(in->seed, count));
            if (w->status) goto END;
        }
        done = prune || giveup;
        orc_trace(w, 3, count, w->NIneq, w->NPentIneq, w->NHeptaIneq, viol3, w->LB, bound);
        store_fracsol(w, in->xfixed, in->solX, fracsol);

        if (!done) {
            int m = w->NIneq + w->NPentIneq + w->NHeptaIneq;
            if (m > w->max_cuts) w->max_cuts = m;
            for (int i = 0; i < w->NIneq; ++i) w->gamma[i] = w->Cuts[i].y;
            for (int i = 0; i < w->NPentIneq; ++i) w->gamma[i + w->NIneq] = w->Pent_Cuts[i].y;
            for (int i = 0; i < w->NHeptaIneq; ++i)
                w->gamma[i + w->NIneq + w->NPentIneq] = w->Hepta_Cuts[i].y;
            orc_ws_fct_eval(w, w->gamma, w->X_test, w->g);
            if (w->status) goto END;
            if (orc_ws_ensure_G(w, (size_t)m * (w->bundle + 1))) { w->status = ORC_ERR_ALLOC; goto END; }
            for (size_t i = 0; i < (size_t)m * w->bundle; ++i) w->G[i] = 1.0;
            for (int i = 0; i < w->bundle; ++i)
                orc_op_B(n, w->Cuts, w->NIneq, w->Pent_Cuts, w->NPentIneq, w->Hepta_Cuts,
                         w->NHeptaIneq, w->G + (size_t)i * m, w->X_bundle + (size_t)i * nn);
            memcpy(w->G + (size_t)w->bundle * m, w->g, sizeof(double) * m);
            w->F[w->bundle] = lx_dot(w->L, w->X_test, n);
            memcpy(w->X_bundle + (size_t)w->bundle * nn, w->X_test, sizeof(double) * nn);
            if (w->bundle == w->max_bundle) { w->status = ORC_ERR_BUNDLE_FULL; goto END; }
            ++w->bundle;
            if (w->bundle > w->max_bundle_seen) w->max_bundle_seen = w->bundle;
            t *= ORC_T_REBUILD;
        }
        bdl_iter += count % 2;
        bdl_iter = (bdl_iter < p->max_bundle_iter) ? bdl_iter : p->max_bundle_iter;
    }
    bound = w->f + fixedvalue;
    if (in->is_root) diff = basic_bound - bound;

END:
    out->bound = bound;
    out->LB = w->LB;
    out->improved = w->improved;
    out->diff = diff;
    out->prune = prune;
    out->giveup = giveup;
    out->outer_iters = count;
    out->ipm_calls = w->ipm_calls;
    out->ipm_iters = w->ipm_iters;
    out->max_cuts = w->max_cuts;
    out->max_bundle = w->max_bundle_seen;
    out->status = w->status;
    if (xbest_out) memcpy(xbest_out, w->xbest, sizeof(int) * B);
    free(x);
}

/* ---------------- single-process branch and bound ---------------- */
typedef struct bnode {
    int *xfixed, *solX;
    double *fracsol;
    double upper_bound;
    int level;
    uint64_t id;
} bnode;

typedef struct bheap { bnode **d; int used, cap; } bheap;

/* heap.c:27-133: max-heap on upper_bound (ties: the comparison returns -1). */
static int cmp_nodes(const bnode *a, const bnode *b) { return a->upper_bound > b->upper_bound ? 1 : -1; }
static void heap_push(bheap *h, bnode *nd)
{
    if (h->used == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 1024;
        h->d = realloc(h->d, sizeof(bnode *) * h->cap);
    }
    int cur = h->used++;
    h->d[cur] = nd;
    while (cur > 0) {
        int par = (cur - 1) / 2;
        if (cmp_nodes(h->d[par], h->d[cur]) >= 0) break;
        bnode *t = h->d[par]; h->d[par] = h->d[cur]; h->d[cur] = t;
        cur = par;
    }
}
static bnode *heap_pop(bheap *h)
{
    bnode *top = h->d[0];
    h->used--;
    h->d[0] = h->d[h->used];
    int cur = 0;
    for (;;) {
        int c = 2 * cur + 1;
        if (c >= h->used) break;
        if (c + 1 < h->used && cmp_nodes(h->d[c + 1], h->d[c]) > 0) c++;
        if (cmp_nodes(h->d[cur], h->d[c]) >= 0) break;
        bnode *t = h->d[c]; h->d[c] = h->d[cur]; h->d[cur] = t;
        cur = c;
    }
    return top;
}

static bnode *new_node(const bnode *parent, int B, uint64_t id)
{
    bnode *nd = calloc(1, sizeof(bnode));
    nd->xfixed = calloc(B, sizeof(int));
    nd->solX = calloc(B, sizeof(int));
    nd->fracsol = calloc(B, sizeof(double));
    nd->id = id;
    if (parent) {
        for (int i = 0; i < B; ++i) {
            nd->xfixed[i] = parent->xfixed[i];
            nd->solX[i] = parent->xfixed[i] ? parent->solX[i] : 0;
        }
        nd->level = parent->level + 1;
        nd->upper_bound = parent->upper_bound;   /* children keyed on the parent bound */
    }
    return nd;
}
static void free_node(bnode *nd) { free(nd->xfixed); free(nd->solX); free(nd->fracsol); free(nd); }

/* bab_functions.c:387-421 */
static int branching_variable(const bnode *nd, int B, int strategy)
{
    int ic = -1;
    if (strategy == 0) {
        double maxv = -ORC_BIG_NUMBER;
        for (int i = 0; i < B; ++i)
            if (!nd->xfixed[i] && fabs(0.5 - nd->fracsol[i]) > maxv) { ic = i; maxv = fabs(0.5 - nd->fracsol[ic]); }
    } else {
        double minv = ORC_BIG_NUMBER;
        for (int i = 0; i < B; ++i)
            if (!nd->xfixed[i] && fabs(0.5 - nd->fracsol[i]) < minv) { ic = i; minv = fabs(0.5 - nd->fracsol[ic]); }
    }
    return ic;
}

void orc_bab_solve(const double *SP_L, int n0, const orc_params *p0, uint64_t seed, int batch,
                   int *sol, orc_bab_out *out)
{
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    orc_params p = *p0;
    if (p.adjust_TriIneq) p.TriIneq = n0 * 10;       /* process_input.c:135-136 */
    int B = n0 - 1;
    orc_ws *w = orc_ws_create(n0, 400, 50000, 50000, 50000, &p);
    memset(out, 0, sizeof(*out));
    double LB = 0.0;                                 /* bab_functions.c:26-35 */
    memset(sol, 0, sizeof(int) * B);
    double diff = 0.0;
    uint64_t next_id = 0;
    bheap h = {0};
    int *xb = malloc(sizeof(int) * B);
    if (batch < 1) batch = 1;

    bnode *root = new_node(NULL, B, next_id++);
    {
        orc_node_in in = {root->xfixed, root->solX, LB, 0.0, 1, orc_key(seed, root->id)};
        orc_node_out o;
        orc_sdpbound(w, SP_L, n0, &in, &o, root->fracsol, xb);
        out->nodes++;
        if (o.status) out->status = o.status;
        if (o.improved && o.LB > LB) { LB = o.LB; memcpy(sol, xb, sizeof(int) * B); }
        diff = o.diff;
        root->upper_bound = o.bound;
        out->root_bound = o.bound;
        if (LB + 1.0 < root->upper_bound && !p.root) {
            int ic = branching_variable(root, B, p.branchingStrategy);
            for (int xic = 0; xic <= 1; ++xic) {
                bnode *c = new_node(root, B, next_id++);
                c->xfixed[ic] = 1;
                c->solX[ic] = xic;
                heap_push(&h, c);
            }
        }
        free_node(root);
    }
    bnode **cur = malloc(sizeof(bnode *) * batch);
    while (h.used > 0 && !out->status) {
        int nb = 0;
        while (nb < batch && h.used > 0) {
            bnode *nd = heap_pop(&h);
            if (LB + 1.0 < nd->upper_bound) cur[nb++] = nd;   /* stale node: prune */
            else free_node(nd);
        }
        double LBsnap = LB;
        for (int b = 0; b < nb; ++b) {
            bnode *nd = cur[b];
            orc_node_in in = {nd->xfixed, nd->solX, LBsnap, diff, 0, orc_key(seed, nd->id)};
            orc_node_out o;
            orc_sdpbound(w, SP_L, n0, &in, &o, nd->fracsol, xb);
            out->nodes++;
            if (o.status) { out->status = o.status; break; }
            if (o.improved && o.LB > LB) { LB = o.LB; memcpy(sol, xb, sizeof(int) * B); }
            nd->upper_bound = o.bound;
        }
        for (int b = 0; b < nb; ++b) {
            bnode *nd = cur[b];
            if (!out->status && LB + 1.0 < nd->upper_bound) {
                int ic = branching_variable(nd, B, p.branchingStrategy);
                if (ic >= 0)
                    for (int xic = 0; xic <= 1; ++xic) {
                        bnode *c = new_node(nd, B, next_id++);
                        c->xfixed[ic] = 1;
                        c->solX[ic] = xic;
                        heap_push(&h, c);
                    }
            }
            free_node(nd);
        }
    }
    while (h.used > 0) free_node(heap_pop(&h));
    free(h.d); free(cur); free(xb);
    orc_ws_destroy(w);
    out->best = LB;
    clock_gettime(CLOCK_MONOTONIC, &t1);
    out->seconds = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}
