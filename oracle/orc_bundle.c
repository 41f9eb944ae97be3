/*
 * orc_bundle.c -- oracle restatement of the proximal bundle method over the
 * cut multipliers: src/bundle.c:9-239 (bundle_method), 242-275 (fct_eval),
 * 292-373 (lambda_eta), 390-653 (solve_lambda).  TEST INFRASTRUCTURE ONLY.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_internal.h"

#define ABSV(v) ((v) > 0 ? (v) : -(v))

int orc_ws_ensure_G(orc_ws *w, size_t need)
{
    if (need <= w->G_cap) return 0;
    size_t cap = need + need / 2 + 1024;
    double *G = realloc(w->G, cap * sizeof(double));
    if (!G) return ORC_ERR_ALLOC;
    w->G = G;
    w->G_cap = cap;
    return 0;
}

/* Cholesky (lower) of a k x k SPD matrix in place; 0 on success. */
static int chol_small(double *A, int k)
{
    for (int j = 0; j < k; ++j) {
        double d = A[j + j * k];
        for (int l = 0; l < j; ++l) d -= A[j + l * k] * A[j + l * k];
        if (!(d > 0.0)) return j + 1;
        d = sqrt(d);
        A[j + j * k] = d;
        for (int i = j + 1; i < k; ++i) {
            double s = A[i + j * k];
            for (int l = 0; l < j; ++l) s -= A[i + l * k] * A[j + l * k];
            A[i + j * k] = s / d;
        }
    }
    return 0;
}
static void cholsolve_small(const double *Lc, int k, double *b)
{
    for (int i = 0; i < k; ++i) {
        double s = b[i];
        for (int l = 0; l < i; ++l) s -= Lc[i + l * k] * b[l];
        b[i] = s / Lc[i + i * k];
    }
    for (int i = k - 1; i >= 0; --i) {
        double s = b[i];
        for (int l = i + 1; l < k; ++l) s -= Lc[l + i * k] * b[l];
        b[i] = s / Lc[i + i * k];
    }
}

/* bundle.c:390-653.  min <lambda,c> + 1/2 <lambda,Q lambda> over the simplex
 * by a primal-dual IPM.  The KKT matrix [-Q-Diag(z./lambda) e; e' 0] is
 * solved through its Schur complement: with A = Q + Diag(z./lambda) (SPD),
 * u = A\e, w = A\r1: dy = (r2 + e'w)/(e'u), dlambda = u*dy - w. */
int orc_solve_lambda(int k, const double *Q, const double *c, double *lambda)
{
    int status = ORC_OK, cnt = 0;
    double *Qlam = calloc(k, sizeof(double));
    double *z = calloc(k, sizeof(double));
    double *A = calloc((size_t)k * k, sizeof(double));
    double *u = calloc(k, sizeof(double));
    double *wv = calloc(k, sizeof(double));
    double *dlam = calloc(k, sizeof(double));
    double *dz = calloc(k, sizeof(double));
    if (!Qlam || !z || !A || !u || !wv || !dlam || !dz) { status = ORC_ERR_ALLOC; goto out; }

    for (int i = 0; i < k; ++i) lambda[i] = 1.0 / k;
    for (int i = 0; i < k; ++i) {
        double s = 0.0;
        for (int j = 0; j < k; ++j) s += Q[i + j * k] * lambda[j];
        Qlam[i] = s;
    }
    double mintmp = ORC_BIG_NUMBER;
    for (int i = 0; i < k; ++i) {
        double tv = Qlam[i] + c[i];
        z[i] = tv;
        mintmp = (tv < mintmp) ? tv : mintmp;
    }
    double y;
    if (mintmp > 1) y = 0.0;
    else {
        y = mintmp - 1.0;
        for (int i = 0; i < k; ++i) z[i] = z[i] - y;
    }
    double mu = 0.0;
    for (int i = 0; i < k; ++i) mu += z[i] * lambda[i];
    mu = mu / k * 0.5;
    double res_p = 1.0;
    { double s = 0.0; for (int i = 0; i < k; ++i) s += lambda[i]; res_p = 1.0 - s; }
    double temp = 0.0;
    for (int i = 0; i < k; ++i) temp += lambda[i] * Qlam[i];
    double D_cost = y - 0.5 * temp;
    double P_cost = 0.0;
    for (int i = 0; i < k; ++i) P_cost += lambda[i] * c[i];
    P_cost += 0.5 * temp;
    double gap = P_cost - D_cost;

    while (ABSV(gap) > 1e-5) {
        ++cnt;
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < k; ++i) A[i + j * k] = Q[i + j * k];
        for (int i = 0; i < k; ++i) A[i + i * k] = -(-Q[i + i * k] - z[i] / lambda[i]);
        for (int i = 0; i < k; ++i) {
            wv[i] = c[i] - y + Qlam[i];        /* r1 */
            u[i] = 1.0;
        }
        double r2 = res_p;
        if (chol_small(A, k)) { status = ORC_ERR_QP; break; }
        cholsolve_small(A, k, u);
        cholsolve_small(A, k, wv);
        double su = 0.0, sw = 0.0;
        for (int i = 0; i < k; ++i) { su += u[i]; sw += wv[i]; }
        double dy = (r2 + sw) / su;
        for (int i = 0; i < k; ++i) dlam[i] = u[i] * dy - wv[i];
        for (int i = 0; i < k; ++i)
            dz[i] = 1.0 / lambda[i] * (mu - lambda[i] * z[i] - z[i] * dlam[i]);

        double alpha_p = -ORC_BIG_NUMBER, alpha_d = -ORC_BIG_NUMBER;
        for (int i = 0; i < k; ++i) {
            double tv = -dlam[i] / lambda[i];
            alpha_p = (tv > alpha_p) ? tv : alpha_p;
            tv = -dz[i] / z[i];
            alpha_d = (tv > alpha_d) ? tv : alpha_d;
        }
        alpha_p = (alpha_p > 0) ? ((0.99 / alpha_p < 1.0) ? 0.99 / alpha_p : 1.0) : 1.0;
        alpha_d = (alpha_d > 0) ? ((0.99 / alpha_d < 1.0) ? 0.99 / alpha_d : 1.0) : 1.0;

        y += alpha_d * dy;
        for (int i = 0; i < k; ++i) {
            lambda[i] = lambda[i] + alpha_p * dlam[i];
            z[i] = z[i] + alpha_d * dz[i];
        }
        { double s = 0.0; for (int i = 0; i < k; ++i) s += lambda[i]; res_p = 1 - s; }
        for (int i = 0; i < k; ++i) {
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += Q[i + j * k] * lambda[j];
            Qlam[i] = s;
        }
        mu = 0.0;
        for (int i = 0; i < k; ++i) mu += lambda[i] * z[i];
        mu = mu / k * 0.4;
        if (alpha_p + alpha_d > 1.8) mu *= 0.2;
        temp = 0.0;
        for (int i = 0; i < k; ++i) temp += lambda[i] * Qlam[i];
        D_cost = y - 0.5 * temp;
        P_cost = 0.0;
        for (int i = 0; i < k; ++i) P_cost += lambda[i] * c[i];
        P_cost += 0.5 * temp;
        gap = P_cost - D_cost;
        if (cnt > 30) break;
    }
out:
    free(Qlam); free(z); free(A); free(u); free(wv); free(dlam); free(dz);
    return status;
}

/* bundle.c:292-373 */
int orc_lambda_eta(int m, int k, const double *zeta, const double *G, int ldG,
                   const double *gamma, double *dgamma, double *lambda, double *eta, double *t)
{
    int done = 0, cnt = 0, status = ORC_OK;
    double dir_prev = 0.0, dir_curr = 0.0;
    for (int i = 0; i < m; ++i) eta[i] = 0.0;
    double *Q = calloc((size_t)k * k, sizeof(double));
    double *tmp = calloc(m > 0 ? m : 1, sizeof(double));
    double *c = calloc(k, sizeof(double));
    if (!Q || !tmp || !c) { free(Q); free(tmp); free(c); return ORC_ERR_ALLOC; }
    /* Q = t * G'G (dsyrk; with m = 0 it yields Q = 0) */
    for (int j = 0; j < k; ++j)
        for (int i = j; i < k; ++i) {
            double s = 0.0;
            for (int l = 0; l < m; ++l) s += G[l + i * ldG] * G[l + j * ldG];
            Q[i + j * k] = Q[j + i * k] = (m > 0) ? *t * s : 0.0;
        }
    while (!done) {
        ++cnt;
        for (int j = 0; j < k; ++j) {          /* c = zeta - t G' eta (dgemv returns early if m = 0) */
            double s = 0.0;
            for (int l = 0; l < m; ++l) s += G[l + j * ldG] * eta[l];
            c[j] = (m > 0) ? zeta[j] + (-(*t)) * s : zeta[j];
        }
        status = orc_solve_lambda(k, Q, c, lambda);
        if (status) break;
        for (int l = 0; l < m; ++l) {          /* tmp = G lambda */
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += G[l + j * ldG] * lambda[j];
            tmp[l] = s;
        }
        double nrm = 0.0;
        for (int i = 0; i < m; ++i) {
            double e = -gamma[i] / (*t) + tmp[i];
            e = (e > 0) ? e : 0;
            eta[i] = e;
            dgamma[i] = (*t) * (e - tmp[i]);
            nrm += dgamma[i] * dgamma[i];
        }
        dir_curr = sqrt(nrm);
        if (ABSV(dir_curr - dir_prev) / (1.0 + dir_curr) < ORC_LE_TOL) done = 1;
        if (cnt >= ORC_LE_MAXIT) { done = 1; *t *= ORC_T_NOCONV; }
        dir_prev = dir_curr;
    }
    free(Q); free(tmp); free(c);
    return status;
}

/* <L, X> formed as bundle.c:152-162 does (upper triangle doubled). */
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

double orc_fct_eval(const double *L, int n, const orc_tri *tri, int ntri,
                    const orc_pent *pent, int npent, const orc_hept *hept, int nhept,
                    const double *gamma, double *X, double *g, orc_ipm_stats *st)
{
    int m = ntri + npent + nhept;
    double *L0 = malloc(sizeof(double) * n * n);
    memcpy(L0, L, sizeof(double) * n * n);
    if (m > 0) orc_op_Bt(n, tri, ntri, pent, npent, hept, nhept, L0, gamma);
    double f;
    orc_ipm_mc_pk(L0, n, X, &f, st);
    if (m > 0) {
        for (int i = 0; i < m; ++i) { f += gamma[i]; g[i] = 1.0; }
        orc_op_B(n, tri, ntri, pent, npent, hept, nhept, g, X);
    }
    free(L0);
    return f;
}

double orc_ws_fct_eval(orc_ws *w, const double *gamma, double *X, double *g)
{
    orc_ipm_stats st;
    double f = orc_fct_eval(w->L, w->n, w->Cuts, w->NIneq, w->Pent_Cuts, w->NPentIneq,
                            w->Hepta_Cuts, w->NHeptaIneq, gamma, X, g, &st);
    w->ipm_calls++;
    w->ipm_iters += st.iters;
    if (st.status && !w->status) w->status = st.status;
    return f;
}

static void bundle_copy(orc_ws *w, int m, int nn, int from, int to)
{
    w->F[to] = w->F[from];
    memmove(w->G + (size_t)m * to, w->G + (size_t)m * from, sizeof(double) * m);
    memmove(w->X_bundle + (size_t)nn * to, w->X_bundle + (size_t)nn * from, sizeof(double) * nn);
}

/* bundle.c:9-239 */
int orc_bundle_method(orc_ws *w, double *t, int bdl_iter)
{
    int m = w->NIneq + w->NPentIneq + w->NHeptaIneq;
    int n = w->n, nn = n * n;
    double *zeta = calloc(w->max_bundle, sizeof(double));
    if (!zeta) return ORC_ERR_ALLOC;
    int bdl_cnt = 0;
    while (bdl_cnt < bdl_iter) {
        ++bdl_cnt;
        int k = w->bundle;
        for (int j = 0; j < k; ++j) {           /* zeta = -F - G' gamma */
            double s = 0.0;
            for (int l = 0; l < m; ++l) s += w->G[l + (size_t)j * m] * w->gamma[l];
            /* with m = 0 dgemv returns before scaling, leaving zeta = F */
            zeta[j] = (m > 0) ? -w->F[j] + (-1.0) * s : w->F[j];
        }
        int st = orc_lambda_eta(m, k, zeta, w->G, m, w->gamma, w->dgamma, w->lambda, w->eta, t);
        if (st) { w->status = st; break; }
        {
            double nd = 0.0;
            for (int i = 0; i < m; ++i) nd += w->dgamma[i] * w->dgamma[i];
            orc_trace(w, 2, w->count, *t, w->lambda[0], nd, k, m, 0);
        }
        for (int i = 0; i < m; ++i) w->gamma_test[i] = w->gamma[i] + w->dgamma[i];
        double f_test = orc_ws_fct_eval(w, w->gamma_test, w->X_test, w->g);
        if (w->status) break;
        double fa = 0.0;                        /* <F + G' gamma_test, lambda> */
        for (int j = 0; j < k; ++j) {
            double s = 0.0;
            for (int l = 0; l < m; ++l) s += w->G[l + (size_t)j * m] * w->gamma_test[l];
            fa += (w->F[j] + s) * w->lambda[j];
        }
        double del = w->f - fa;
        double lmax = -ORC_BIG_NUMBER;
        for (int i = 0; i < k; ++i) lmax = (w->lambda[i] > lmax) ? w->lambda[i] : lmax;

        const int serious = (w->f - f_test > ORC_SERIOUS_FRAC * del) ? 1 : 0;
        orc_trace(w, 5, w->count, del, w->f, f_test, lmax, k, serious);
        for (int i = 0; i < k; ++i) orc_trace(w, 6, w->count, i, w->lambda[i], lmax, k, serious, 0);
        if (serious) {                          /* serious step */
            memcpy(w->gamma, w->gamma_test, sizeof(double) * m);
            w->f = f_test;
            for (int i = 0; i < nn; ++i) {      /* X = X_bundle * lambda */
                double s = 0.0;
                for (int j = 0; j < k; ++j) s += w->X_bundle[i + (size_t)j * nn] * w->lambda[j];
                w->X[i] = s;
            }
            *t *= ORC_T_STEP;
            int subtracted = 0, next = 0;
            for (int i = 0; i < k; ++i) {
                if (w->lambda[i] < ORC_PURGE_LAMBDA * lmax) ++subtracted;
                else {
                    if (i > next) bundle_copy(w, m, nn, i, next);
                    ++next;
                }
            }
            w->bundle -= subtracted;
            if (w->bundle == w->max_bundle) { w->status = ORC_ERR_BUNDLE_FULL; break; }
            if (orc_ws_ensure_G(w, (size_t)m * (w->bundle + 1))) { w->status = ORC_ERR_ALLOC; break; }
            memcpy(w->G + (size_t)m * w->bundle, w->g, sizeof(double) * m);
            memcpy(w->X_bundle + (size_t)nn * w->bundle, w->X_test, sizeof(double) * nn);
            w->F[w->bundle] = lx_dot(w->L, w->X_test, n);
            ++w->bundle;
        } else {                                /* null step */
            *t /= ORC_T_STEP;
            int subtracted = 0, next = 0;
            for (int i = 0; i < k - 1; ++i) {
                if (w->lambda[i] < ORC_PURGE_LAMBDA * lmax) ++subtracted;
                else {
                    if (i > next) bundle_copy(w, m, nn, i, next);
                    ++next;
                }
            }
            w->bundle -= subtracted;
            if (w->bundle == w->max_bundle) { w->status = ORC_ERR_BUNDLE_FULL; break; }
            if (orc_ws_ensure_G(w, (size_t)m * (w->bundle + 1))) { w->status = ORC_ERR_ALLOC; break; }
            bundle_copy(w, m, nn, k - 1, w->bundle);
            memcpy(w->G + (size_t)m * (w->bundle - 1), w->g, sizeof(double) * m);
            memcpy(w->X_bundle + (size_t)nn * (w->bundle - 1), w->X_test, sizeof(double) * nn);
            w->F[w->bundle - 1] = lx_dot(w->L, w->X_test, n);
            ++w->bundle;
        }
        if (w->bundle > w->max_bundle_seen) w->max_bundle_seen = w->bundle;
        orc_trace(w, 1, w->count, f_test, w->f, *t, w->bundle, m, serious);
    }
    free(zeta);
    return w->status;
}
