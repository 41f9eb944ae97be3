/*
 * orc_ipm.c -- oracle restatement of the basic Max-Cut SDP interior-point
 * method (reference src/ipm_mc_pk.c:28-359) and the dense kernels it needs
 * (the LAPACK routines dpotrf/dpotri/dposv/dsymm/dsymv it calls), written as
 * plain loops.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * All matrices are column-major: A[i + j*n] is row i, column j.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define IPM_MAX_ITERS 500        /* the reference loop (ipm_mc_pk.c:126) is unbounded */
#define LINESEARCH_MAX_TRIALS 500 /* ipm_mc_pk.c:283-290 and 299-306 are unbounded */

/* dpotrf('U') semantics: A = U'U, U in the upper triangle.  Returns 0 on
 * success or j+1 where the j-th leading minor is not positive definite. */
int orc_chol_upper(double *A, int n)
{
    for (int j = 0; j < n; ++j) {
        double ajj = A[j + j * n];
        for (int k = 0; k < j; ++k)
            ajj -= A[k + j * n] * A[k + j * n];
        if (!(ajj > 0.0))
            return j + 1;
        ajj = sqrt(ajj);
        A[j + j * n] = ajj;
        for (int i = j + 1; i < n; ++i) {
            double s = A[j + i * n];
            for (int k = 0; k < j; ++k)
                s -= A[k + j * n] * A[k + i * n];
            A[j + i * n] = s / ajj;
        }
    }
    return 0;
}

/* dpotrf('L') semantics: A = LL', L in the lower triangle; the strictly upper
 * triangle is zeroed afterwards as heuristic.c:76-79 does. */
int orc_chol_lower(double *A, int n)
{
    for (int j = 0; j < n; ++j) {
        double ajj = A[j + j * n];
        for (int k = 0; k < j; ++k)
            ajj -= A[j + k * n] * A[j + k * n];
        if (!(ajj > 0.0))
            return j + 1;
        ajj = sqrt(ajj);
        A[j + j * n] = ajj;
        for (int i = j + 1; i < n; ++i) {
            double s = A[i + j * n];
            for (int k = 0; k < j; ++k)
                s -= A[i + k * n] * A[j + k * n];
            A[i + j * n] = s / ajj;
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j)
            A[j + i * n] = 0.0;
    return 0;
}

/* dpotri('U') after dpotrf('U'): upper triangle of A becomes inv(U'U).
 * W = inv(U) (upper triangular), then inv = W W'. */
static void potri_upper(double *A, int n, double *W)
{
    memset(W, 0, sizeof(double) * (size_t)n * n);
    for (int j = 0; j < n; ++j) {
        W[j + j * n] = 1.0 / A[j + j * n];
        for (int i = j - 1; i >= 0; --i) {
            /* W(i,j) = -( sum_{k=i+1..j} U(i,k) W(k,j) ) / U(i,i) */
            double s = 0.0;
            for (int k = i + 1; k <= j; ++k)
                s += A[i + k * n] * W[k + j * n];
            W[i + j * n] = -s / A[i + i * n];
        }
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) {
            double s = 0.0;
            for (int k = j; k < n; ++k)
                s += W[i + k * n] * W[j + k * n];
            A[i + j * n] = s;
        }
}

/* Solve (U'U) x = b in place, U from orc_chol_upper (dposv's solve phase). */
static void potrs_upper(const double *U, int n, double *b)
{
    for (int i = 0; i < n; ++i) {            /* U' z = b */
        double s = b[i];
        for (int k = 0; k < i; ++k)
            s -= U[k + i * n] * b[k];
        b[i] = s / U[i + i * n];
    }
    for (int i = n - 1; i >= 0; --i) {       /* U x = z */
        double s = b[i];
        for (int k = i + 1; k < n; ++k)
            s -= U[i + k * n] * b[k];
        b[i] = s / U[i + i * n];
    }
}

/* C = A*B, A symmetric (full storage, already mirrored), all n x n. */
static void symm(const double *A, const double *B, double *C, int n)
{
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i)
            C[i + j * n] = 0.0;
        for (int k = 0; k < n; ++k) {
            double b = B[k + j * n];
            for (int i = 0; i < n; ++i)
                C[i + j * n] += A[i + k * n] * b;
        }
    }
}

static void symmetrize(double *A, int n)
{
    for (int j = 0; j < n; ++j)
        for (int k = 0; k <= j; ++k)
            A[k + n * j] = A[j + n * k] = 0.5 * (A[j + n * k] + A[k + n * j]);
}

void orc_ipm_mc_pk_ex(const double *L, int n, double *X, double *y_out, double *phi_out,
                      orc_ipm_stats *st, double *alpha_trace, int trace_cap)
{
    const int nn = n * n;
    double *y = calloc(n, sizeof(double));
    double *Z = calloc(nn, sizeof(double));
    double *dX = calloc(nn, sizeof(double));
    double *dy = calloc(n, sizeof(double));
    double *dy1 = calloc(n, sizeof(double));
    double *dX1 = calloc(nn, sizeof(double));
    double *tmp = calloc(nn, sizeof(double));
    double *tmp2 = calloc(nn, sizeof(double));
    double *Zi = calloc(nn, sizeof(double));
    double *M = calloc(nn, sizeof(double));
    double *W = calloc(nn, sizeof(double));
    orc_ipm_stats s = {0, 0, ORC_OK};
    double phi = 0.0;

    if (!y || !Z || !dX || !dy || !dy1 || !dX1 || !tmp || !tmp2 || !Zi || !M || !W) {
        s.status = ORC_ERR_ALLOC;
        goto done;
    }

    /* ipm_mc_pk.c:69-76: y = 1.1 + sum(abs(L))' (column sums) */
    for (int j = 0; j < n; ++j) {
        double a = 0.0;
        for (int i = 0; i < n; ++i)
            a += fabs(L[i + j * n]);
        y[j] = a + 1.1;
    }
    /* X = I; Z = Diag(y) - L (ipm_mc_pk.c:83-89) */
    for (int i = 0; i < nn; ++i) {
        X[i] = 0.0;
        Z[i] = -L[i];
    }
    for (int i = 0; i < n; ++i) {
        X[i + i * n] = 1.0;
        Z[i + i * n] = y[i] - L[i + i * n];
    }
    phi = 0.0;
    for (int i = 0; i < n; ++i) phi += y[i];
    double psi = 0.0;
    for (int i = 0; i < nn; ++i) psi += L[i] * X[i];
    double mu = 0.0;
    for (int i = 0; i < nn; ++i) mu += Z[i] * X[i];
    mu /= (2.0 * n);

    double gap = fabs(phi - psi);
    while (gap > 1e-2) {                         /* ipm_mc_pk.c:126 */
        if (s.iters >= IPM_MAX_ITERS) { s.status = ORC_ERR_IPM_ITERS; break; }
        ++s.iters;

        /* Zi = inv(Z) via Cholesky (ipm_mc_pk.c:129-150) */
        memcpy(Zi, Z, sizeof(double) * nn);
        if (orc_chol_upper(Zi, n)) { s.status = ORC_ERR_CHOL_Z; break; }
        potri_upper(Zi, n, W);
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < j; ++k)
                Zi[j + n * k] = Zi[k + n * j];

        /* predictor (ipm_mc_pk.c:156-199) */
        for (int i = 0; i < nn; ++i) M[i] = Zi[i] * X[i];
        for (int i = 0; i < n; ++i) dy1[i] = -1.0;
        if (orc_chol_upper(M, n)) { s.status = ORC_ERR_SCHUR; break; }
        potrs_upper(M, n, dy1);
        for (int j = 0; j < nn; ++j) tmp[j] = -dy1[j % n] * X[j];
        symm(Zi, tmp, dX1, n);
        for (int i = 0; i < nn; ++i) dX1[i] = dX1[i] - X[i];
        symmetrize(dX1, n);

        /* corrector (ipm_mc_pk.c:207-272).  M is rebuilt and refactored by
         * the reference because dposv overwrote it; the factor is identical,
         * so the factor from the predictor is reused. */
        for (int i = 0; i < n; ++i) {
            double a = 0.0;
            for (int j = 0; j < n; ++j)
                a += Zi[i + j * n] * dX1[i + j * n] * dy1[j];
            dy[i] = mu * Zi[i + i * n] - a;
        }
        potrs_upper(M, n, dy);
        for (int j = 0; j < nn; ++j) {
            tmp[j] = -dy[j % n] * X[j];
            tmp2[j] = -dy1[j % n] * dX1[j];
        }
        for (int i = 0; i < nn; ++i) tmp[i] = tmp[i] + tmp2[i];
        symm(Zi, tmp, dX, n);
        for (int i = 0; i < nn; ++i) dX[i] = mu * Zi[i] + dX[i];
        for (int i = 0; i < n; ++i) dy[i] = dy1[i] + dy[i];
        for (int i = 0; i < nn; ++i) dX[i] = dX1[i] + dX[i];
        symmetrize(dX, n);

        /* primal line search (ipm_mc_pk.c:281-293) */
        double alpha_p = 1.0;
        int trials = 0, ok = 0;
        while (trials < LINESEARCH_MAX_TRIALS) {
            for (int i = 0; i < nn; ++i) tmp[i] = X[i] + alpha_p * dX[i];
            ++trials;
            if (orc_chol_upper(tmp, n) == 0) { ok = 1; break; }
            alpha_p *= 0.8;
        }
        s.chol_trials += trials;
        if (!ok) { s.status = ORC_ERR_LINESEARCH; break; }
        if (alpha_p < 1.0) alpha_p *= 0.95;

        /* dual line search, dZ = Diag(dy) (ipm_mc_pk.c:297-309) */
        double alpha_d = 1.0;
        trials = 0; ok = 0;
        while (trials < LINESEARCH_MAX_TRIALS) {
            memcpy(tmp, Z, sizeof(double) * nn);
            for (int i = 0; i < n; ++i) tmp[i + i * n] = Z[i + i * n] + alpha_d * dy[i];
            ++trials;
            if (orc_chol_upper(tmp, n) == 0) { ok = 1; break; }
            alpha_d *= 0.8;
        }
        s.chol_trials += trials;
        if (!ok) { s.status = ORC_ERR_LINESEARCH; break; }
        if (alpha_d < 1.0) alpha_d *= 0.95;

        if (alpha_trace && 2 * s.iters <= trace_cap) {
            alpha_trace[2 * (s.iters - 1)] = alpha_p;
            alpha_trace[2 * (s.iters - 1) + 1] = alpha_d;
        }

        /* update (ipm_mc_pk.c:313-334) */
        for (int i = 0; i < nn; ++i) X[i] = X[i] + alpha_p * dX[i];
        for (int i = 0; i < n; ++i) {
            y[i] = y[i] + alpha_d * dy[i];
            Z[i + i * n] = Z[i + i * n] + alpha_d * dy[i];
        }
        mu = 0.0;
        for (int i = 0; i < nn; ++i) mu += Z[i] * X[i];
        mu /= (2.0 * n);
        if (alpha_p + alpha_d > 1.6) mu *= 0.5;
        if (alpha_p + alpha_d > 1.9) mu *= 0.2;
        phi = 0.0;
        for (int i = 0; i < n; ++i) phi += y[i];
        psi = 0.0;
        for (int i = 0; i < nn; ++i) psi += L[i] * X[i];
        gap = fabs(phi - psi);
    }

done:
    if (y_out && y) memcpy(y_out, y, sizeof(double) * n);
    *phi_out = phi;
    if (st) *st = s;
    free(y); free(Z); free(dX); free(dy); free(dy1); free(dX1);
    free(tmp); free(tmp2); free(Zi); free(M); free(W);
}

void orc_ipm_mc_pk(const double *L, int n, double *X, double *phi, orc_ipm_stats *st)
{
    orc_ipm_mc_pk_ex(L, n, X, NULL, phi, st, NULL, 0);
}
