/*
 * bqb_runtime.c -- host side of the HIP bounding library, written in C.
 *
 * Device selection, the per-thread error string, and the host-pointer entry
 * points for the batched IPM (used by the drop-in ipm_mc_pk and by tests).
 * Device memory for a call is taken from a grow-only arena owned by the
 * library (the reference's allocMemory/freeMemory, src/allocate_free.c:8-76,
 * likewise sizes its scratch once per process).
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "bqb_kernels.h"
#include "bqb_runtime.h"

static _Thread_local char g_err[512];
static int g_device = -1;

void bqb_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

const char *bqb_last_error(void) { return g_err; }

int bqb_check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    bqb_set_error("%s: %s", what, hipGetErrorString(e));
    return -1;
}

int bqb_init(int device)
{
    int count = 0;
    if (bqb_check(hipGetDeviceCount(&count), "hipGetDeviceCount")) return -1;
    if (count < 1) { bqb_set_error("no HIP device"); return -1; }
    if (device < 0) device = 0;
    if (device >= count) { bqb_set_error("device %d out of range (%d devices)", device, count); return -1; }
    if (bqb_check(hipSetDevice(device), "hipSetDevice")) return -1;
    if (bqb_noalloc_setup()) return -1;
    g_device = device;
    return 0;
}

int bqb_ensure_init(void)
{
    if (g_device >= 0) return hipSetDevice(g_device) == hipSuccess ? 0 : -1;
    return bqb_init(0);
}

/* ---------------------------------------------------------------- arenas */
int bqb_buf_reserve(bqb_buf *b, size_t bytes)
{
    if (b->cap >= bytes && b->ptr) return 0;
    if (b->ptr) { hipFree(b->ptr); b->ptr = NULL; b->cap = 0; }
    if (bytes == 0) return 0;
    if (bqb_check(hipMalloc(&b->ptr, bytes), "hipMalloc")) return -1;
    b->cap = bytes;
    return 0;
}

void bqb_buf_free(bqb_buf *b)
{
    if (b->ptr) hipFree(b->ptr);
    b->ptr = NULL;
    b->cap = 0;
}

int bqb_hbuf_reserve(bqb_buf *b, size_t bytes)
{
    if (b->cap >= bytes && b->ptr) return 0;
    if (b->ptr) { hipHostFree(b->ptr); b->ptr = NULL; b->cap = 0; }
    if (bytes == 0) return 0;
    if (bqb_check(hipHostMalloc(&b->ptr, bytes, hipHostMallocDefault), "hipHostMalloc")) return -1;
    b->cap = bytes;
    return 0;
}

void bqb_hbuf_free(bqb_buf *b)
{
    if (b->ptr) hipHostFree(b->ptr);
    b->ptr = NULL;
    b->cap = 0;
}

/* ---------------------------------------------------------------- IPM */
static bqb_buf g_ipm_dev, g_ipm_host;

int bqb_ipm_ld_for(int nmax)
{
    int tile = bqb_ipm_tile_for(nmax);
    if (!tile) return 0;
    int P = tile / 100, R = tile % 100;
    return P * R;
}

static int ipm_batch_impl(int nodes, const int *n, const double *L, double *X, double *phi,
                          int *iters, int *trials, int *status, double *y_out, double *alpha_trace,
                          int trace_cap, int max_iters, double *work_out, int mix_slices, double *phi_cert,
                          double *ms_out);

int bqb_ipm_batch(int nodes, const int *n, const double *L, double *X, double *phi, int *iters,
                  int *trials, int *status, double *y_out, double *alpha_trace, int trace_cap)
{
    return ipm_batch_impl(nodes, n, L, X, phi, iters, trials, status, y_out, alpha_trace, trace_cap, 0, NULL, 0,
                          NULL, NULL);
}

int bqb_ipm_batch_mixed(int nodes, const int *n, const double *L, double *X, double *phi, double *phi_cert,
                        int *iters, int *trials, int *status, double *y_out, int slices, double *ms)
{
    if (slices < 0 || slices == 1 || slices > BQB_MIX_MAX_SLICES) {
        bqb_set_error("slices = %d: 0 (FP64 products) or 2..%d", slices, BQB_MIX_MAX_SLICES);
        return -1;
    }
    return ipm_batch_impl(nodes, n, L, X, phi, iters, trials, status, y_out, NULL, 0, 0, NULL, slices, phi_cert,
                          ms);
}

int bqb_ipm_debug(int n, const double *L, int max_iters, double *X, double *Zi, double *Mi,
                  double *dX, double *y, double *phi, int *iters, double *alpha_trace, double *dbg_out)
{
    int ld = bqb_ipm_ld_for(n);
    if (!ld) { bqb_set_error("n too large"); return -1; }
    double *w = malloc(sizeof(double) * (3 * (size_t)ld * ld + 4 * (size_t)ld));
    if (!w) { bqb_set_error("malloc"); return -1; }
    int st = 0;
    int rc = ipm_batch_impl(1, &n, L, X, phi, iters, NULL, &st, y, alpha_trace, max_iters, max_iters, w, 0, NULL,
                            NULL);
    if (rc == 0) {
        double *outs[3] = {Zi, Mi, dX};
        for (int m = 0; m < 3; ++m)
            if (outs[m])
                for (int j = 0; j < n; ++j)
                    memcpy(outs[m] + (size_t)j * n, w + (size_t)m * ld * ld + (size_t)j * ld, sizeof(double) * n);
        if (dbg_out) memcpy(dbg_out, w + 3 * (size_t)ld * ld, sizeof(double) * 4 * ld);
    }
    free(w);
    return rc;
}

static int ipm_batch_impl(int nodes, const int *n, const double *L, double *X, double *phi,
                          int *iters, int *trials, int *status, double *y_out, double *alpha_trace,
                          int trace_cap, int max_iters, double *work_out, int mix_slices, double *phi_cert,
                          double *ms_out)
{
    int wslabs = BQB_IPM_WORK_SLABS(mix_slices);
    if (nodes <= 0) return 0;
    if (bqb_ensure_init()) return -1;
    int nmax = 0;
    for (int b = 0; b < nodes; ++b) {
        if (n[b] < 1) { bqb_set_error("node %d: n = %d", b, n[b]); return -1; }
        if (n[b] > nmax) nmax = n[b];
    }
    int tile = bqb_ipm_tile_for(nmax);
    if (!tile) { bqb_set_error("n = %d exceeds the largest compiled IPM tile", nmax); return -1; }
    if (tile % 100 == 1 && wslabs < 4) wslabs = 4;      /* the blocked solve's four work slabs */
    const int ld = bqb_ipm_ld_for(nmax);
    const size_t nn = (size_t)ld * ld;
    if (trace_cap < 0) trace_cap = 0;

    /* device layout: n | L slabs | X slabs | work (wslabs/node) | phi | iters | trials | status | y | trace
     * | debug | certified phi */
    size_t off_n = 0;
    size_t off_L = (off_n + sizeof(int) * nodes + 255) & ~(size_t)255;
    size_t off_X = off_L + sizeof(double) * nn * nodes;
    size_t off_W = off_X + sizeof(double) * nn * nodes;
    size_t off_phi = off_W + sizeof(double) * wslabs * nn * nodes;
    size_t off_it = off_phi + sizeof(double) * nodes;
    size_t off_tr = off_it + sizeof(int) * nodes;
    size_t off_st = off_tr + sizeof(int) * nodes;
    size_t off_y = (off_st + sizeof(int) * nodes + 255) & ~(size_t)255;
    size_t off_a = off_y + sizeof(double) * ld * nodes;
    size_t off_dbg = off_a + sizeof(double) * 2 * (size_t)trace_cap * nodes;
    size_t off_cert = off_dbg + (work_out ? sizeof(double) * 4 * ld * nodes : 0);
    size_t off_m2 = off_cert + sizeof(double) * nodes;
    size_t total = off_m2 + sizeof(int) * nodes;
    if (bqb_buf_reserve(&g_ipm_dev, total)) return -1;
    size_t hbytes = off_W;                      /* n + L slabs + X slabs */
    if (bqb_hbuf_reserve(&g_ipm_host, hbytes)) return -1;

    char *h = (char *)g_ipm_host.ptr;
    char *d = (char *)g_ipm_dev.ptr;
    memcpy(h + off_n, n, sizeof(int) * nodes);
    double *hL = (double *)(h + off_L);
    memset(hL, 0, sizeof(double) * nn * nodes);
    size_t src = 0;
    for (int b = 0; b < nodes; ++b) {
        for (int j = 0; j < n[b]; ++j)
            memcpy(hL + b * nn + (size_t)j * ld, L + src + (size_t)j * n[b], sizeof(double) * n[b]);
        src += (size_t)n[b] * n[b];
    }
    if (bqb_check(hipMemcpy(d, h, off_X, hipMemcpyHostToDevice), "hipMemcpy H2D")) return -1;
    if (bqb_check(hipMemset(d + off_X, 0, off_phi - off_X), "hipMemset")) return -1;

    IpmBatchArgs a;
    memset(&a, 0, sizeof a);
    a.n = (const int *)(d + off_n);
    a.L = (const double *)(d + off_L);
    a.X = (double *)(d + off_X);
    a.work = (double *)(d + off_W);
    a.ld = ld;
    a.phi = (double *)(d + off_phi);
    a.iters = (int *)(d + off_it);
    a.trials = (int *)(d + off_tr);
    a.status = (int *)(d + off_st);
    a.active = NULL;
    a.y_out = (double *)(d + off_y);
    a.ystride = ld;
    a.alpha_trace = trace_cap ? (double *)(d + off_a) : NULL;
    a.trace_cap = trace_cap;
    a.max_iters = max_iters;
    a.dbg = work_out ? (double *)(d + off_dbg) : NULL;
    a.mix_slices = mix_slices;
    a.wslabs = wslabs;
    hipEvent_t e0 = NULL, e1 = NULL;
    if (ms_out && (bqb_check(hipEventCreate(&e0), "event") || bqb_check(hipEventCreate(&e1), "event"))) return -1;
    if (ms_out) hipEventRecord(e0, 0);
    if (bqb_check(bqb_launch_ipm_batch(&a, nodes, tile, 0), "ipm_batch_kernel launch")) return -1;
    /* mixed solves that failed are repeated with FP64 products */
    if (mix_slices &&
        bqb_check(bqb_launch_ipm_retry(&a, nodes, tile, (int *)(d + off_m2), NULL, 0, 0, 0), "retry launch"))
        return -1;
    /* blocked solves that failed are repeated with the register tile */
    if (!mix_slices && tile % 100 == 1 && !work_out &&
        bqb_check(bqb_launch_ipm_retry(&a, nodes, bqb_ipm_reg_tile_for_ld(tile / 100), (int *)(d + off_m2), NULL, 0,
                                       0, 0), "retry launch"))
        return -1;
    if (ms_out) hipEventRecord(e1, 0);
    if (phi_cert &&
        bqb_check(bqb_launch_cert_ipm(a.n, a.L, a.y_out, a.ystride, a.work, ld, wslabs, a.status,
                                      (double *)(d + off_cert), nodes, 0), "certificate launch"))
        return -1;
    if (bqb_check(hipDeviceSynchronize(), "ipm_batch_kernel")) return -1;
    if (ms_out) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        *ms_out = ms;
        hipEventDestroy(e0);
        hipEventDestroy(e1);
    }
    if (phi_cert &&
        bqb_check(hipMemcpy(phi_cert, d + off_cert, sizeof(double) * nodes, hipMemcpyDeviceToHost), "D2H cert"))
        return -1;

    double *hX = (double *)(h + off_L);        /* reuse the L staging area */
    if (bqb_check(hipMemcpy(hX, d + off_X, sizeof(double) * nn * nodes, hipMemcpyDeviceToHost), "hipMemcpy D2H")) return -1;
    size_t dst = 0;
    for (int b = 0; b < nodes; ++b) {
        if (X)
            for (int j = 0; j < n[b]; ++j)
                memcpy(X + dst + (size_t)j * n[b], hX + b * nn + (size_t)j * ld, sizeof(double) * n[b]);
        dst += (size_t)n[b] * n[b];
    }
    if (phi && bqb_check(hipMemcpy(phi, d + off_phi, sizeof(double) * nodes, hipMemcpyDeviceToHost), "D2H phi")) return -1;
    if (iters && bqb_check(hipMemcpy(iters, d + off_it, sizeof(int) * nodes, hipMemcpyDeviceToHost), "D2H iters")) return -1;
    if (trials && bqb_check(hipMemcpy(trials, d + off_tr, sizeof(int) * nodes, hipMemcpyDeviceToHost), "D2H trials")) return -1;
    if (status && bqb_check(hipMemcpy(status, d + off_st, sizeof(int) * nodes, hipMemcpyDeviceToHost), "D2H status")) return -1;
    if (y_out) {
        double *hy = malloc(sizeof(double) * ld * nodes);
        if (!hy) { bqb_set_error("malloc"); return -1; }
        if (bqb_check(hipMemcpy(hy, d + off_y, sizeof(double) * ld * nodes, hipMemcpyDeviceToHost), "D2H y")) { free(hy); return -1; }
        size_t o = 0;
        for (int b = 0; b < nodes; ++b) { memcpy(y_out + o, hy + (size_t)b * ld, sizeof(double) * n[b]); o += n[b]; }
        free(hy);
    }
    if (work_out &&
        bqb_check(hipMemcpy(work_out, d + off_W, sizeof(double) * 3 * nn * nodes, hipMemcpyDeviceToHost), "D2H work"))
        return -1;
    if (work_out &&
        bqb_check(hipMemcpy(work_out + 3 * nn * nodes, d + off_dbg, sizeof(double) * 4 * ld * nodes, hipMemcpyDeviceToHost), "D2H dbg"))
        return -1;
    if (alpha_trace && trace_cap &&
        bqb_check(hipMemcpy(alpha_trace, d + off_a, sizeof(double) * 2 * trace_cap * nodes, hipMemcpyDeviceToHost), "D2H trace"))
        return -1;
    return 0;
}

/* Drop-in for the reference's ipm_mc_pk (src/ipm_mc_pk.c:28, biqbin.h:176):
 * same arguments and meaning; failures abort with exit code 10 as the
 * reference's MPI_Abort(MPI_COMM_WORLD, 10) does. */
void ipm_mc_pk(double *L, int n, double *X, double *phi, int print)
{
    int it = 0, st = 0;
    (void)print;
    if (bqb_ipm_batch(1, &n, L, X, phi, &it, NULL, &st, NULL, NULL, 0) || st) {
        fprintf(stderr, "ipm_mc_pk: %s (status %d)\n", st ? "IPM failed" : bqb_last_error(), st);
        exit(10);
    }
}

/* ---------------------------------------------------------------- self-test */
hipError_t bqb_launch_dist_selftest(int tile, int op, int n, int ld, const double *in, double *out,
                                    int *flag);

int bqb_dist_selftest(int tile, int op, int n, const double *in, double *out, int *ok)
{
    if (bqb_ensure_init()) return -1;
    const int P = tile / 100, R = tile % 100;
    const int ld = P * R;
    if (n < 1 || n > ld) { bqb_set_error("n out of range for tile"); return -1; }
    size_t bytes = sizeof(double) * (size_t)ld * ld;
    double *h = calloc((size_t)ld * ld, sizeof(double));
    void *d = NULL;
    if (!h) { bqb_set_error("calloc"); return -1; }
    for (int j = 0; j < n; ++j) memcpy(h + (size_t)j * ld, in + (size_t)j * n, sizeof(double) * n);
    int rc = -1;
    if (bqb_check(hipMalloc(&d, 2 * bytes + 256), "hipMalloc")) goto out;
    if (bqb_check(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice), "H2D")) goto out;
    if (bqb_check(hipMemset((char *)d + bytes, 0, bytes + 256), "memset")) goto out;
    if (bqb_check(bqb_launch_dist_selftest(tile, op, n, ld, (const double *)d, (double *)((char *)d + bytes),
                                           (int *)((char *)d + 2 * bytes)), "selftest launch")) goto out;
    if (bqb_check(hipDeviceSynchronize(), "selftest")) goto out;
    if (bqb_check(hipMemcpy(h, (char *)d + bytes, bytes, hipMemcpyDeviceToHost), "D2H")) goto out;
    if (bqb_check(hipMemcpy(ok, (char *)d + 2 * bytes, sizeof(int), hipMemcpyDeviceToHost), "D2H")) goto out;
    for (int j = 0; j < n; ++j) memcpy(out + (size_t)j * n, h + (size_t)j * ld, sizeof(double) * n);
    rc = 0;
out:
    if (d) hipFree(d);
    free(h);
    return rc;
}

hipError_t bqb_launch_mx_selftest(int n, int s, const double *A, const double *B, double *C, unsigned short *S);

/* Diagnostic: C = A B through the sliced bf16 matrix-core product. */
int bqb_mx_selftest(int n, int slices, const double *A, const double *B, double *C)
{
    if (bqb_ensure_init()) return -1;
    if (n < 1 || n > 256 || slices < 1 || slices > BQB_MIX_MAX_SLICES) { bqb_set_error("mx selftest: n or slices out of range"); return -1; }
    const size_t nn = sizeof(double) * (size_t)n * n;
    const size_t sl = 2 * (size_t)2 * slices * 256 * 256;
    const size_t soff = (3 * nn + 255) & ~(size_t)255;   /* the slices take 16-byte loads */
    char *d = NULL;
    int rc = -1;
    if (bqb_check(hipMalloc((void **)&d, soff + sl), "hipMalloc")) return -1;
    if (bqb_check(hipMemcpy(d, A, nn, hipMemcpyHostToDevice), "H2D") ||
        bqb_check(hipMemcpy(d + nn, B, nn, hipMemcpyHostToDevice), "H2D") ||
        bqb_check(hipMemset(d + 2 * nn, 0, nn), "memset"))
        goto out;
    if (bqb_check(bqb_launch_mx_selftest(n, slices, (const double *)d, (const double *)(d + nn), (double *)(d + 2 * nn),
                                         (unsigned short *)(d + soff)), "mx selftest launch"))
        goto out;
    if (bqb_check(hipDeviceSynchronize(), "mx selftest")) goto out;
    if (bqb_check(hipMemcpy(C, d + 2 * nn, nn, hipMemcpyDeviceToHost), "D2H")) goto out;
    rc = 0;
out:
    hipFree(d);
    return rc;
}

hipError_t bqb_launch_copy_probe(const void *src, void *dst, size_t bytes, int reps);

/* Copies `bytes` (a multiple of 16) between two device buffers `reps` times
 * with a plain streaming kernel; *ms receives the average time per copy.
 * Calibration of the HBM byte counters (each copy moves 2 x bytes). */
int bqb_copy_probe(size_t bytes, int reps, double *ms)
{
    if (bqb_ensure_init()) return -1;
    if (bytes == 0 || bytes % 16 || reps < 1) { bqb_set_error("copy probe: bytes must be a positive multiple of 16"); return -1; }
    void *a = NULL, *b = NULL;
    hipEvent_t e0 = NULL, e1 = NULL;
    int rc = -1;
    if (bqb_check(hipMalloc(&a, bytes), "hipMalloc") || bqb_check(hipMalloc(&b, bytes), "hipMalloc")) goto out;
    if (bqb_check(hipMemset(a, 1, bytes), "hipMemset")) goto out;
    if (bqb_check(hipEventCreate(&e0), "event") || bqb_check(hipEventCreate(&e1), "event")) goto out;
    if (bqb_check(bqb_launch_copy_probe(a, b, bytes, 1), "copy probe")) goto out;   /* warm-up */
    hipEventRecord(e0, 0);
    if (bqb_check(bqb_launch_copy_probe(a, b, bytes, reps), "copy probe")) goto out;
    hipEventRecord(e1, 0);
    if (bqb_check(hipEventSynchronize(e1), "sync")) goto out;
    float t = 0.f;
    hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t / reps;
    rc = 0;
out:
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (a) hipFree(a);
    if (b) hipFree(b);
    return rc;
}

/* ---- IPM tile autotuner ------------------------------------------------ */
static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* Median launch time of the batch with the given tile forced and the
 * products in FP64 (slices 0) or on the bf16 matrix cores. */
static int time_tile(int tile, int slices, int nodes, const int *nv, const double *L, double *X, double *phi, int *it,
                     int *tr, int *st, int reps, double *med)
{
    double ms[16];
    if (reps > 16) reps = 16;
    bqb_set_ipm_tile(tile);
    for (int r = -1; r < reps; ++r) {
        double t = 0.0;
        if (ipm_batch_impl(nodes, nv, L, X, phi, it, tr, st, NULL, NULL, 0, 0, NULL, slices, NULL, &t)) {
            bqb_set_ipm_tile(0);
            return -1;
        }
        if (r >= 0) ms[r] = t;
    }
    bqb_set_ipm_tile(0);
    qsort(ms, (size_t)reps, sizeof(double), cmp_double);
    *med = ms[reps / 2];
    return 0;
}

/* The subproblems an autotuner times for class ld: `nodes` Max-Cut L of
 * G(ld, 0.3) with weights +-1 (the class's largest size). */
static int tune_batch(int ld, int nodes, double **L, int **nv)
{
    const int n = ld;
    double *adj = calloc((size_t)n * n, sizeof(double));
    *L = malloc(sizeof(double) * (size_t)n * n * nodes);
    *nv = malloc(sizeof(int) * nodes);
    if (!adj || !*L || !*nv) { free(adj); bqb_set_error("malloc"); return -1; }
    uint64_t z = 0x9E3779B97F4A7C15ull * (uint64_t)(ld + 1);
    int rc = 0;
    for (int b = 0; b < nodes && !rc; ++b) {
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                z = z * 6364136223846793005ull + 1442695040888963407ull;
                const double w = ((z >> 33) % 10 < 3) ? (((z >> 20) & 1) ? 1.0 : -1.0) : 0.0;
                adj[i + (size_t)j * n] = adj[j + (size_t)i * n] = w;
            }
        if (bqb_build_L(n, adj, *L + (size_t)b * n * n)) rc = -1;
        (*nv)[b] = n;
    }
    free(adj);
    return rc;
}

int bqb_autotune_ipm(int nmin, int nmax, int nodes, int reps, double *ms_out, int *tiles)
{
    if (nodes < 1 || reps < 1 || nmin < 1 || nmax < nmin) { bqb_set_error("autotune: bad arguments"); return -1; }
    if (bqb_ensure_init()) return -1;
    int table[BQB_TILE_CLASSES];
    bqb_ipm_tile_table(table);
    int rc = 0;
    for (int c = 0; c < BQB_TILE_CLASSES && !rc; ++c) {
        const int ld = bqb_ipm_class_ld(c), lo = c ? bqb_ipm_class_ld(c - 1) + 1 : 1;
        if (ld < nmin || lo > nmax) continue;
        double *L = NULL, *X = malloc(sizeof(double) * (size_t)ld * ld * nodes), *phi = malloc(sizeof(double) * nodes);
        int *nv = NULL, *it = malloc(sizeof(int) * nodes), *tr = malloc(sizeof(int) * nodes),
            *st = malloc(sizeof(int) * nodes);
        if (!X || !phi || !it || !tr || !st) { bqb_set_error("malloc"); rc = -1; }
        if (!rc) rc = tune_batch(ld, nodes, &L, &nv);
        double mr = 0.0, mb = 0.0;
        const int reg = bqb_ipm_reg_tile_for_ld(ld), blk = ld * 100 + 1;
        if (!rc) rc = time_tile(reg, 0, nodes, nv, L, X, phi, it, tr, st, reps, &mr);
        if (!rc) rc = time_tile(blk, 0, nodes, nv, L, X, phi, it, tr, st, reps, &mb);
        if (!rc) {
            const int cur = table[c], other = cur == reg ? blk : reg;
            const double tcur = cur == reg ? mr : mb, toth = cur == reg ? mb : mr;
            if (toth < 0.95 * tcur) table[c] = other;
            if (ms_out) { ms_out[2 * c] = mr; ms_out[2 * c + 1] = mb; }
        }
        free(L); free(X); free(phi); free(nv); free(it); free(tr); free(st);
    }
    if (!rc && bqb_set_ipm_tile_table(table)) { bqb_set_error("autotune: table rejected"); rc = -1; }
    if (tiles) bqb_ipm_tile_table(tiles);
    return rc;
}

int bqb_autotune_ipm_auto(int nmin, int nmax, int nodes, int reps, double *ms_out, int *tiles, int *slices)
{
    if (nodes < 1 || reps < 1 || nmin < 1 || nmax < nmin) { bqb_set_error("autotune: bad arguments"); return -1; }
    if (bqb_ensure_init()) return -1;
    int ttab[BQB_TILE_CLASSES], stab[BQB_TILE_CLASSES];
    bqb_ipm_auto_table(ttab, stab);
    int rc = 0;
    for (int c = 0; c < BQB_TILE_CLASSES && !rc; ++c) {
        const int ld = bqb_ipm_class_ld(c), lo = c ? bqb_ipm_class_ld(c - 1) + 1 : 1;
        if (ld < nmin || lo > nmax) continue;
        double *L = NULL, *X = malloc(sizeof(double) * (size_t)ld * ld * nodes), *phi = malloc(sizeof(double) * nodes);
        int *nv = NULL, *it = malloc(sizeof(int) * nodes), *tr = malloc(sizeof(int) * nodes),
            *st = malloc(sizeof(int) * nodes);
        if (!X || !phi || !it || !tr || !st) { bqb_set_error("malloc"); rc = -1; }
        if (!rc) rc = tune_batch(ld, nodes, &L, &nv);
        /* variants: register / blocked, FP64 / 4 bf16 slices */
        const int vt[4] = {bqb_ipm_reg_tile_for_ld(ld), ld * 100 + 1, bqb_ipm_reg_tile_for_ld(ld), ld * 100 + 1};
        const int vs[4] = {0, 0, 4, 4};
        double ms[4] = {0, 0, 0, 0};
        for (int v = 0; v < 4 && !rc; ++v) rc = time_tile(vt[v], vs[v], nodes, nv, L, X, phi, it, tr, st, reps, &ms[v]);
        if (!rc) {
            int cur = 0;
            for (int v = 0; v < 4; ++v)
                if (vt[v] == ttab[c] && vs[v] == stab[c]) cur = v;
            int best = cur;
            for (int v = 0; v < 4; ++v)
                if (ms[v] < 0.95 * ms[cur] && ms[v] < ms[best]) best = v;
            ttab[c] = vt[best];
            stab[c] = vs[best];
            if (ms_out)
                for (int v = 0; v < 4; ++v) ms_out[4 * c + v] = ms[v];
        }
        free(L); free(X); free(phi); free(nv); free(it); free(tr); free(st);
    }
    if (!rc && bqb_set_ipm_auto_table(ttab, stab)) { bqb_set_error("autotune: table rejected"); rc = -1; }
    bqb_ipm_auto_table(tiles, slices);
    return rc;
}
