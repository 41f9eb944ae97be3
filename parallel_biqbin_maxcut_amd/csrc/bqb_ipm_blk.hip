// bqb_ipm_blk.hip -- blocked variant of the batched FP64 primal-dual
// interior-point solve of the basic Max-Cut SDP (reference
// src/ipm_mc_pk.c:28-359) for large subproblems.  Tile codes ld*100 + 1
// select it (ld the slab leading dimension).
//
// The register-tile kernel (bqb_ipm_kernel.h) holds one n x n tile in
// registers; above n ~ 160 that tile no longer fits the registers of a
// 1024-thread workgroup and most of it lives in scratch memory.  Here every
// matrix stays in the node's slab (its working set of 4 slabs stays in L2 /
// MALL between phases) and each phase streams it through 8 x 8 register
// micro-tiles of a 16 x 16 thread grid (one 128 x 128 tile per pass):
//   * SPD inverse by the symmetric sweep in panels of BK pivots: the panel's
//     BK columns are swept in LDS, one barrier per pivot (double-buffered
//     panel), then the rest of the matrix takes one rank-BK update
//         A_RR <- A_RR - (A_RK G) A_RK',   G = inv(A_KK),
//     on its lower triangle (the block form of the same sweep operator);
//   * Cholesky tests: the same panel sweep restricted to the trailing
//     matrix: its pivots are the Cholesky pivots, so a test fails exactly
//     where dpotrf would;
//   * the two products: 128 x 128 output tiles, LDS stages of 16 columns
//     loaded one stage ahead through registers, accumulated in k order with
//     the register kernel's operand formulas.
// The iteration, its formulas, line searches and stopping rule are those of
// the register-tile kernel; only the rounding of the sweeps' updates and of
// the reductions differs.
#include <hip/hip_runtime.h>

#include "bqb_dist.h"
#include "bqb_kernels.h"
#include "bqb_mixed.h"

namespace bqb {
namespace {

constexpr int NT = 256;       // threads: 16 x 16 grid
constexpr int BK = 16;        // panel width (pivots per panel)
constexpr int TM = 128;       // output tile (rows and columns)
constexpr int MT = 8;         // micro tile per thread: rows tr + 16a, columns tc + 16b
constexpr int NB = 256;       // largest n
constexpr int NWV = NT / 64;  // waves
constexpr int MAX_ITERS = 500;
constexpr int LS_TRIALS = 500;

struct BlkLds {
    // the swept panel (column kk, row i) and the panel before its sweep,
    // read by the rank update; together also the product stages
    double pan[BK][NB];
    double old[BK][NB];
    double prow[2][BK];       // pivot row, double buffered
    double y[NB], zd[NB], dy1[NB], dy2[NB], v[NB], dg[NB];
    double red[NWV];
};
static_assert(NT == NB, "one panel row per thread");

struct Tid {
    int tid, tr, tc, lane, wave;
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the workgroup in a fixed order, identical in every thread.
__device__ __forceinline__ double block_sum(double v, double *red, const Tid &t)
{
    v = wave_sum(v);
    __syncthreads();
    if (t.lane == 0) red[t.wave] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NWV; ++w) s += red[w];
    return s;
}

// out[j] = sum_i term(i, j) for j < n: each wave takes CW columns at a
// time, lanes over the rows (n <= 256: four per lane), fixed order.
template <class F>
__device__ __forceinline__ void col_sums(int n, double *out, const Tid &t, F term)
{
    constexpr int CW = 8;
    for (int j0 = t.wave * CW; j0 < n; j0 += NWV * CW) {
        double s[CW];
#pragma unroll
        for (int c = 0; c < CW; ++c) s[c] = 0.0;
#pragma unroll
        for (int m = 0; m < NB / 64; ++m) {
            const int i = t.lane + 64 * m;
#pragma unroll
            for (int c = 0; c < CW; ++c)
                if (i < n && j0 + c < n) s[c] += term(i, j0 + c);
        }
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            const double v = wave_sum(s[c]);
            if (t.lane == 0 && j0 + c < n) out[j0 + c] = v;
        }
    }
    __syncthreads();
}

// Runs an elementwise pass over the n x n elements, EU per thread per step:
// load(i, j, v) fills v[0..NL) for element (i, j), then store(i, j, v)
// writes; all loads of a step are issued before its first store.
constexpr int EU = 8;
template <int NL, class LD, class ST>
__device__ __forceinline__ void ew_pass(int n, const Tid &t, LD load, ST store)
{
    const int nsq = n * n;
    for (int base = 0; base < nsq; base += NT * EU) {
        double v[EU][NL];
        int ii[EU], jj[EU];
#pragma unroll
        for (int u = 0; u < EU; ++u) {
            const int e = base + t.tid + NT * u;
            const int j = e / n, i = e - j * n;
            ii[u] = i;
            jj[u] = j;
            if (e < nsq) load(i, j, v[u]);
        }
#pragma unroll
        for (int u = 0; u < EU; ++u)
            if (base + t.tid + NT * u < nsq) store(ii[u], jj[u], v[u]);
    }
}

// ---------------------------------------------------------------- sweeps
// The panel is columns k0 .. k0+BK-1; thread i holds row i of it in
// registers.  panel_load reads the symmetric matrix's lower triangle
// (element (i, k) at (max, min)) for rows lo <= i < n (zero elsewhere) and
// saves the rows as they were in S.old.
__device__ __forceinline__ void panel_load(const double *A, int ld, int n, int lo, int k0, double (&p)[BK],
                                           BlkLds &S, const Tid &t)
{
    const int i = t.tid;
    const bool act = i >= lo && i < n;
#pragma unroll
    for (int kk = 0; kk < BK; ++kk) {
        const int k = k0 + kk;
        p[kk] = (act && k < n) ? A[(i >= k) ? i + (size_t)k * ld : k + (size_t)i * ld] : 0.0;
    }
#pragma unroll
    for (int kk = 0; kk < BK; ++kk) S.old[kk][i] = p[kk];
}

// Sweeps the panel's pivots k0 .. min(k0+BK, n)-1 over its columns, rows
// lo .. n-1: pivot k maps a_ij -> a_ij - a_ik (a_kj / a_kk),
// a_ik -> a_ik / a_kk, a_kj -> a_kj / a_kk, a_kk -> -1 / a_kk.  The owner of
// row k publishes it in LDS (one barrier per pivot).  Returns false at the
// first pivot that is not positive (uniform: every thread reads the same
// LDS value); on success the swept rows are also in S.pan.
__device__ __forceinline__ bool panel_sweep(double (&p)[BK], BlkLds &S, int n, int lo, int k0, const Tid &t)
{
    const int i = t.tid;
    const bool act = i >= lo && i < n;
    bool ok = true;
    static_for<0, BK>([&](auto kc) {
        constexpr int kk = decltype(kc)::value;
        const int k = k0 + kk;
        if (!ok || k >= n) return;
        double *prow = S.prow[kk & 1];
        if (i == k) {
#pragma unroll
            for (int jj = 0; jj < BK; ++jj) prow[jj] = p[jj];
        }
        __syncthreads();
        const double d = prow[kk];
        if (!(d > 0.0)) {
            ok = false;
            return;
        }
        const double dinv = pivot_recip(d);
        double tj[BK];
#pragma unroll
        for (int jj = 0; jj < BK; ++jj) tj[jj] = prow[jj] * dinv;
        if (act) {
            if (i == k) {
#pragma unroll
                for (int jj = 0; jj < BK; ++jj) p[jj] = (jj == kk) ? -dinv : tj[jj];
            } else {
                const double c = p[kk];
#pragma unroll
                for (int jj = 0; jj < BK; ++jj) p[jj] = (jj == kk) ? c * dinv : fma(-c, tj[jj], p[jj]);
            }
        }
    });
    if (!ok) return false;
#pragma unroll
    for (int kk = 0; kk < BK; ++kk) S.pan[kk][i] = p[kk];
    __syncthreads();
    return true;
}

// A(I, J) -= sum_kk U[kk][I] V[kk][J] on the lower triangle, for I, J in
// [lo, n) outside the panel [k0, k0 + BK).
__device__ __forceinline__ void rank_update(double *A, int ld, int n, int lo, int k0, const double (*U)[NB],
                                            const double (*V)[NB], const Tid &t)
{
    const int t0 = lo / TM, nt = (n + TM - 1) / TM;
    for (int ti = t0; ti < nt; ++ti)
        for (int tj = t0; tj <= ti; ++tj) {
            const int i0 = ti * TM + t.tr, j0 = tj * TM + t.tc;
            bool rv[MT], cv[MT];
#pragma unroll
            for (int a = 0; a < MT; ++a) {
                const int I = i0 + 16 * a;
                rv[a] = I >= lo && I < n && (I < k0 || I >= k0 + BK);
            }
#pragma unroll
            for (int b = 0; b < MT; ++b) {
                const int J = j0 + 16 * b;
                cv[b] = J >= lo && J < n && (J < k0 || J >= k0 + BK);
            }
            double acc[MT][MT];
#pragma unroll
            for (int b = 0; b < MT; ++b)
#pragma unroll
                for (int a = 0; a < MT; ++a) {
                    const int I = i0 + 16 * a, J = j0 + 16 * b;
                    acc[a][b] = (rv[a] && cv[b] && I >= J) ? A[I + (size_t)J * ld] : 0.0;
                }
#pragma unroll 4
            for (int kk = 0; kk < BK; ++kk) {
                double u[MT], w[MT];
#pragma unroll
                for (int a = 0; a < MT; ++a) u[a] = U[kk][i0 + 16 * a];
#pragma unroll
                for (int b = 0; b < MT; ++b) w[b] = V[kk][j0 + 16 * b];
#pragma unroll
                for (int a = 0; a < MT; ++a)
#pragma unroll
                    for (int b = 0; b < MT; ++b) acc[a][b] = fma(-u[a], w[b], acc[a][b]);
            }
#pragma unroll
            for (int b = 0; b < MT; ++b)
#pragma unroll
                for (int a = 0; a < MT; ++a) {
                    const int I = i0 + 16 * a, J = j0 + 16 * b;
                    if (rv[a] && cv[b] && I >= J) A[I + (size_t)J * ld] = acc[a][b];
                }
        }
}

// In-place inverse of the SPD matrix A (n x n, full storage in, full
// storage out) by the panel sweep; false (uniformly) at the first pivot
// that is not positive.
__device__ bool blk_inverse(double *A, int ld, int n, BlkLds &S, const Tid &t)
{
    const int i = t.tid;
    for (int k0 = 0; k0 < n; k0 += BK) {
        double p[BK];
        __syncthreads();
        panel_load(A, ld, n, 0, k0, p, S, t);
        if (!panel_sweep(p, S, n, 0, k0, t)) return false;
        // the panel's columns as swept; rows above the panel go to row k
        // of the lower storage; inside the panel only (i >= k) is written
        if (i < n) {
#pragma unroll
            for (int kk = 0; kk < BK; ++kk) {
                const int k = k0 + kk;
                if (k >= n) break;
                if (i >= k) A[i + (size_t)k * ld] = p[kk];
                else if (i < k0) A[k + (size_t)i * ld] = p[kk];
            }
        }
        rank_update(A, ld, n, 0, k0, S.pan, S.old, t);
    }
    __syncthreads();
    // the sweep leaves -inv(A) in the lower triangle: negate and mirror
    ew_pass<1>(
        n, t, [&](int i, int j, double *v) { v[0] = (i >= j) ? A[i + (size_t)j * ld] : 0.0; },
        [&](int i, int j, const double *v) {
            if (i < j) return;
            A[i + (size_t)j * ld] = -v[0];
            A[j + (size_t)i * ld] = -v[0];
        });
    __syncthreads();
    return true;
}

// Positive-definiteness test of A (lower triangle; destroyed): the panel
// sweep of the trailing matrix, i.e. right-looking blocked Cholesky.
__device__ bool blk_chol_test(double *A, int ld, int n, BlkLds &S, const Tid &t)
{
    for (int k0 = 0; k0 < n; k0 += BK) {
        double p[BK];
        __syncthreads();
        panel_load(A, ld, n, k0, k0, p, S, t);
        if (!panel_sweep(p, S, n, k0, k0, t)) return false;
        if (k0 + BK < n) rank_update(A, ld, n, k0 + BK, k0, S.pan, S.old, t);
    }
    __syncthreads();
    return true;
}

// ---------------------------------------------------------------- products
// C = Zi * T with T(k, J) = -s1[k] X(J, k)  (TWO: (-s2[k] X(J, k)) +
// (-s1[k] W(J, k))), all n x n; every operand is symmetric, so row k of T
// is formed from column k.  Summed in k order.
template <bool TWO>
__device__ void gemm_blk(double *C, const double *Zi, const double *X, const double *W, const double *s1,
                         const double *s2, int ld, int n, BlkLds &S, const Tid &t)
{
    double *As = &S.pan[0][0];             // [2][BK][TM]: S.pan and S.old, contiguous
    double *Bs = As + 2 * BK * TM;         // [2][BK][TM]
    constexpr int PER = BK * TM / NT;      // 8 elements of each stage per thread
    const int nt = (n + TM - 1) / TM;
    for (int ti = 0; ti < nt; ++ti)
        for (int tj = 0; tj < nt; ++tj) {
            const int i0 = ti * TM, j0 = tj * TM;
            double ra[PER], rb[PER];
            auto fetch = [&](int k0) {
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    const int e = t.tid + NT * q;
                    const int kk = e / TM, r = e - kk * TM;
                    const int k = k0 + kk;
                    const int I = i0 + r, J = j0 + r;
                    ra[q] = (k < n && I < n) ? Zi[I + (size_t)k * ld] : 0.0;
                    double b = 0.0;
                    if (k < n && J < n) {
                        const double m1 = -s1[k];
                        if (TWO) {
                            const double m2 = -s2[k];
                            b = (m2 * X[J + (size_t)k * ld]) + (m1 * W[J + (size_t)k * ld]);
                        } else {
                            b = m1 * X[J + (size_t)k * ld];
                        }
                    }
                    rb[q] = b;
                }
            };
            auto put = [&](int buf) {
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    const int e = t.tid + NT * q;
                    As[buf * BK * TM + e] = ra[q];
                    Bs[buf * BK * TM + e] = rb[q];
                }
            };
            double acc[MT][MT];
#pragma unroll
            for (int a = 0; a < MT; ++a)
#pragma unroll
                for (int b = 0; b < MT; ++b) acc[a][b] = 0.0;
            __syncthreads();            // the stages may still be read
            fetch(0);
            put(0);
            __syncthreads();
            int buf = 0;
            for (int k0 = 0; k0 < n; k0 += BK) {
                const bool more = k0 + BK < n;
                if (more) fetch(k0 + BK);
                const double *Ap = As + buf * BK * TM, *Bp = Bs + buf * BK * TM;
#pragma unroll 4
                for (int kk = 0; kk < BK; ++kk) {
                    double aI[MT], tJ[MT];
#pragma unroll
                    for (int a = 0; a < MT; ++a) aI[a] = Ap[kk * TM + t.tr + 16 * a];
#pragma unroll
                    for (int b = 0; b < MT; ++b) tJ[b] = Bp[kk * TM + t.tc + 16 * b];
#pragma unroll
                    for (int a = 0; a < MT; ++a)
#pragma unroll
                        for (int b = 0; b < MT; ++b) acc[a][b] = fma(aI[a], tJ[b], acc[a][b]);
                }
                if (more) put(buf ^ 1);
                __syncthreads();
                buf ^= 1;
            }
#pragma unroll
            for (int b = 0; b < MT; ++b)
#pragma unroll
                for (int a = 0; a < MT; ++a) {
                    const int I = i0 + t.tr + 16 * a, J = j0 + t.tc + 16 * b;
                    if (I < n && J < n) C[I + (size_t)J * ld] = acc[a][b];
                }
        }
    __syncthreads();
}

// ---------------------------------------------------------------- the solve
// MIX: the two products on the bf16 matrix cores (bqb_mixed.h, slices in
// work slabs 4 and 5); everything else in FP64.
template <bool MIX>
__device__ void blk_solve(const IpmBatchArgs &args, const int node, const int n, BlkLds &S, const Tid &t)
{
    const int ld = args.ld;
    const size_t nn = (size_t)ld * ld;
    const double *L = args.L + (size_t)node * nn;
    double *X = args.X + (size_t)node * nn;
    double *W0 = args.work + (size_t)node * args.wslabs * nn;
    double *Wzi = W0;             // inv(Z)
    double *Wmi = W0 + nn;        // inv(Zi o X)
    double *Wdx = W0 + 2 * nn;    // dX1, then dX
    double *Ws = W0 + 3 * nn;     // products, line-search trial matrices
    unsigned short *Wsl = (unsigned short *)(W0 + 4 * nn);   // bf16 slices of the product operands
    const int mixs = MIX ? args.mix_slices : 0;
    const bool dbg = args.dbg != nullptr;

    // ---- initial point (ipm_mc_pk.c:65-99) -------------------------------
    col_sums(n, S.y, t, [&](int i, int j) { return fabs(L[i + (size_t)j * ld]); });
    for (int i = t.tid; i < NB; i += NT) {
        const double dl = (i < n) ? L[i + (size_t)i * ld] : 0.0;
        S.dg[i] = dl;
        S.y[i] = (i < n) ? S.y[i] + 1.1 : 0.0;
        S.zd[i] = (i < n) ? S.y[i] - dl : 0.0;
    }
    ew_pass<1>(
        n, t, [&](int, int, double *) {},
        [&](int i, int j, const double *) { X[i + (size_t)j * ld] = (i == j) ? 1.0 : 0.0; });
    __syncthreads();
    double phi, psi, mu;
    {
        double py = 0.0, pl = 0.0, pz = 0.0;
        for (int i = t.tid; i < n; i += NT) { py += S.y[i]; pl += S.dg[i]; pz += S.zd[i]; }
        phi = block_sum(py, S.red, t);
        psi = block_sum(pl, S.red, t);
        mu = block_sum(pz, S.red, t) / (2.0 * n);
    }
    if (dbg) {
        double *o = args.dbg + (size_t)node * 4 * ld;
        for (int i = t.tid; i < ld; i += NT) {
            o[i] = (i < NB) ? S.y[i] : 0.0;
            o[ld + i] = (i < NB) ? S.zd[i] : 0.0;
            o[2 * ld + i] = (i < NB) ? S.dg[i] : 0.0;
        }
        if (t.tid == 0) { o[3 * ld] = phi; o[3 * ld + 1] = psi; o[3 * ld + 2] = mu; }
    }
    double gap = fabs(phi - psi);
    int iter = 0, trials = 0, status = 0;
    const int max_iters = args.max_iters > 0 ? args.max_iters : MAX_ITERS;
    bool zi_ready = false;
    // mixed products while the iterate is far from optimal; FP64 products
    // after the first iteration whose relative gap is below 1e-3, or, from
    // the fourth on, whose primal or dual step was cut below 0.3
    bool mx_on = MIX;
    int iters_mx = 0;
    while (gap > 1e-2) {
        if (iter >= max_iters) {
            if (args.max_iters <= 0) status = ERR_IPM_ITERS;
            break;
        }
        ++iter;
        const bool use_mx = MIX && mx_on;
        iters_mx += use_mx ? 1 : 0;

        // ---- Zi = inv(Z), Z = Diag(zd) - offdiag(L) -------------------------
        if (!zi_ready) {
            ew_pass<1>(
                n, t, [&](int i, int j, double *v) { v[0] = L[i + (size_t)j * ld]; },
                [&](int i, int j, const double *v) { Wzi[i + (size_t)j * ld] = (i == j) ? S.zd[i] : -v[0]; });
            if (!blk_inverse(Wzi, ld, n, S, t)) { status = ERR_CHOL_Z; break; }
        }
        zi_ready = false;
        for (int i = t.tid; i < n; i += NT) S.dg[i] = Wzi[i + (size_t)i * ld];

        // ---- Mi = inv(Zi o X), dy1 = Mi (-e) ------------------------------
        ew_pass<2>(
            n, t,
            [&](int i, int j, double *v) {
                const size_t o = i + (size_t)j * ld;
                v[0] = Wzi[o];
                v[1] = X[o];
            },
            [&](int i, int j, const double *v) { Wmi[i + (size_t)j * ld] = v[0] * v[1]; });
        if (!blk_inverse(Wmi, ld, n, S, t)) { status = ERR_SCHUR; break; }
        col_sums(n, S.v, t, [&](int i, int j) { return Wmi[i + (size_t)j * ld]; });
        for (int i = t.tid; i < NB; i += NT) S.dy1[i] = (i < n) ? -S.v[i] : 0.0;
        __syncthreads();

        // ---- dX1 = sym(Zi * (-Diag(dy1) X) - X) ---------------------------
        if (MIX && use_mx) {
            // row i of Zi is its column i; column j of -Diag(dy1) X is
            // -dy1 o X(:, j) (X symmetric)
            const double *zi = Wzi;
            const auto zi_line = [&](int i, int k) { return zi[k + (size_t)i * ld]; };
            const auto t1_line = [&](int j, int k) { return -S.dy1[k] * X[k + (size_t)j * ld]; };
            mx_product<NT>(zi_line, t1_line, n, mixs, Wsl, (int *)&S.pan[0][0], Ws, ld);
        } else {
            gemm_blk<false>(Ws, Wzi, X, nullptr, S.dy1, nullptr, ld, n, S, t);
        }
        ew_pass<4>(
            n, t,
            [&](int i, int j, double *v) {
                if (i < j) return;
                const size_t o = i + (size_t)j * ld, p = j + (size_t)i * ld;
                v[0] = Ws[o];
                v[1] = X[o];
                v[2] = Ws[p];
                v[3] = X[p];
            },
            [&](int i, int j, const double *v) {
                if (i < j) return;
                const double a = v[0] - v[1], b = v[2] - v[3];
                const double w = (i == j) ? a : 0.5 * (a + b);
                Wdx[i + (size_t)j * ld] = w;
                Wdx[j + (size_t)i * ld] = w;
            });
        __syncthreads();

        // ---- corrector rhs: mu diag(Zi) - (Zi o dX1) dy1; dy2 = Mi rhs -----
        col_sums(n, S.v, t, [&](int i, int j) {
            const size_t o = i + (size_t)j * ld;
            return Wzi[o] * Wdx[o] * S.dy1[i];
        });
        for (int i = t.tid; i < NB; i += NT) S.v[i] = (i < n) ? mu * S.dg[i] - S.v[i] : 0.0;
        __syncthreads();
        col_sums(n, S.dy2, t, [&](int i, int j) { return Wmi[i + (size_t)j * ld] * S.v[i]; });
        for (int i = n + t.tid; i < NB; i += NT) S.dy2[i] = 0.0;
        __syncthreads();

        // ---- dX = sym(Zi * (-Diag(dy2) X - Diag(dy1) dX1) + mu Zi + dX1) --
        if (MIX && use_mx) {
            const double *zi = Wzi;
            const auto zi_line = [&](int i, int k) { return zi[k + (size_t)i * ld]; };
            const auto t2_line = [&](int j, int k) {
                return (-S.dy2[k] * X[k + (size_t)j * ld]) + (-S.dy1[k] * Wdx[k + (size_t)j * ld]);
            };
            mx_product<NT>(zi_line, t2_line, n, mixs, Wsl, (int *)&S.pan[0][0], Ws, ld);
        } else {
            gemm_blk<true>(Ws, Wzi, X, Wdx, S.dy1, S.dy2, ld, n, S, t);
        }
        ew_pass<6>(
            n, t,
            [&](int i, int j, double *v) {
                if (i < j) return;
                const size_t o = i + (size_t)j * ld, p = j + (size_t)i * ld;
                v[0] = Ws[o];
                v[1] = Wzi[o];
                v[2] = Wdx[o];
                v[3] = Ws[p];
                v[4] = Wzi[p];
                v[5] = Wdx[p];
            },
            [&](int i, int j, const double *v) {
                if (i < j) return;
                const double a = (v[0] + mu * v[1]) + v[2];
                const double b = (v[3] + mu * v[4]) + v[5];
                const double w = (i == j) ? a : 0.5 * (a + b);
                Wdx[i + (size_t)j * ld] = w;
                Wdx[j + (size_t)i * ld] = w;
            });
        for (int i = t.tid; i < NB; i += NT) S.v[i] = S.dy2[i] + S.dy1[i];   // dy
        __syncthreads();

        // ---- primal line search: X + a dX psd -----------------------------
        double alpha_p = 1.0;
        {
            int tr_ = 0;
            bool ok = false;
            while (tr_ < LS_TRIALS) {
                ew_pass<2>(
                    n, t,
                    [&](int i, int j, double *v) {
                        if (i < j) return;
                        const size_t o = i + (size_t)j * ld;
                        v[0] = X[o];
                        v[1] = Wdx[o];
                    },
                    [&](int i, int j, const double *v) {
                        if (i >= j) Ws[i + (size_t)j * ld] = v[0] + alpha_p * v[1];
                    });
                ++tr_;
                if (blk_chol_test(Ws, ld, n, S, t)) { ok = true; break; }
                alpha_p *= 0.8;
            }
            trials += tr_;
            if (!ok) { status = ERR_LINESEARCH; break; }
            if (alpha_p < 1.0) alpha_p *= 0.95;
        }
        // ---- dual line search: Z + a Diag(dy) psd --------------------------
        // The first trial (a = 1) is tested by the inverse sweep (the same
        // pivots); when it passes, the step is taken in full and the inverse
        // is the next iteration's Zi (not in the debug run, which reads Zi
        // from the first work slab).
        double alpha_d = 1.0;
        {
            int tr_ = 0;
            bool ok = false;
            while (tr_ < LS_TRIALS) {
                ew_pass<1>(
                    n, t, [&](int i, int j, double *v) { v[0] = L[i + (size_t)j * ld]; },
                    [&](int i, int j, const double *v) {
                        Ws[i + (size_t)j * ld] = (i == j) ? S.zd[i] + alpha_d * S.v[i] : -v[0];
                    });
                ++tr_;
                if (tr_ == 1 && !dbg) {
                    if (blk_inverse(Ws, ld, n, S, t)) {
                        double *tmp = Wzi;
                        Wzi = Ws;
                        Ws = tmp;
                        zi_ready = true;
                        ok = true;
                        break;
                    }
                } else if (blk_chol_test(Ws, ld, n, S, t)) {
                    ok = true;
                    break;
                }
                alpha_d *= 0.8;
            }
            trials += tr_;
            if (!ok) { status = ERR_LINESEARCH; break; }
            if (alpha_d < 1.0) alpha_d *= 0.95;
        }
        if (args.alpha_trace && t.tid == 0 && iter <= args.trace_cap) {
            args.alpha_trace[(size_t)node * 2 * args.trace_cap + 2 * (iter - 1)] = alpha_p;
            args.alpha_trace[(size_t)node * 2 * args.trace_cap + 2 * (iter - 1) + 1] = alpha_d;
        }

        // ---- update -------------------------------------------------------
        for (int i = t.tid; i < n; i += NT) {
            S.y[i] = S.y[i] + alpha_d * S.v[i];
            S.zd[i] = S.zd[i] + alpha_d * S.v[i];
        }
        __syncthreads();
        double pz = 0.0, pl = 0.0;
        ew_pass<3>(
            n, t,
            [&](int i, int j, double *v) {
                const size_t o = i + (size_t)j * ld;
                v[0] = X[o];
                v[1] = Wdx[o];
                v[2] = L[o];
            },
            [&](int i, int j, const double *v) {
                const double x = v[0] + alpha_p * v[1];
                X[i + (size_t)j * ld] = x;
                const double z = (i == j) ? S.zd[i] : -v[2];
                pz += z * x;
                pl += v[2] * x;
            });
        double py = 0.0;
        for (int i = t.tid; i < n; i += NT) py += S.y[i];
        mu = block_sum(pz, S.red, t) / (2.0 * n);
        psi = block_sum(pl, S.red, t);
        phi = block_sum(py, S.red, t);
        if (alpha_p + alpha_d > 1.6) mu *= 0.5;
        if (alpha_p + alpha_d > 1.9) mu *= 0.2;
        gap = fabs(phi - psi);
        if (MIX && (gap < 1e-3 * fabs(phi) || (iter > 3 && (alpha_p < 0.3 || alpha_d < 0.3)))) mx_on = false;
    }

    if (t.tid == 0) {
        args.phi[node] = phi;
        args.iters[node] = iter;
        if (args.trials) args.trials[node] = trials;
        args.status[node] = status;
        if (args.flops)
            atomicAdd(args.flops, (unsigned long long)(MIX ? BQB_IPM_FLOPS_MIXED(n, iter, trials) +
                                                                 BQB_IPM_FLOPS_PRODUCTS(n, iter - iters_mx)
                                                           : BQB_IPM_FLOPS(n, iter, trials)));
        if (MIX && args.flops_mx)
            atomicAdd(args.flops_mx, (unsigned long long)BQB_IPM_FLOPS_MX(n, iters_mx, mixs));
    }
    if (args.y_out)
        for (int i = t.tid; i < n; i += NT) args.y_out[(size_t)node * args.ystride + i] = S.y[i];
    if (args.warm_ok && t.tid == 0) args.warm_ok[node] = 0;     // starts cold
}

template <bool MIX>
__global__ __launch_bounds__(NT, 1) void ipm_blk_kernel(IpmBatchArgs args)
{
    __shared__ BlkLds S;
    const int node = blockIdx.x;
    if (args.active && !args.active[node]) return;
    const int n = args.n[node];
    if (n < 1 || n > NB || n > args.ld || args.wslabs < (MIX ? 6 : 4)) {
        if (threadIdx.x == 0) {
            args.status[node] = ERR_SIZE;
            args.iters[node] = 0;
            if (args.warm_ok) args.warm_ok[node] = 0;
        }
        return;
    }
    Tid t;
    t.tid = threadIdx.x;
    t.tr = t.tid & 15;
    t.tc = t.tid >> 4;
    t.lane = t.tid & 63;
    t.wave = t.tid >> 6;
    blk_solve<MIX>(args, node, n, S, t);
}

}  // namespace
}  // namespace bqb

// The blocked solve, n <= 256: FP64 products with 4 work slabs per node,
// or the products on the matrix cores (a->mix_slices 2..4) with 6.
extern "C" hipError_t bqb_launch_ipm_blk(const IpmBatchArgs *a, int nodes, hipStream_t s)
{
    if (a->mix_slices) {
        if (a->wslabs < 6 || a->mix_slices < 2 || a->mix_slices > bqb::MX_MAX_SLICES) return hipErrorInvalidValue;
        hipLaunchKernelGGL(bqb::ipm_blk_kernel<true>, dim3(nodes), dim3(bqb::NT), 0, s, *a);
    } else {
        if (a->wslabs < 4) return hipErrorInvalidValue;
        hipLaunchKernelGGL(bqb::ipm_blk_kernel<false>, dim3(nodes), dim3(bqb::NT), 0, s, *a);
    }
    return hipGetLastError();
}
