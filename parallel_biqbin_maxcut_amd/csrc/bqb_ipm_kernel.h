// bqb_ipm_kernel.h -- batched FP64 primal-dual interior-point solve of the basic
// Max-Cut SDP  max <L,X> s.t. diag(X) = e, X psd  for many independent
// subproblems of different size (reference src/ipm_mc_pk.c:28-359).
//
// One workgroup bounds one subproblem and runs the whole predictor-corrector
// loop without leaving the kernel.  Per iteration:
//   Zi = inv(Diag(zd) - L)          symmetric sweep, registers
//   Mi = inv(Zi o X)                symmetric sweep, registers
//   dy1 = Mi (-e)
//   dX1 = sym(Zi (-Diag(dy1) X) - X)          register-tile GEMM, LDS panels
//   dy2 = Mi (mu diag(Zi) - (Zi o dX1) dy1)
//   dX  = sym(Zi (-Diag(dy2) X - Diag(dy1) dX1) + mu Zi + dX1)
//   alpha_p / alpha_d by Cholesky tests of X + a dX and Z + a Diag(dy),
//   a <- 0.8 a on failure, then 0.95 a if a < 1
//   X += alpha_p dX, y += alpha_d dy, Z += alpha_d Diag(dy),
//   mu = <Z,X>/2n (x0.5 if alpha_p+alpha_d > 1.6, x0.2 if > 1.9)
// until |e'y - <L,X>| <= 1e-2.  Z is kept as its diagonal zd (its off-diagonal
// is -L and never changes, as in the reference where dZ is diagonal).
//
// Only one n x n register tile is live at a time; Zi, Mi and dX1/dX are kept
// in the node's HBM slab between phases and streamed back (column-contiguous
// panels for the GEMMs).
//
// The reference factors Zi o X twice with dposv; here it is inverted once and
// both right-hand sides are matrix-vector products (same solution in exact
// arithmetic).  Zi = inv(Z) uses the sweep instead of dpotrf+dpotri.
#pragma once
#include <hip/hip_runtime.h>

#include "bqb_dist.h"
#include "bqb_kernels.h"
#include "bqb_mixed.h"

namespace bqb {

constexpr int IPM_MAX_ITERS = 500;

// Keeps the loads of one register column from being hoisted above the
// previous column: caps the registers held by loads in flight, so the kernel
// fits two waves per SIMD (two subproblems per CU).
#define BQB_COLUMN_FENCE() __builtin_amdgcn_sched_barrier(0)
constexpr int LS_MAX_TRIALS = 500;
constexpr int KB = 8;   // GEMM panel depth

// Copies columns k0 .. k0+KB-1 (rows 0 .. NMAX-1) of a slab into LDS at
// dst[kk*NMAX + J] with global_load_lds (16 bytes per lane, written by the
// memory pipeline straight into LDS: no registers).  Every wave covers 64
// consecutive 16-byte chunks per instruction; the slab's zero padding
// supplies rows and columns past n.  Completion is waited for with vmcnt.
template <int P, int R>
__device__ __forceinline__ void panel_to_lds(double *dst, const double *g, int ld, int k0, const Thr &t)
{
    constexpr int NMAX = P * R;
    static_assert(NMAX % 2 == 0, "16-byte chunks");
    constexpr int CPC = NMAX / 2;        // chunks per column
    constexpr int CH = KB * CPC;         // chunks per panel
    constexpr int NW = P * P / 64;
    const int lane = t.tid & 63, wave = t.tid >> 6;
    for (int j = wave; j * 64 < CH; j += NW) {
        const int c = j * 64 + lane;
        const int kk = c / CPC, r = c - kk * CPC;
        if (c < CH)
            __builtin_amdgcn_global_load_lds((const void *)(g + (size_t)(k0 + kk) * ld + 2 * r),
                                             (__attribute__((address_space(3))) void *)(dst + j * 128), 16, 0, 0);
    }
}

__device__ __forceinline__ void wait_panel_dma() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// C += A * T for a symmetric A held in HBM (column-major slab, ld) and
// T(k, J) = -s1[k] X(J, k)  (or, with W, (-s2[k] X(J, k)) + (-s1[k] W(J, k)));
// every operand is symmetric, so row k of T is formed from column k.
// Panels of KB columns of A, X (and W) are copied to LDS by DMA, double
// buffered: the next panel's copy runs during the current panel's products,
// one barrier per panel.  pan holds 2 * (2 or 3) * KB * NMAX doubles.
template <int P, int R, bool TWO>
__device__ __forceinline__ void gemm_sym(double (&C)[R][R], const double *Ag, const double *Xg, const double *Wg,
                                         const double *s1, const double *s2, int ld, int n, double *pan,
                                         const Thr &t)
{
    constexpr int NMAX = P * R;
    constexpr int NP = TWO ? 3 : 2;                 // panels per stage
    constexpr int STAGE = NP * KB * NMAX;
    auto issue = [&](int k0, int buf) {
        double *st = pan + buf * STAGE;
        panel_to_lds<P, R>(st, Ag, ld, k0, t);
        panel_to_lds<P, R>(st + KB * NMAX, Xg, ld, k0, t);
        if (TWO) panel_to_lds<P, R>(st + 2 * KB * NMAX, Wg, ld, k0, t);
    };
    __syncthreads();   // pan may still be read by the previous user
    int buf = 0;
    issue(0, 0);
    for (int k0 = 0; k0 < n; k0 += KB) {
        wait_panel_dma();
        __syncthreads();
        if (k0 + KB < n) issue(k0 + KB, buf ^ 1);
        const double *Ap = pan + buf * STAGE;
        const double *Xp = Ap + KB * NMAX;
        const double *Wp = Xp + KB * NMAX;
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            const int k = k0 + q;
            const double m1 = -s1[k];
            const double m2 = TWO ? -s2[k] : 0.0;
            double aI[R], tJ[R];
#pragma unroll
            for (int a = 0; a < R; ++a) aI[a] = Ap[q * NMAX + t.tr + P * a];
#pragma unroll
            for (int b = 0; b < R; ++b) {
                const int e = q * NMAX + t.tc + P * b;
                tJ[b] = TWO ? (m2 * Xp[e]) + (m1 * Wp[e]) : m1 * Xp[e];
            }
#pragma unroll
            for (int a = 0; a < R; ++a)
#pragma unroll
                for (int b = 0; b < R; ++b) C[a][b] = fma(aI[a], tJ[b], C[a][b]);
        }
        buf ^= 1;
    }
    __syncthreads();
}

// LDS of one solve (sized by the kernel for its largest tile).
struct IpmLds {
    double *col, *scr, *y, *zd, *dy1, *dy2, *v, *dg, *red;
};

// The solve of one subproblem with an R x R register tile per thread (the
// smallest tile that covers its n; the slabs keep the kernel's leading
// dimension).  MIX: the two products run on the bf16 matrix cores
// (bqb_mixed.h) with args.mix_slices slices; everything else is FP64.
template <int P, int R, bool MIX>
__device__ __forceinline__ void ipm_solve(const IpmBatchArgs &args, const int node, const int n,
                                          const IpmLds &S)
{
    constexpr int NMAX = P * R;
    double *const s_col = S.col, *const s_scr = S.scr, *const s_y = S.y, *const s_zd = S.zd;
    double *const s_dy1 = S.dy1, *const s_dy2 = S.dy2, *const s_v = S.v, *const s_dg = S.dg;
    double *const s_red = S.red;
    const Thr t = thr_id<P>();
    const int ld = args.ld;
    const size_t nn = (size_t)ld * ld;
    const double *L = args.L + (size_t)node * nn;
    double *X = args.X + (size_t)node * nn;
    double *Wzi = args.work + (size_t)node * (MIX ? 6 : 3) * nn;   // inv(Z)
    double *Wmi = Wzi + nn;                              // inv(Zi o X)
    double *Wdx = Wmi + nn;                              // dX1, then dX
    // mixed products: bf16 slices of both operands (3 slabs)
    unsigned short *Wsl = MIX ? (unsigned short *)(Wdx + nn) : nullptr;
    const int mixs = MIX ? args.mix_slices : 0;
    const Slab sL = make_slab<P>(L, ld, t), sX = make_slab<P>(X, ld, t);
    const Slab sZi = make_slab<P>(Wzi, ld, t), sMi = make_slab<P>(Wmi, ld, t);
    const Slab sDx = make_slab<P>(Wdx, ld, t);

    double A[R][R];
    double phi, psi, mu;

    // The slot's last solve (warm_ok) left its X, y and L: a bundle step
    // changes L by B'(dgamma) only, and starting there takes about a third
    // fewer iterations than the reference's fixed point.
    const bool keep_L = args.lprev && args.warm_ok && args.y_out && args.warm_mode == 1;
    bool warm = keep_L && args.warm_ok[node] != 0;
    const Slab sLp = make_slab<P>(keep_L ? args.lprev + (size_t)node * nn : L, ld, t);

    // ---- initial point (ipm_mc_pk.c:65-99) -------------------------------
    // y_j = 1.1 + sum_i |L_ij| (column sums: reduce over tr inside the wave)
    auto cold_start = [&]() {
        dload<P, R>(A, sL);
        {
            double cs[R];
#pragma unroll
            for (int b = 0; b < R; ++b) {
                double s = 0.0;
#pragma unroll
                for (int a = 0; a < R; ++a) s += fabs(A[a][b]);
#pragma unroll
                for (int o = 1; o < P; o <<= 1) s += __shfl_xor(s, o, 64);
                cs[b] = s;
            }
            if (t.tr == 0) {
#pragma unroll
                for (int b = 0; b < R; ++b) {
                    const int J = t.tc + P * b;
                    s_y[J] = (J < n) ? cs[b] + 1.1 : 0.0;
                }
            }
#pragma unroll
            for (int a = 0; a < R; ++a)
                if (t.tr == t.tc) s_dg[t.tr + P * a] = A[a][a];
        }
        if (keep_L) dstore<P, R>(A, sLp);                  // the next warm start's L
        __syncthreads();
        for (int I = t.tid; I < NMAX; I += P * P) s_zd[I] = (I < n) ? s_y[I] - s_dg[I] : 0.0;
#pragma unroll
        for (int a = 0; a < R; ++a)
#pragma unroll
            for (int b = 0; b < R; ++b) A[a][b] = (t.tr + P * a == t.tc + P * b && t.tr + P * a < n) ? 1.0 : 0.0;
        dstore<P, R>(A, sX);                                // X = I
        if constexpr (MIX) {
            // the products write only the n x n interior of their output slabs,
            // whose padding the later tile loads need as zeros
#pragma unroll
            for (int a = 0; a < R; ++a)
#pragma unroll
                for (int b = 0; b < R; ++b) A[a][b] = 0.0;
            dstore<P, R>(A, sDx);
            dstore<P, R>(A, sMi);
        }
        __syncthreads();
        {
            double py = 0.0, pl = 0.0, pz = 0.0;
            for (int I = t.tid; I < n; I += P * P) { py += s_y[I]; pl += s_dg[I]; pz += s_zd[I]; }
            phi = wg_sum<P>(py, s_red, t);
            psi = wg_sum<P>(pl, s_red, t);
            mu = wg_sum<P>(pz, s_red, t) / (2.0 * n);
        }
    };

    // X0 = (1 - th) X + th I from the last solution, and
    // y0_i = y_i + dL_ii + sum_{j != i} |dL_ij| + margin with dL = L - Lprev:
    // Z0 = Z_last + (Diag(y0 - y) - dL) is the last (positive definite) dual
    // slack plus a diagonally dominant matrix, so positive definite.  Any
    // y with Z positive definite gives a valid bound e'y.
    auto warm_start = [&]() {
        constexpr double th = 0.05, margin = 0.01;
        const double *yl = args.y_out + (size_t)node * args.ystride;
        dload<P, R>(A, sL);
        double cs[R];
#pragma unroll
        for (int b = 0; b < R; ++b) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < R; ++a) {
                const double dl = A[a][b] - sld<P>(sLp, a, b);
                s += (t.tr + P * a == t.tc + P * b) ? dl : fabs(dl);
            }
#pragma unroll
            for (int o = 1; o < P; o <<= 1) s += __shfl_xor(s, o, 64);
            cs[b] = s;
        }
        if (t.tr == 0) {
#pragma unroll
            for (int b = 0; b < R; ++b) {
                const int J = t.tc + P * b;
                s_y[J] = (J < n) ? yl[J] + cs[b] + margin : 0.0;
            }
        }
#pragma unroll
        for (int a = 0; a < R; ++a)
            if (t.tr == t.tc) s_dg[t.tr + P * a] = A[a][a];
        dstore<P, R>(A, sLp);                           // each thread its own tile
        dload<P, R>(A, sX);
#pragma unroll
        for (int a = 0; a < R; ++a)
#pragma unroll
            for (int b = 0; b < R; ++b) {
                const int I = t.tr + P * a, J = t.tc + P * b;
                A[a][b] = (1.0 - th) * A[a][b] + ((I == J && I < n) ? th : 0.0);
            }
        dstore<P, R>(A, sX);
        if constexpr (MIX) {
            double Zr[R][R];
#pragma unroll
            for (int a = 0; a < R; ++a)
#pragma unroll
                for (int b = 0; b < R; ++b) Zr[a][b] = 0.0;
            dstore<P, R>(Zr, sDx);
            dstore<P, R>(Zr, sMi);
        }
        __syncthreads();
        for (int I = t.tid; I < NMAX; I += P * P) s_zd[I] = (I < n) ? s_y[I] - s_dg[I] : 0.0;
        __syncthreads();
        double pz = 0.0, pl = 0.0;
#pragma unroll
        for (int b = 0; b < R; ++b) {
#pragma unroll
            for (int a = 0; a < R; ++a) {
                const int I = t.tr + P * a, J = t.tc + P * b;
                const double l = sld<P>(sL, a, b);
                const double z = (I == J) ? s_zd[I] : -l;
                pz += z * A[a][b];
                pl += l * A[a][b];
            }
        }
        double py = 0.0;
        for (int I = t.tid; I < n; I += P * P) py += s_y[I];
        mu = wg_sum<P>(pz, s_red, t) / (2.0 * n);
        psi = wg_sum<P>(pl, s_red, t);
        phi = wg_sum<P>(py, s_red, t);
    };

    if (warm) warm_start();
    else cold_start();
    double gap = fabs(phi - psi);
    int iter = 0, trials = 0, status = 0;
    if (args.dbg) {
        double *dbg = args.dbg + (size_t)node * 4 * ld;
        for (int I = t.tid; I < NMAX; I += P * P) {
            dbg[I] = s_y[I];
            dbg[ld + I] = s_zd[I];
            dbg[2 * ld + I] = s_dg[I];
        }
        if (t.tid == 0) { dbg[3 * ld] = phi; dbg[3 * ld + 1] = psi; dbg[3 * ld + 2] = mu; }
    }

    const int max_iters = args.max_iters > 0 ? args.max_iters : IPM_MAX_ITERS;
    // set when the dual line search accepted the full step at its first
    // trial: that trial inverted exactly the next iteration's Z (same
    // diagonal zd + 1 * dy, same off-diagonal), so Zi is already in its slab
    bool zi_ready = false;
    // mixed products while the iterate is far from optimal; FP64 products
    // after the first iteration whose relative gap is below 1e-3, or, from
    // the fourth on, whose primal or dual step was cut below 0.3 (near the
    // optimum the slices' product error would dominate dX and stall the
    // line searches)
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

        // ---- Zi = inv(Z) --------------------------------------------------
        if (zi_ready) {
            dload<P, R>(A, sZi);
        } else {
            dload<P, R>(A, sL);
#pragma unroll
            for (int a = 0; a < R; ++a)
#pragma unroll
                for (int b = 0; b < R; ++b) {
                    const int I = t.tr + P * a, J = t.tc + P * b;
                    A[a][b] = (I == J) ? s_zd[I] : -A[a][b];   // Z = Diag(zd) - offdiag(L)
                }
            if (!sweep_inverse<P, R>(A, n, s_col, s_scr, t)) {
                if (warm && iter == 1) {       // rounding: the reference's start
                    warm = false;
                    __syncthreads();
                    cold_start();
                    gap = fabs(phi - psi);
                    iter = 0;
                    continue;
                }
                status = ERR_CHOL_Z;
                break;
            }
            dstore<P, R>(A, sZi);
        }
        zi_ready = false;
#pragma unroll
        for (int a = 0; a < R; ++a)
            if (t.tr == t.tc) s_dg[t.tr + P * a] = A[a][a];

        // ---- Mi = inv(Zi o X), dy1 = Mi (-e) ------------------------------
#pragma unroll
        for (int b = 0; b < R; ++b)
#pragma unroll
            for (int a = 0; a < R; ++a) A[a][b] = A[a][b] * sld<P>(sX, a, b);
        if (!sweep_inverse<P, R>(A, n, s_col, s_scr, t)) { status = ERR_SCHUR; break; }
        {
            double part[R];
#pragma unroll
            for (int a = 0; a < R; ++a) {
                double s = 0.0;
#pragma unroll
                for (int b = 0; b < R; ++b) s += A[a][b];
                part[a] = s;
            }
            row_sums<P, R>(part, s_scr, s_v, t);
            for (int I = t.tid; I < NMAX; I += P * P) s_dy1[I] = -s_v[I];
        }
        dstore<P, R>(A, sMi);
        __syncthreads();

        // ---- dX1 = sym(Zi * (-Diag(dy1) X) - X) ---------------------------
        if (MIX && use_mx) {
            // row i of Zi is its column i; column j of -Diag(dy1) X is
            // -dy1 o X(:, j) (X symmetric)
            const auto zi_line = [&](int i, int k) { return Wzi[k + (size_t)i * ld]; };
            const auto t1_line = [&](int j, int k) { return -s_dy1[k] * X[k + (size_t)j * ld]; };
            mx_product<P * P>(zi_line, t1_line, n, mixs, Wsl, (int *)s_scr, Wdx, ld);
            dload<P, R>(A, sDx);
        } else {
#pragma unroll
            for (int a = 0; a < R; ++a)
#pragma unroll
                for (int b = 0; b < R; ++b) A[a][b] = 0.0;
            gemm_sym<P, R, false>(A, Wzi, X, nullptr, s_dy1, nullptr, ld, n, s_scr, t);
        }
#pragma unroll
        for (int b = 0; b < R; ++b)
#pragma unroll
            for (int a = 0; a < R; ++a) A[a][b] = A[a][b] - sld<P>(sX, a, b);
        symmetrize<P, R>(A, s_scr, t);
        dstore<P, R>(A, sDx);

        // ---- corrector rhs: mu diag(Zi) - (Zi o dX1) dy1 ------------------
        {
            double part[R];
#pragma unroll
            for (int a = 0; a < R; ++a) part[a] = 0.0;
#pragma unroll
            for (int b = 0; b < R; ++b) {
                const double vj = s_dy1[t.tc + P * b];
#pragma unroll
                for (int a = 0; a < R; ++a) part[a] += sld<P>(sZi, a, b) * A[a][b] * vj;
            }
            row_sums<P, R>(part, s_scr, s_v, t);
            for (int I = t.tid; I < NMAX; I += P * P) s_v[I] = (I < n) ? mu * s_dg[I] - s_v[I] : 0.0;
            __syncthreads();
        }
        // dy2 = Mi * rhs (Mi streamed back)
        {
            double part[R];
#pragma unroll
            for (int a = 0; a < R; ++a) part[a] = 0.0;
#pragma unroll
            for (int b = 0; b < R; ++b) {
                const double vj = s_v[t.tc + P * b];
#pragma unroll
                for (int a = 0; a < R; ++a) part[a] += sld<P>(sMi, a, b) * vj;
            }
            row_sums<P, R>(part, s_scr, s_dy2, t);
        }

        // ---- dX = sym(Zi * (-Diag(dy2) X - Diag(dy1) dX1) + mu Zi + dX1) --
        if (MIX && use_mx) {
            // the product goes to the Mi slab (Mi is not read again in this
            // iteration); dX1 stays in its slab for the sum below
            const auto zi_line = [&](int i, int k) { return Wzi[k + (size_t)i * ld]; };
            const auto t2_line = [&](int j, int k) {
                return (-s_dy2[k] * X[k + (size_t)j * ld]) + (-s_dy1[k] * Wdx[k + (size_t)j * ld]);
            };
            mx_product<P * P>(zi_line, t2_line, n, mixs, Wsl, (int *)s_scr, Wmi, ld);
            dload<P, R>(A, sMi);
        } else {
#pragma unroll
            for (int a = 0; a < R; ++a)
#pragma unroll
                for (int b = 0; b < R; ++b) A[a][b] = 0.0;
            gemm_sym<P, R, true>(A, Wzi, X, Wdx, s_dy1, s_dy2, ld, n, s_scr, t);
        }
#pragma unroll
        for (int b = 0; b < R; ++b) {
#pragma unroll
            for (int a = 0; a < R; ++a)
                A[a][b] = (A[a][b] + mu * sld<P>(sZi, a, b)) + sld<P>(sDx, a, b);
            BQB_COLUMN_FENCE();
        }
        symmetrize<P, R>(A, s_scr, t);
        dstore<P, R>(A, sDx);                         // dX
        for (int I = t.tid; I < NMAX; I += P * P) s_v[I] = s_dy2[I] + s_dy1[I];   // dy
        __syncthreads();

        // ---- primal line search: X + a dX psd -----------------------------
        double alpha_p = 1.0;
        {
            int tr_ = 0;
            bool ok = false;
            while (tr_ < LS_MAX_TRIALS) {
#pragma unroll
                for (int b = 0; b < R; ++b) {
#pragma unroll
                    for (int a = 0; a < R; ++a)
                        A[a][b] = sld<P>(sX, a, b) + alpha_p * sld<P>(sDx, a, b);
                    BQB_COLUMN_FENCE();
                }
                ++tr_;
                if (chol_test<P, R>(A, n, s_col, t)) { ok = true; break; }
                alpha_p *= 0.8;
            }
            trials += tr_;
            if (!ok) { status = ERR_LINESEARCH; break; }
            if (alpha_p < 1.0) alpha_p *= 0.95;
        }
        // ---- dual line search: Z + a Diag(dy) psd --------------------------
        // The first trial (a = 1) is tested by the sweep inverse: its pivots
        // are the Cholesky test's (the same trailing updates), and when it
        // passes, the step is taken in full and the inverse is the next
        // iteration's Zi (not in the debug run, which reads the Zi slab).
        double alpha_d = 1.0;
        {
            int tr_ = 0;
            bool ok = false;
            while (tr_ < LS_MAX_TRIALS) {
                dload<P, R>(A, sL);
#pragma unroll
                for (int a = 0; a < R; ++a)
#pragma unroll
                    for (int b = 0; b < R; ++b) {
                        const int I = t.tr + P * a, J = t.tc + P * b;
                        A[a][b] = (I == J) ? s_zd[I] + alpha_d * s_v[I] : -A[a][b];
                    }
                ++tr_;
                if (tr_ == 1 && !args.dbg) {
                    if (sweep_inverse<P, R>(A, n, s_col, s_scr, t)) {
                        dstore<P, R>(A, sZi);
                        zi_ready = true;
                        ok = true;
                        break;
                    }
                } else if (chol_test<P, R>(A, n, s_col, t)) {
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
#pragma unroll
        for (int b = 0; b < R; ++b) {
#pragma unroll
            for (int a = 0; a < R; ++a)
                A[a][b] = sld<P>(sX, a, b) + alpha_p * sld<P>(sDx, a, b);
            BQB_COLUMN_FENCE();
        }
        dstore<P, R>(A, sX);
        for (int I = t.tid; I < n; I += P * P) {
            s_y[I] = s_y[I] + alpha_d * s_v[I];
            s_zd[I] = s_zd[I] + alpha_d * s_v[I];
        }
        __syncthreads();
        {
            double pz = 0.0, pl = 0.0;
#pragma unroll
            for (int b = 0; b < R; ++b) {
#pragma unroll
                for (int a = 0; a < R; ++a) {
                    const int I = t.tr + P * a, J = t.tc + P * b;
                    const double l = sld<P>(sL, a, b);
                    const double z = (I == J) ? s_zd[I] : -l;
                    pz += z * A[a][b];
                    pl += l * A[a][b];
                }
                BQB_COLUMN_FENCE();
            }
            double py = 0.0;
            for (int I = t.tid; I < n; I += P * P) py += s_y[I];
            mu = wg_sum<P>(pz, s_red, t) / (2.0 * n);
            psi = wg_sum<P>(pl, s_red, t);
            phi = wg_sum<P>(py, s_red, t);
        }
        if (alpha_p + alpha_d > 1.6) mu *= 0.5;
        if (alpha_p + alpha_d > 1.9) mu *= 0.2;
        gap = fabs(phi - psi);
        if (MIX && (gap < 1e-3 * fabs(phi) || (iter > 3 && (alpha_p < 0.3 || alpha_d < 0.3)))) mx_on = false;
    }

    if (t.tid == 0) {
        args.phi[node] = phi;
        args.iters[node] = iter;
        args.trials[node] = trials;
        args.status[node] = status;
        if (args.flops)
            atomicAdd(args.flops, (unsigned long long)(MIX ? BQB_IPM_FLOPS_MIXED(n, iter, trials) +
                                                                 BQB_IPM_FLOPS_PRODUCTS(n, iter - iters_mx)
                                                           : BQB_IPM_FLOPS(n, iter, trials)));
        if (MIX && args.flops_mx)
            atomicAdd(args.flops_mx, (unsigned long long)BQB_IPM_FLOPS_MX(n, iters_mx, mixs));
    }
    if (args.y_out)
        for (int I = t.tid; I < n; I += P * P) args.y_out[(size_t)node * args.ystride + I] = s_y[I];
    if (args.warm_ok && t.tid == 0) args.warm_ok[node] = (keep_L && status == 0) ? 1 : 0;
}

// One workgroup per subproblem.  The register tile is chosen per
// subproblem: ceil(n / P) blocks, so a small node does not pay for the
// padding of the largest one.
// 256-thread tiles are held to two workgroups per CU (the 16 x 7 tile then
// keeps part of its state in scratch): with several slot groups in flight,
// a second resident workgroup per CU gains more than the spills cost.
template <int P, int R, bool MIX>
__global__ __launch_bounds__(P * P, (P == 16) ? 2 : 1) void ipm_batch_kernel(IpmBatchArgs args)
{
    constexpr int NMAX = P * R;
    constexpr int NW = P * P / 64;
    using L_ = Lds<P, R, KB>;
    __shared__ double s_col[L_::COL];
    __shared__ __align__(16) double s_scr[L_::UNION];
    __shared__ double s_y[NMAX], s_zd[NMAX], s_dy1[NMAX], s_dy2[NMAX], s_v[NMAX], s_dg[NMAX];
    __shared__ double s_red[NW];

    const int node = blockIdx.x;
    if (args.active && !args.active[node]) return;
    const int n = args.n[node];
    if (n < 1 || n > NMAX || args.ld < NMAX) {
        if (threadIdx.x == 0) {
            args.status[node] = ERR_SIZE;
            args.iters[node] = 0;
            if (args.warm_ok) args.warm_ok[node] = 0;
        }
        return;
    }
    const IpmLds S = {s_col, s_scr, s_y, s_zd, s_dy1, s_dy2, s_v, s_dg, s_red};
    const int re = (n + P - 1) / P;
    static_for<1, R + 1>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        if (re == r) ipm_solve<P, r, MIX>(args, node, n, S);
    });
}

}  // namespace bqb
