// bqb_bundle.hip -- proximal bundle method over the cut multipliers of a
// batch of nodes (reference src/bundle.c:9-373, 390-653).
//   bundle_qp_kernel:   zeta = -F - G'gamma, Q = t G'G, then up to 50 rounds of
//                       { c = zeta - t G'eta; lambda = argmin of the simplex QP;
//                         eta = max(0, -gamma/t + G lambda);
//                         dgamma = t (eta - G lambda) }  and gamma_test;
//                       G is read once per round (G lambda and G'eta of the
//                       next round from the same loads).
//   bundle_step_kernel: after fct_eval(gamma_test) -- the serious / null step
//                       test, X = X_bundle lambda, purge and append.
// The simplex QP (k <= 64 bundle elements) is solved by the reference's
// primal-dual interior-point loop on one wave with LDS-resident data; its
// (k+1) x (k+1) KKT system [-Q-Diag(z./lambda) e; e' 0] is solved through the
// Schur complement with a Cholesky factor of Q + Diag(z./lambda).
#include <hip/hip_runtime.h>

#include "bqb_batch.h"
#include "bqb_cutops.h"
#include "bqb_dist.h"
#include "bqb_kernels.h"
#include "bqb_wave.h"

namespace bqb {

constexpr int BT = BQB_QP_THREADS;
constexpr int KMAX = 64;

__device__ __forceinline__ double wsum(double v) { return wave_sum64(v); }
__device__ __forceinline__ double wmax(double v) { return wave_max64(v); }
__device__ __forceinline__ double wmin(double v) { return wave_min64(v); }

__device__ __forceinline__ double bsum(double v, double *red)
{
    v = wsum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < BT / 64; ++w) s += red[w];
    __syncthreads();
    return s;
}

// Cholesky L L' = A of a k x k SPD matrix in LDS (column-major, ld k), one
// wave, lane i owns row i.  Returns false if a pivot is not positive.
__device__ __forceinline__ bool wave_chol(double *A, int k, int lane)
{
    for (int j = 0; j < k; ++j) {
        const double d = A[j + j * k];
        if (!(d > 0.0)) return false;
        const double ljj = sqrt(d);
        if (lane > j && lane < k) A[lane + j * k] = A[lane + j * k] / ljj;
        __builtin_amdgcn_wave_barrier();
        if (lane == j) A[j + j * k] = ljj;
        __builtin_amdgcn_wave_barrier();
        if (lane > j && lane < k) {
            const double lij = A[lane + j * k];
            for (int l = j + 1; l <= lane; ++l) A[lane + l * k] -= lij * A[l + j * k];
        }
        __builtin_amdgcn_wave_barrier();
    }
    return true;
}

// Solve (L L') x = b in place (b in LDS), one wave.
__device__ __forceinline__ void wave_cholsolve(const double *Lc, int k, double *b, int lane)
{
    for (int j = 0; j < k; ++j) {
        if (lane == j) b[j] = b[j] / Lc[j + j * k];
        __builtin_amdgcn_wave_barrier();
        const double xj = b[j];
        if (lane > j && lane < k) b[lane] -= Lc[lane + j * k] * xj;
        __builtin_amdgcn_wave_barrier();
    }
    for (int j = k - 1; j >= 0; --j) {
        if (lane == j) b[j] = b[j] / Lc[j + j * k];
        __builtin_amdgcn_wave_barrier();
        const double xj = b[j];
        if (lane < j) b[lane] -= Lc[j + lane * k] * xj;
        __builtin_amdgcn_wave_barrier();
    }
}

// bundle.c:390-653 on one wave.  Q (k x k), c, lambda in LDS.
__device__ __forceinline__ int solve_lambda_wave(int k, const double *Q, const double *c, double *lam, double *A,
                                 double *u, double *w, int lane)
{
    const bool in = lane < k;
    double l = in ? 1.0 / k : 0.0;
    if (in) lam[lane] = l;
    __builtin_amdgcn_wave_barrier();
    auto qlam = [&](double lv_unused) {
        (void)lv_unused;
        double s = 0.0;
        if (in)
            for (int j = 0; j < k; ++j) s += Q[lane + j * k] * lam[j];
        return s;
    };
    double Ql = qlam(0.0);
    const double ci = in ? c[lane] : 0.0;
    const double tmpv = in ? Ql + ci : 1e300;
    const double mintmp = wmin(tmpv);
    double y, z;
    if (mintmp > 1) { y = 0.0; z = in ? tmpv : 0.0; }
    else { y = mintmp - 1.0; z = in ? tmpv - y : 0.0; }
    double mu = wsum(in ? z * l : 0.0) / k * 0.5;
    double res_p = 1.0 - wsum(in ? l : 0.0);
    double temp = wsum(in ? l * Ql : 0.0);
    double D_cost = y - 0.5 * temp;
    double P_cost = wsum(in ? l * ci : 0.0) + 0.5 * temp;
    double gap = P_cost - D_cost;
    int cnt = 0;
    while (fabs(gap) > 1e-5) {
        ++cnt;
        for (int j = 0; j < k; ++j)
            if (in) A[lane + j * k] = Q[lane + j * k];
        __builtin_amdgcn_wave_barrier();
        if (in) {
            A[lane + lane * k] = -(-Q[lane + lane * k] - z / l);
            w[lane] = ci - y + Ql;
            u[lane] = 1.0;
        }
        __builtin_amdgcn_wave_barrier();
        if (!wave_chol(A, k, lane)) return ERR_QP;
        wave_cholsolve(A, k, u, lane);
        wave_cholsolve(A, k, w, lane);
        const double ui = in ? u[lane] : 0.0, wi = in ? w[lane] : 0.0;
        const double su = wsum(ui), sw = wsum(wi);
        const double dy = (res_p + sw) / su;
        const double dl = in ? ui * dy - wi : 0.0;
        const double dz = in ? 1.0 / l * (mu - l * z - z * dl) : 0.0;
        double ap = wmax(in ? -dl / l : -1e9);
        double ad = wmax(in ? -dz / z : -1e9);
        ap = (ap > 0) ? ((0.99 / ap < 1.0) ? 0.99 / ap : 1.0) : 1.0;
        ad = (ad > 0) ? ((0.99 / ad < 1.0) ? 0.99 / ad : 1.0) : 1.0;
        y += ad * dy;
        if (in) { l = l + ap * dl; z = z + ad * dz; lam[lane] = l; }
        __builtin_amdgcn_wave_barrier();
        res_p = 1 - wsum(in ? l : 0.0);
        Ql = qlam(0.0);
        mu = wsum(in ? l * z : 0.0) / k * 0.4;
        if (ap + ad > 1.8) mu *= 0.2;
        temp = wsum(in ? l * Ql : 0.0);
        D_cost = y - 0.5 * temp;
        P_cost = wsum(in ? l * ci : 0.0) + 0.5 * temp;
        gap = P_cost - D_cost;
        if (cnt > 30) break;
    }
    return 0;
}

// The same algorithm for k <= K (K = 8, 12, 16, 20, 24, 28, 32) with the factor in
// registers: lane i holds row i of A = Q + Diag(z./lambda) and of its
// Cholesky factor.  Pivots are read with v_readlane, each step's column is
// stored to LDS (the back substitution reads the transpose from there).
// The rank-1 updates run unmasked: lanes at or above k and entries above
// the diagonal collect values that are never read.  Divisions by a pivot
// are multiplications by its reciprocal (one divide per column), so results
// agree with the LDS version to rounding, not bit for bit.
constexpr int KR = 32;
// bundle elements whose G'eta partial sums stay in registers in the row pass
constexpr int PE = BQB_QP_REG_PARTIALS;

template <int K>
__device__ __forceinline__ int solve_lambda_reg(int k, const double *Q, const double *c, double *lam, double *T,
                                int lane)
{
    const bool in = lane < k;
    double l = in ? 1.0 / k : 0.0;
    if (in) lam[lane] = l;
    __builtin_amdgcn_wave_barrier();
    auto qlam = [&]() {
        double s = 0.0;
        if (in)
            for (int j = 0; j < k; ++j) s += Q[lane + j * k] * lam[j];
        return s;
    };
    double Ql = qlam();
    const double ci = in ? c[lane] : 0.0;
    const double tmpv = in ? Ql + ci : 1e300;
    const double mintmp = wave_min32(tmpv);
    double y, z;
    if (mintmp > 1) { y = 0.0; z = in ? tmpv : 0.0; }
    else { y = mintmp - 1.0; z = in ? tmpv - y : 0.0; }
    double mu = wave_sum32(in ? z * l : 0.0) / k * 0.5;
    double res_p = 1.0 - wave_sum32(in ? l : 0.0);
    double temp = wave_sum32(in ? l * Ql : 0.0);
    double D_cost = y - 0.5 * temp;
    double P_cost = wave_sum32(in ? l * ci : 0.0) + 0.5 * temp;
    double gap = P_cost - D_cost;
    int cnt = 0;
    const int qrow = in ? lane : 0;
    double Ar[K], Lt[K];
    while (fabs(gap) > 1e-5) {
        ++cnt;
        // row `lane` of A = Q with the diagonal -(-Q_ii - z/l)
        const double zl = in ? z / l : 0.0;
        static_for<0, K>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const double q = Q[qrow + (j < k ? j : 0) * k];
            Ar[j] = (j == lane) ? -(-q - zl) : q;
        });
        // right-looking Cholesky (wave_chol).  The entries L_l2j of column
        // j reach every lane by v_readlane (no LDS round trip between
        // pivots); the column also goes to LDS at T[j*(K+1) + i] for the
        // transpose below; lane j keeps 1/L_jj in rd.
        // The pivots run without branches, so that the compiler can
        // interleave them: pivots past k only touch padding rows and columns,
        // and a non-positive real pivot is reported after the loop (the
        // factor is not used then).
        bool ok = true;
        double rd = 0.0;
        static_for<0, K>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const double dj = lane_value(Ar[j], j);
            ok = ok & ((j >= k) | (dj > 0.0));
            const double ljj = sqrt(dj);
            const double rj = 1.0 / ljj;
            const double lij = (lane == j) ? ljj : Ar[j] * rj;
            Ar[j] = lij;
            if (lane == j) rd = rj;
            T[j * (K + 1) + lane] = lij;
            static_for<j + 1, K>([&](auto lc) {
                constexpr int l2 = decltype(lc)::value;
                Ar[l2] -= lij * lane_value(lij, l2);
            });
        });
        __builtin_amdgcn_wave_barrier();
        if (!ok) return ERR_QP;
        // the factor's transpose: lane i gets row i of L' (L[j][i], j >= i)
        static_for<0, K>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            Lt[j] = T[lane * (K + 1) + j];
        });
        __builtin_amdgcn_wave_barrier();
        // A u = e and A w = r1 (wave_cholsolve on both right-hand sides)
        double u = 1.0, w = ci - y + Ql;
        static_for<0, K>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j >= k) return;
            const double rjj = lane_value(rd, j);
            const double uj = lane_value(u, j) * rjj, wj = lane_value(w, j) * rjj;
            const bool below = lane > j && in;
            u = (lane == j) ? uj : (below ? u - Ar[j] * uj : u);
            w = (lane == j) ? wj : (below ? w - Ar[j] * wj : w);
        });
        static_for<0, K>([&](auto jc) {
            constexpr int j = K - 1 - decltype(jc)::value;
            if (j >= k) return;
            const double rjj = lane_value(rd, j);
            const double uj = lane_value(u, j) * rjj, wj = lane_value(w, j) * rjj;
            u = (lane == j) ? uj : (lane < j ? u - Lt[j] * uj : u);
            w = (lane == j) ? wj : (lane < j ? w - Lt[j] * wj : w);
        });
        const double ui = in ? u : 0.0, wi = in ? w : 0.0;
        const double su = wave_sum32(ui), sw = wave_sum32(wi);
        const double dy = (res_p + sw) / su;
        const double dl = in ? ui * dy - wi : 0.0;
        const double dz = in ? 1.0 / l * (mu - l * z - z * dl) : 0.0;
        double ap = wave_max32(in ? -dl / l : -1e9);
        double ad = wave_max32(in ? -dz / z : -1e9);
        ap = (ap > 0) ? ((0.99 / ap < 1.0) ? 0.99 / ap : 1.0) : 1.0;
        ad = (ad > 0) ? ((0.99 / ad < 1.0) ? 0.99 / ad : 1.0) : 1.0;
        y += ad * dy;
        if (in) { l = l + ap * dl; z = z + ad * dz; lam[lane] = l; }
        __builtin_amdgcn_wave_barrier();
        res_p = 1 - wave_sum32(in ? l : 0.0);
        Ql = qlam();
        mu = wave_sum32(in ? l * z : 0.0) / k * 0.4;
        if (ap + ad > 1.8) mu *= 0.2;
        temp = wave_sum32(in ? l * Ql : 0.0);
        D_cost = y - 0.5 * temp;
        P_cost = wave_sum32(in ? l * ci : 0.0) + 0.5 * temp;
        gap = P_cost - D_cost;
        if (cnt > 30) break;
    }
    return 0;
}

// KM = the largest bundle the kernel holds: 64 (every case), or 16 -- then
// Q and the LDS scratch shrink (19 KB instead of 68 KB) and three
// workgroups fit on a CU.  Both variants do the same arithmetic in the same
// order.  split: the launch is one of a pair, the 16 variant taking the
// nodes with at most 16 elements and the 64 variant the others.
template <int KM>
__global__ __launch_bounds__(BT, (KM == KMAX) ? 2 : 3) void bundle_qp_kernel(BatchDims d, BatchPtrs p, int split)
{
    // s_A: the Gram pass's row tiles, the solver's factor, the wave partials
    constexpr int SA = (KM == KMAX) ? KMAX * KMAX : 2048;
    static_assert(SA >= KR * (KR + 1) || KM <= 16, "factor storage");
    __shared__ double s_Q[KM * KM];
    __shared__ double s_A[SA];
    __shared__ double s_zeta[KM], s_c[KM], s_lam[KM], s_u[KM], s_w[KM];
    __shared__ double s_red[BT / 64];
    __shared__ int s_done, s_stat;
    __shared__ double s_t;

    const int node = blockIdx.x;
    if (node >= d.nodes || !p.mask[node]) return;
    NodeState *sp = p.st + node;
    const NodeState s = *sp;
    const int m = s.NIneq + s.NPentIneq + s.NHeptaIneq;
    const int k = s.bundle;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (split && (KM == KMAX ? k <= 16 : k > KM)) return;   // the other launch's node
    if (k > KM || k < 1) {
        if (threadIdx.x == 0) bundle_full_giveup(sp, s, p.mask + node, d.certify);
        return;
    }
    const double *G = p.G + (size_t)node * d.max_bundle * d.max_m;
    const double *F = p.F + (size_t)node * d.max_bundle;
    const double *gam = p.gamma + (size_t)node * d.max_m;
    double *gt = p.gtest + (size_t)node * d.max_m;
    // G[l + j*m] through a buffer resource: byte 8*l [VGPR] + j*8*m [SGPR];
    // columns past the bundle read as 0
    const __amdgpu_buffer_rsrc_t Gr =
        __builtin_amdgcn_make_buffer_rsrc((void *)G, (short)0, 8 * m * k, 0x00020000);
    const int colb = 8 * m;
    // per-thread partial sums of G'eta for bundle elements past PE
    // ((k - PE) * BT <= tmp_ld doubles of this node's scratch)
    double *tmp = p.tmpm + (size_t)node * d.tmp_ld;
    double t = s.t;
    if (k > PE && (size_t)(k - PE) * BT > (size_t)d.tmp_ld) {
        if (threadIdx.x == 0) bundle_full_giveup(sp, s, p.mask + node, d.certify);
        return;
    }
    for (int j = PE; j < k; ++j) tmp[(size_t)(j - PE) * BT + threadIdx.x] = 0.0;

    // zeta = -F - G' gamma (bundle.c:57-63) and Q = t G'G (bundle.c:309-315)
    // in one pass over the rows of [G gamma], staged through LDS (s_A) in
    // tiles of R rows.  Each thread owns up to GP of the k(k+1)/2 + k dot
    // products and sums its rows in order l = 0..m-1.
    {
        const int kc = k + 1;                       // column k holds gamma
        const int RS = min(256 + 1, SA / kc);   // column stride (odd: bank spread)
        const int R = RS - 1;
        const int npq = k * (k + 1) / 2;
        const int np = npq + k;
        constexpr int GP = (KM * (KM + 1) / 2 + KM + BT - 1) / BT;
        double acc[GP];
        int pa[GP], pb[GP];
        static_for<0, GP>([&](auto sc) {
            constexpr int q = decltype(sc)::value;
            const int pp = threadIdx.x + q * BT;
            acc[q] = 0.0;
            int i = 0, j = 0;
            if (pp < npq) {                 // column-major lower triangle, i >= j
                int r = pp;
                while (r >= k - j) { r -= k - j; ++j; }
                i = j + r;
            } else if (pp < np) {
                i = pp - npq;
                j = k;
            }
            pa[q] = i * RS;
            pb[q] = j * RS;
        });
        for (int l0 = 0; l0 < m; l0 += R) {
            const int rows = min(R, m - l0);
            __syncthreads();
            for (int e = threadIdx.x; e < rows * kc; e += BT) {
                const int c = e / rows, l = e - c * rows;
                s_A[c * RS + l] = (c < k) ? G[l0 + l + (size_t)c * m] : gam[l0 + l];
            }
            __syncthreads();
            static_for<0, GP>([&](auto sc) {
                constexpr int q = decltype(sc)::value;
                if (threadIdx.x + q * BT >= np) return;
                const double *x = s_A + pa[q], *y = s_A + pb[q];
                double v = acc[q];
                for (int l = 0; l < rows; ++l) v += x[l] * y[l];
                acc[q] = v;
            });
        }
        static_for<0, GP>([&](auto sc) {
            constexpr int q = decltype(sc)::value;
            const int pp = threadIdx.x + q * BT;
            if (pp >= np) return;
            const int i = pa[q] / RS, j = pb[q] / RS;
            if (pp < npq) s_Q[i + j * k] = s_Q[j + i * k] = (m > 0) ? t * acc[q] : 0.0;
            // with no cuts dgemv returns early and leaves zeta = F
            else s_zeta[i] = (m > 0) ? -F[i] + (-1.0) * acc[q] : F[i];
        });
        __syncthreads();
    }
    if (threadIdx.x == 0) { s_done = 0; s_stat = 0; }
    // first round: eta = 0, so c = zeta - t G'0
    for (int j = threadIdx.x; j < k; j += BT) s_c[j] = (m > 0) ? s_zeta[j] + (-t) * 0.0 : s_zeta[j];
    __syncthreads();

    // lambda-eta loop (bundle.c:316-372).  One pass over G per round: each
    // row of G is loaded once and gives (G lambda)_l, eta_l, dgamma_l and
    // its share of G'eta for the next round's c.
    double dir_prev = 0.0;
    int cnt = 0;
    while (!s_done) {
        ++cnt;
        if (wave == 0) {
            int st;
            // the smallest register factor that holds k (the unrolled
            // updates of a larger one run over padding)
            if constexpr (KM <= 16)
                st = (k <= 8)    ? solve_lambda_reg<8>(k, s_Q, s_c, s_lam, s_A, lane)
                   : (k <= 12) ? solve_lambda_reg<12>(k, s_Q, s_c, s_lam, s_A, lane)
                               : solve_lambda_reg<16>(k, s_Q, s_c, s_lam, s_A, lane);
            else
                st = (k <= 8)    ? solve_lambda_reg<8>(k, s_Q, s_c, s_lam, s_A, lane)
                   : (k <= 12) ? solve_lambda_reg<12>(k, s_Q, s_c, s_lam, s_A, lane)
                   : (k <= 16) ? solve_lambda_reg<16>(k, s_Q, s_c, s_lam, s_A, lane)
                   : (k <= 20) ? solve_lambda_reg<20>(k, s_Q, s_c, s_lam, s_A, lane)
                   : (k <= 24) ? solve_lambda_reg<24>(k, s_Q, s_c, s_lam, s_A, lane)
                   : (k <= 28) ? solve_lambda_reg<28>(k, s_Q, s_c, s_lam, s_A, lane)
                   : (k <= KR) ? solve_lambda_reg<KR>(k, s_Q, s_c, s_lam, s_A, lane)
                               : solve_lambda_wave(k, s_Q, s_c, s_lam, s_A, s_u, s_w, lane);
            if (lane == 0 && st) s_stat = st;
        }
        __syncthreads();
        if (s_stat) break;
        double nrm = 0.0;
        double pe[PE];
        static_for<0, PE>([&](auto jc) { pe[decltype(jc)::value] = 0.0; });
        for (int l = threadIdx.x; l < m; l += BT) {
            // row l of G: the first PE columns are read once, in chunks of 8
            // while the chunk starts inside the bundle, and kept for G'eta
            const int vo = 8 * l;
            double g[PE];
            double acc = 0.0;
            static_for<0, PE / 8>([&](auto cc) {
                constexpr int c0 = 8 * decltype(cc)::value;
                if (c0 >= k) return;
                static_for<0, 8>([&](auto jc) {
                    constexpr int j = c0 + decltype(jc)::value;
                    g[j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(Gr, vo, j * colb, 0));
                });
                static_for<0, 8>([&](auto jc) {
                    constexpr int j = c0 + decltype(jc)::value;
                    if (j < k) acc += g[j] * s_lam[j];
                });
            });
            for (int j = PE; j < k; ++j)
                acc += __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(Gr, vo, j * colb, 0)) * s_lam[j];
            double e = -gam[l] / t + acc;
            e = (e > 0) ? e : 0;
            const double dgl = t * (e - acc);
            gt[l] = gam[l] + dgl;
            nrm += dgl * dgl;
            static_for<0, PE / 8>([&](auto cc) {
                constexpr int c0 = 8 * decltype(cc)::value;
                if (c0 >= k) return;
                static_for<0, 8>([&](auto jc) {
                    constexpr int j = c0 + decltype(jc)::value;
                    pe[j] += g[j] * e;
                });
            });
            // columns past PE are read again (from the cache)
            for (int j = PE; j < k; ++j)
                tmp[(size_t)(j - PE) * BT + threadIdx.x] +=
                    __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(Gr, vo, j * colb, 0)) * e;
        }
        nrm = bsum(nrm, s_red);
        const double dir = sqrt(nrm);
        if (threadIdx.x == 0) {
            int done = 0;
            if (fabs(dir - dir_prev) / (1.0 + dir) < 1e-5) done = 1;
            if (cnt >= 50) { done = 1; t *= 0.95; }
            s_done = done;
            s_t = t;
        }
        dir_prev = dir;
        __syncthreads();
        t = s_t;
        if (s_done) break;
        // c = zeta - t G'eta for the next round: each wave reduces its
        // partial sums of every column, then one thread per column adds the
        // wave results in order (the arithmetic of bsum, one barrier)
        static_for<0, PE>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j >= k) return;
            const double v = wsum(pe[j]);
            if (lane == 0) s_A[wave * KM + j] = v;
        });
        for (int j = PE; j < k; ++j) {
            const double v = wsum(tmp[(size_t)(j - PE) * BT + threadIdx.x]);
            if (lane == 0) s_A[wave * KM + j] = v;
        }
        __syncthreads();
        for (int j = threadIdx.x; j < k; j += BT) {
            double v = 0.0;
            for (int w = 0; w < BT / 64; ++w) v += s_A[w * KM + j];
            s_c[j] = (m > 0) ? s_zeta[j] + (-t) * v : s_zeta[j];
        }
        for (int j = PE; j < k; ++j) tmp[(size_t)(j - PE) * BT + threadIdx.x] = 0.0;
        __syncthreads();
    }
    if (s_stat) {
        if (threadIdx.x == 0) { sp->status = s_stat; sp->active = 0; p.mask[node] = 0; }
        return;
    }
    double *lam = p.lambda + (size_t)node * d.max_bundle;
    for (int j = threadIdx.x; j < k; j += BT) lam[j] = s_lam[j];
    if (threadIdx.x == 0) sp->t = t;
}

// Column / slab copy inside one node's bundle (memmove order preserved by
// the caller: destinations precede sources).
__device__ void bundle_move(const BatchDims &d, const BatchPtrs &p, int node, int m, int n, int from, int to)
{
    const size_t slab = (size_t)d.ld * d.ld;
    double *G = p.G + (size_t)node * d.max_bundle * d.max_m;
    double *Xb = p.Xb + (size_t)node * d.max_bundle * slab;
    for (int l = threadIdx.x; l < m; l += BT) G[(size_t)to * m + l] = G[(size_t)from * m + l];
    for (InteriorWalk w(n, threadIdx.x, BT); w.more(); w.next()) {
        const size_t e = w.off(d.ld);
        Xb[to * slab + e] = Xb[from * slab + e];
    }
    if (threadIdx.x == 0) p.F[(size_t)node * d.max_bundle + to] = p.F[(size_t)node * d.max_bundle + from];
    __syncthreads();
}

__global__ __launch_bounds__(BT) void bundle_step_kernel(BatchDims d, BatchPtrs p)
{
    __shared__ double s_red[BT / 64];
    __shared__ double s_lam[KMAX];
    __shared__ int s_keep[KMAX];
    const int node = blockIdx.x;
    if (node >= d.nodes || !p.mask[node]) return;
    NodeState *sp = p.st + node;
    const NodeState s = *sp;
    const int m = s.NIneq + s.NPentIneq + s.NHeptaIneq;
    const int k = s.bundle;
    const size_t slab = (size_t)d.ld * d.ld;
    double *G = p.G + (size_t)node * d.max_bundle * d.max_m;
    double *F = p.F + (size_t)node * d.max_bundle;
    const double *lam = p.lambda + (size_t)node * d.max_bundle;
    double *gam = p.gamma + (size_t)node * d.max_m;
    const double *gt = p.gtest + (size_t)node * d.max_m;
    double *g = p.g + (size_t)node * d.max_m;
    const double *Xt = p.Xt + node * slab;
    const int ipm_st = p.ipm_status[node];

    for (int j = threadIdx.x; j < k; j += BT) s_lam[j] = lam[j];
    // f_test = phi + sum(gamma_test), g = 1 - B(X_test)  (bundle.c:259-270)
    double sg = 0.0;
    for (int l = threadIdx.x; l < m; l += BT) sg += gt[l];
    sg = bsum(sg, s_red);
    const double f_test = p.phi[node] + (m > 0 ? sg : 0.0);
    {
        const TriCut *tri = p.tri + (size_t)node * d.max_tri;
        const KgonCut *pent = p.pent + (size_t)node * d.max_pent;
        const KgonCut *hept = p.hept + (size_t)node * d.max_hept;
        for (int c = threadIdx.x; c < m; c += BT)
            g[c] = 1.0 - cut_B(Xt, d.ld, c, s.NIneq, s.NPentIneq, tri, pent, hept);
    }
    // f_appr = <F + G' gamma_test, lambda>  (bundle.c:76-83)
    double fa = 0.0;
    for (int j = threadIdx.x >> 6; j < k; j += BT / 64) {
        double acc = 0.0;
        for (int l = threadIdx.x & 63; l < m; l += 64) acc += G[l + (size_t)j * m] * gt[l];
        acc = wsum(acc);
        if ((threadIdx.x & 63) == 0) fa += (F[j] + acc) * s_lam[j];
    }
    fa = bsum(fa, s_red);
    const double del = s.f - fa;
    double lmax = -1e9;
    for (int i = 0; i < k; ++i) lmax = (s_lam[i] > lmax) ? s_lam[i] : lmax;
    const int serious = (s.f - f_test > 0.05 * del) ? 1 : 0;
    if (ipm_st) {
        if (threadIdx.x == 0) { sp->status = ipm_st; sp->active = 0; p.mask[node] = 0; }
        return;
    }
    double t = s.t;
    double f = s.f;
    if (serious) {
        for (int l = threadIdx.x; l < m; l += BT) gam[l] = gt[l];
        f = f_test;
        // the new center's y, for its certificate at the round end
        if (d.certify) {
            const double *yt = p.ytest + (size_t)node * d.ld;
            double *yc = p.ycen + (size_t)node * d.ld;
            for (int i = threadIdx.x; i < s.n; i += BT) yc[i] = yt[i];
        }
        // X = X_bundle * lambda, each entry summed over the bundle in order
        double *X = p.X + node * slab;
        const double *Xb = p.Xb + (size_t)node * d.max_bundle * slab;
        for (InteriorWalk w(s.n, threadIdx.x, BT); w.more(); w.next()) {
            const size_t e = w.off(d.ld);
            double acc = 0.0;
            for (int j = 0; j < k; ++j) acc += Xb[j * slab + e] * s_lam[j];
            X[e] = acc;
        }
        t *= 1.01;
    } else {
        t /= 1.01;
    }
    __syncthreads();
    // purge (all k on a serious step, the first k-1 on a null step)
    const int kp = serious ? k : k - 1;
    int next = 0, subtracted = 0;
    for (int i = 0; i < kp; ++i) {
        if (s_lam[i] < 0.01 * lmax) ++subtracted;
        else {
            if (i > next) bundle_move(d, p, node, m, s.n, i, next);
            ++next;
        }
    }
    int bundle = k - subtracted;
    if (bundle >= d.max_bundle) {
        if (threadIdx.x == 0) bundle_full_giveup(sp, s, p.mask + node, d.certify);
        return;
    }
    int pos;
    if (serious) pos = bundle;
    else {
        bundle_move(d, p, node, m, s.n, k - 1, bundle);
        pos = bundle - 1;
    }
    for (int l = threadIdx.x; l < m; l += BT) G[(size_t)pos * m + l] = g[l];
    double *Xdst = p.Xb + ((size_t)node * d.max_bundle + pos) * slab;
    for (InteriorWalk w(s.n, threadIdx.x, BT); w.more(); w.next()) {
        const size_t e = w.off(d.ld);
        Xdst[e] = Xt[e];
    }
    // F = <L, X_test> (bundle.c:152-163)
    double acc = 0.0;
    const double *L = p.L + node * slab;
    for (int e = threadIdx.x; e < s.n * s.n; e += BT) {
        const int i = e % s.n, j = e / s.n;
        acc += L[i + (size_t)j * d.ld] * Xt[i + (size_t)j * d.ld];
    }
    acc = bsum(acc, s_red);
    if (threadIdx.x == 0) {
        F[pos] = acc;
        ++bundle;
        sp->bundle = bundle;
        if (bundle > s.max_bundle_seen) sp->max_bundle_seen = bundle;
        sp->t = t;
        sp->f = f;
        sp->f_test = f_test;
        sp->del = del;
        sp->lmax = lmax;
        sp->serious = serious;
        if (serious && d.certify) sp->cert_stale = 1;
        sp->ipm_calls = s.ipm_calls + 1;
        sp->ipm_iters = s.ipm_iters + p.ipm_iters[node];
        sp->bdl_left = s.bdl_left - 1;
        if (sp->bdl_left <= 0) sp->rphase = 2;
    }
    (void)s_keep;
}

}  // namespace bqb

// kbound: an upper bound on the bundle size of every node of the launch
// (0: unknown -- one launch of the 64 variant).  Above 16 the nodes are split
// between the two variants by their bundle size (two launches in order).
extern "C" hipError_t bqb_launch_bundle_qp(const BatchDims *d, const BatchPtrs *p, int kbound, hipStream_t s)
{
    using namespace bqb;
    if (kbound <= 0) {
        hipLaunchKernelGGL(bundle_qp_kernel<KMAX>, dim3(d->nodes), dim3(BT), 0, s, *d, *p, 0);
    } else if (kbound <= 16) {
        hipLaunchKernelGGL(bundle_qp_kernel<16>, dim3(d->nodes), dim3(BT), 0, s, *d, *p, 0);
    } else {
        hipLaunchKernelGGL(bundle_qp_kernel<16>, dim3(d->nodes), dim3(BT), 0, s, *d, *p, 1);
        hipLaunchKernelGGL(bundle_qp_kernel<KMAX>, dim3(d->nodes), dim3(BT), 0, s, *d, *p, 1);
    }
    return hipGetLastError();
}

extern "C" hipError_t bqb_launch_bundle_step(const BatchDims *d, const BatchPtrs *p, hipStream_t s)
{
    hipLaunchKernelGGL(bqb::bundle_step_kernel, dim3(d->nodes), dim3(bqb::BT), 0, s, *d, *p);
    return hipGetLastError();
}
