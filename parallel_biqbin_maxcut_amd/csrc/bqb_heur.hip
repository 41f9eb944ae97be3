// bqb_heur.hip -- primal heuristic of a batch of nodes (reference
// src/heuristic.c:19-317, src/bab_functions.c:96-134):
//   Z = X; repeat { Z = chol_lower(Z); x = best of n0 random-hyperplane cuts
//   sign(Z v), each improved by 1-opt; if it beats the best so far, restart
//   from Z = (1-c) X + c xh xh' } until two rounds in a row do not improve;
// then the node's incumbent is updated with the last round's best cut.
// Random numbers come from the node's counter streams: hyperplane h of
// round r draws v_j = U(key(heur key, r+1), h*1024 + j) - 1/2.
// The Cholesky factor uses the workgroup-distributed tile; the hyperplane
// products are formed per (hyperplane, row) in the reference's order without
// fused multiply-add; 1-opt and the cut values run on integer-valued data
// and are exact.
#include <hip/hip_runtime.h>

#include "bqb_batch.h"
#include "bqb_dist.h"
#include "bqb_kernels.h"
#include "bqb_rng.h"
#include "bqb_wave.h"

namespace bqb {

constexpr int HEUR_MAX_ROUNDS = 1000;
constexpr int NMAX_GW = 256;      // vertices (n0) supported
constexpr int GW_BATCH = 16;      // hyperplanes held in LDS at a time

// argmax with the first index on ties across a wave
__device__ __forceinline__ void wave_argmax(double &v, int &i) { wave_argmax64(v, i); }

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// FL: LDS copies as float (dynamic shared memory), read whenever every
// copied entry converts exactly (integer weights give quarter-integers), so
// the 1-opt and the cut evaluations read LDS instead of HBM/L2 with the same
// arithmetic.  FL = 1: the original objective SP_L and the node's L; FL = 2:
// L only (for integer SP_L, whose cut values then come from the 1-opt's Lx
// and which is read from memory for the two full evaluations per node).
template <int P, int R, int FL>
__global__ __launch_bounds__(P * P) void heur_kernel(BatchDims d, BatchPtrs p)
{
    extern __shared__ float s_f[];     // FL 1: [n0*n0] SP_L, then [n*n] L; FL 2: [n*n] L
    constexpr int NT = P * P;
    constexpr int NW = NT / 64;
    constexpr int NMAX = P * R;
    constexpr int PL = (NMAX + 63) / 64;
    __shared__ double s_col[2 * NMAX];
    __shared__ double s_xh[NMAX];
    __shared__ signed char s_sign[GW_BATCH][NMAX];
    __shared__ double s_v[GW_BATCH][NMAX];        // this batch's hyperplanes
    __shared__ signed char s_lift[GW_BATCH][NMAX_GW];
    __shared__ int s_sol[NMAX_GW];
    __shared__ double s_val[GW_BATCH];
    __shared__ int s_free[NMAX_GW];
    __shared__ int s_best_sign[NMAX];
    __shared__ double s_red[NW];
    __shared__ double s_best;
    __shared__ int s_flag;
    __shared__ int s_exact;
    __shared__ int s_quarter;

    const int node = blockIdx.x;
    if (node >= d.nodes || !p.mask[node]) return;
    NodeState *sp = p.st + node;
    const NodeState s = *sp;
    const Thr t = thr_id<P>();
    const int n = s.n, ld = d.ld, n0 = d.n0, B = n0 - 1;
    const size_t slab = (size_t)ld * ld;
    const double *X = p.X + node * slab;
    const double *L = p.L + node * slab;
    double *Zs = p.Zh + node * slab;
    const int *xf = p.xfixed + (size_t)node * B;
    const int *sx = p.solX + (size_t)node * B;
    int *xo = p.xheur + (size_t)node * B;
    const uint64_t key = BQB_KEY_HEUR(s.seed, s.heur_idx);
    const int lane = t.lane, wave = t.wave;
    if (n > NMAX || n0 > NMAX_GW || n < 1) {
        if (t.tid == 0) { sp->status = ERR_SIZE; sp->active = 0; p.mask[node] = 0; }
        return;
    }
    if (t.tid == 0) {
        int nf = 0;
        for (int i = 0; i < B; ++i)
            if (!xf[i]) s_free[nf++] = i;
        s_exact = FL ? 1 : 0;
        s_quarter = FL ? 1 : 0;
    }
    float *const s_spl = s_f, *const s_lf = s_f + (FL == 1 ? n0 * n0 : 0);
    if (FL) {
        __syncthreads();
        // bad: an entry that float does not hold exactly; frac: one that is
        // not a multiple of 1/4 of modest size (then sums over the entries
        // are not all exact)
        int bad = 0, frac = 0;
        const auto quarter = [](double v) { return 4.0 * v == rint(4.0 * v) && fabs(v) <= 1048576.0; };
        for (int e = t.tid; e < n0 * n0; e += NT) {
            const double v = p.SP_L[e];
            const float f = (float)v;
            if (FL == 1) bad |= ((double)f != v);
            frac |= !quarter(v);
            if (FL == 1) s_spl[e] = f;
        }
        for (int e = t.tid; e < n * n; e += NT) {
            const int i = e % n, j = e / n;
            const double v = L[i + (size_t)j * ld];
            const float f = (float)v;
            bad |= ((double)f != v);
            frac |= !quarter(v);
            s_lf[e] = f;
        }
        if (bad) s_exact = 0;
        if (frac) s_quarter = 0;
    }
    // use_f is read from s_exact after the barrier below
    bool use_f = false;
    const auto spl = [&](int j, int i) -> double {
        return (FL == 1 && use_f) ? (double)s_spl[j + i * n0] : p.SP_L[j + (size_t)i * n0];
    };
    const auto splx = [&](auto uf, int j, int i) -> double {
        if constexpr (decltype(uf)::value && FL == 1) return (double)s_spl[j + i * n0];
        else return p.SP_L[j + (size_t)i * n0];
    };
    const auto lvx = [&](auto uf, int i, int j) -> double {
        if constexpr (decltype(uf)::value) return (double)s_lf[i + j * n];
        else return L[i + (size_t)j * ld];
    };
    const auto lv0 = [&](int i, int j) -> double { return lvx(std::true_type{}, i, j); };
    // first random cut xh in {-1,1}^n (heuristic.c:38-39)
    for (int i = t.tid; i < NMAX; i += NT)
        s_xh[i] = (i < n) ? (double)(2 * (int)(bqb_draw(key, (uint64_t)i) >> 63) - 1) : 0.0;
    __syncthreads();
    use_f = FL != 0 && s_exact;
    // quarter-integer data: every sum below is exact, so a cut's value is
    // the quadratic form of the subproblem plus a per-node constant (below)
    const bool use_q = use_f && s_quarter;

    // fh = value of its lift (heuristic.c:42-54)
    if (t.tid == 0) {
        int index = 0;
        for (int i = 0; i < B; ++i) {
            if (xf[i]) s_sol[i] = sx[i];
            else { s_sol[i] = (int)((s_xh[index] + 1) / 2.0); ++index; }
        }
    }
    __syncthreads();
    double fh;
    {
        double acc = 0.0;
        for (int i = wave; i < B; i += NW) {
            if (!s_sol[i]) continue;
            for (int j = lane; j < B; j += 64)
                if (s_sol[j]) acc += spl(j, i);
        }
        fh = wg_sum<P>(acc, s_red, t);
    }
    // The lift of x (free vertex a takes (x_a + 1) / 2) has the value
    // y'Ly + cq, where y is x with y_{n-1} = +1 (createSubproblem folds the
    // fixed vertices into the last row and column of L and the constant
    // cq).  With exact sums cq is read off the first cut: cq = fh - y'Ly.
    // Flipping coordinate n-1 of x changes the form by -4 x_{n-1} (Lx)_{n-1}
    // + 4 L(n-1, n-1).
    double cq = 0.0;
    if (use_q) {
        double acc = 0.0, accl = 0.0;
        for (int i = wave; i < n; i += NW) {
            double r = 0.0;
            for (int j = lane; j < n; j += 64) r += lv0(j, i) * s_xh[j];
            r = wave_sum(r);
            acc += s_xh[i] * r;
            if (i == n - 1) accl = r;
        }
        if (lane != 0) acc = accl = 0.0;
        double q = wg_sum<P>(acc, s_red, t);
        const double lxl = wg_sum<P>(accl, s_red, t);
        if (s_xh[n - 1] < 0) q = q + 4.0 * lxl + 4.0 * lv0(n - 1, n - 1);
        cq = fh - q;
    }

    double heur_val = -1e9;
    int done = 0, round = 0, status = 0;
    double A[R][R];
    const Slab sX = make_slab<P>(X, ld, t);
    const Slab sZ = make_slab<P>(Zs, ld, t);
    dload<P, R>(A, sX);                               // Z = X
    while (done < 2) {
        ++done;
        if (round >= HEUR_MAX_ROUNDS) break;
        const uint64_t rkey = bqb_key(key, (uint64_t)round + 1);
        ++round;
        if (!chol_lower<P, R>(A, n, s_col, t)) { status = ERR_HEUR_CHOL; break; }
        dstore<P, R>(A, sZ);
        __syncthreads();

        // ---- GW_heuristic (heuristic.c:119-186): n0 hyperplanes ----------
        for (int h0 = 0; h0 < n0; h0 += GW_BATCH) {
            const int hb = min(GW_BATCH, n0 - h0);
            // the batch's random vectors v_j = U - 1/2, drawn once
            for (int e = t.tid; e < hb * n; e += NT) {
                const int hh = e / n, j = e - hh * n;
                s_v[hh][j] = bqb_u01(bqb_draw(rkey, (uint64_t)(h0 + hh) * BQB_GW_STRIDE + j)) - 0.5;
            }
            __syncthreads();
            // sign(Z v) per (hyperplane, row), products summed in j order
            // without fused multiply-add (as the reference's scalar loop); a
            // thread takes one row for HG hyperplanes, so each entry of Z
            // is loaded once for all of them (rows past hb are discarded)
            {
                constexpr int HG = 4;
                static_assert(GW_BATCH % HG == 0, "hyperplane groups");
                const int ng = (hb + HG - 1) / HG;
                for (int e = t.tid; e < ng * n; e += NT) {
                    const int g = e / n, i = e - g * n;
                    double sca[HG];
#pragma unroll
                    for (int h = 0; h < HG; ++h) sca[h] = 0.0;
#pragma unroll 4
                    for (int j = 0; j <= i; ++j) {
                        const double z = Zs[i + (size_t)j * ld];
#pragma unroll
                        for (int h = 0; h < HG; ++h) sca[h] = __dadd_rn(sca[h], __dmul_rn(s_v[g * HG + h][j], z));
                    }
#pragma unroll
                    for (int h = 0; h < HG; ++h)
                        if (g * HG + h < hb) s_sign[g * HG + h][i] = (sca[h] < 0) ? -1 : 1;
                }
            }
            __syncthreads();
            // 1-opt and cut values with the L / SP_L storage fixed at compile
            // time, so the loads of the loops are scheduled together
            const auto opt_eval = [&](auto uf) {
                // 1-opt (heuristic.c:195-296), one wave per hyperplane: the cut
                // stays in LDS, Lx in registers (row i = lane + 64 q)
                for (int hh = wave; hh < hb; hh += NW) {
                    signed char *xcur = s_sign[hh];
                    // Lx_i = sum_j L(i, j) x_j in j order; L is symmetric, so
                    // column j is read (lanes on consecutive rows)
                    double Lx[PL];
#pragma unroll
                    for (int q = 0; q < PL; ++q) Lx[q] = 0.0;
                    for (int j = 0; j < n; ++j) {
                        const double xj = (double)xcur[j];
#pragma unroll
                        for (int q = 0; q < PL; ++q) {
                            const int i = lane + 64 * q;
                            if (i < n) Lx[q] += lvx(uf, i, j) * xj;
                        }
                    }
                    // the diagonal and the cut stay in registers for the flips
                    double dg[PL], xr[PL];
#pragma unroll
                    for (int q = 0; q < PL; ++q) {
                        const int i = lane + 64 * q;
                        dg[q] = (i < n) ? lvx(uf, i, i) : 0.0;
                        xr[q] = (i < n) ? (double)xcur[i] : 0.0;
                    }
                    double bv = -1e9;
                    int bi = 0x7fffffff;
#pragma unroll
                    for (int q = 0; q < PL; ++q) {
                        const int i = lane + 64 * q;
                        if (i < n) {
                            const double dl = dg[q] - xr[q] * Lx[q];
                            if (dl > bv || (dl == bv && i < bi)) { bv = dl; bi = i; }
                        }
                    }
                    wave_argmax(bv, bi);
                    int flips = 0;
                    while (bv > 0.001 && flips < 100 * n + 100) {
                        ++flips;
                        const int index = bi;
                        const int xi = xcur[index];
#pragma unroll
                        for (int q = 0; q < PL; ++q) {
                            const int i = lane + 64 * q;
                            if (i < n) {
                                const double lij = lvx(uf, i, index);   // = L(index, i)
                                if (fabs(lij) > 0.001) Lx[q] = (xi > 0) ? Lx[q] - 2 * lij : Lx[q] + 2 * lij;
                            }
                        }
                        __builtin_amdgcn_wave_barrier();
                        if (lane == (index & 63)) xcur[index] = (signed char)(-xi);
                        __builtin_amdgcn_wave_barrier();
#pragma unroll
                        for (int q = 0; q < PL; ++q)
                            if (lane + 64 * q == index) xr[q] = -xr[q];
                        bv = -1e9;
                        bi = 0x7fffffff;
#pragma unroll
                        for (int q = 0; q < PL; ++q) {
                            const int i = lane + 64 * q;
                            if (i < n) {
                                const double dl = dg[q] - xr[q] * Lx[q];
                                if (dl > bv || (dl == bv && i < bi)) { bv = dl; bi = i; }
                            }
                        }
                        wave_argmax(bv, bi);
                    }
                    double acc = 0.0;
                    if (use_q) {
                        // the quadratic form from the 1-opt's Lx (exact sums)
                        double tl = 0.0;
#pragma unroll
                        for (int q = 0; q < PL; ++q) {
                            const int i = lane + 64 * q;
                            if (i < n) acc += (double)xcur[i] * Lx[q];
                            if (i == n - 1) tl = Lx[q];
                        }
                        acc = wave_sum(acc);
                        tl = wave_sum(tl);
                        if (xcur[n - 1] < 0) acc = acc + 4.0 * tl + 4.0 * lvx(uf, n - 1, n - 1);
                        acc = acc + cq;
                    } else {
                        // lift to the original vertices (fixed ones keep their value)
                        for (int i = lane; i < B; i += 64) s_lift[hh][i] = xf[i] ? (signed char)sx[i] : (signed char)0;
                        __builtin_amdgcn_wave_barrier();
                        for (int a = lane; a < n - 1; a += 64) s_lift[hh][s_free[a]] = (signed char)((xcur[a] + 1) / 2);
                        __builtin_amdgcn_wave_barrier();
                        // its cut value (update_best -> evaluateSolution): the sum of
                        // SP_L over the pairs with sol_i = sol_j = 1, lanes over j and
                        // the 0/1 values as factors; integer data, so the sum is exact
                        // whatever the order
                        {
                            double yj[(NMAX_GW + 63) / 64];
#pragma unroll
                            for (int q = 0; q < (NMAX_GW + 63) / 64; ++q) {
                                const int j = lane + 64 * q;
                                yj[q] = (j < B) ? (double)s_lift[hh][j] : 0.0;
                            }
#pragma unroll 4
                            for (int i = 0; i < B; ++i) {
                                const double yi = (double)s_lift[hh][i];
                                double r = 0.0;
#pragma unroll
                                for (int q = 0; q < (NMAX_GW + 63) / 64; ++q) {
                                    const int j = lane + 64 * q;
                                    if (64 * q < B) r += (j < B) ? yj[q] * splx(uf, j, i) : 0.0;
                                }
                                acc += yi * r;
                            }
                        }
                        acc = wave_sum(acc);
                    }
                    if (lane == 0) s_val[hh] = acc;
                }
            };
            if (FL != 0 && use_f) opt_eval(std::true_type{});
            else opt_eval(std::false_type{});
            __syncthreads();
            // update_best: first strictly better hyperplane in order
            if (t.tid == 0) {
                double bb = (h0 == 0) ? -1e9 : s_best;
                int bh = -1;
                for (int hh = 0; hh < hb; ++hh)
                    if (bb < s_val[hh]) { bb = s_val[hh]; bh = hh; }
                s_best = bb;
                s_flag = bh;
            }
            __syncthreads();
            if (s_flag >= 0)
                for (int i = t.tid; i < n; i += NT) s_best_sign[i] = s_sign[s_flag][i];
            __syncthreads();
        }
        heur_val = s_best;
        // x = lift of this round's best cut
        if (t.tid == 0) {
            int index = 0;
            for (int i = 0; i < B; ++i) {
                if (xf[i]) xo[i] = sx[i];
                else { xo[i] = (s_best_sign[index] + 1) / 2; ++index; }
            }
        }
        __syncthreads();
        if (heur_val > fh) {
            done = 0;
            fh = heur_val;
            if (t.tid == 0) {
                int index = 0;
                for (int i = 0; i < B; ++i)
                    if (!xf[i]) { s_xh[index] = 2 * xo[i] - 1; ++index; }
                s_xh[n - 1] = -1.0;
            }
            __syncthreads();
        }
        const double c = 0.3 + 0.6 * bqb_u01(bqb_draw(rkey, (uint64_t)n0 * BQB_GW_STRIDE));
        const double alpha = 1.0 - c;
        // Z = (1-c) X, lower triangle += c xh xh' (dscal + dsyr 'L')
        dload<P, R>(A, sX);
#pragma unroll
        for (int a = 0; a < R; ++a)
#pragma unroll
            for (int b = 0; b < R; ++b) {
                const int I = t.tr + P * a, J = t.tc + P * b;
                double z = alpha * A[a][b];
                if (I >= J && I < n) z = z + c * s_xh[J] * s_xh[I];
                A[a][b] = z;
            }
    }

    // updateSolution (bab_functions.c:114-134) against the node's incumbent
    __syncthreads();
    if (status) {
        if (t.tid == 0) { sp->status = status; sp->active = 0; p.mask[node] = 0; }
        return;
    }
    for (int i = t.tid; i < B; i += NT) s_sol[i] = xo[i];
    __syncthreads();
    double acc = 0.0;
    for (int i = wave; i < B; i += NW) {
        if (!s_sol[i]) continue;
        for (int j = lane; j < B; j += 64)
            if (s_sol[j]) acc += spl(j, i);
    }
    acc = wg_sum<P>(acc, s_red, t);
    if (acc > s.LB) {
        int *xb = p.xbest + (size_t)node * B;
        for (int i = t.tid; i < B; i += NT) xb[i] = xo[i];
        if (t.tid == 0) { sp->LB = acc; sp->improved = 1; }
    }
}

template <int P, int R>
static hipError_t launch_heur(const BatchDims &d, const BatchPtrs &p, int spl_integer, hipStream_t s)
{
    // float copies of L (and of SP_L unless it is integer-valued; n <= n0)
    // while they fit beside the static LDS
    static const size_t static_lds = []() {
        hipFuncAttributes a;
        return (hipFuncGetAttributes(&a, (const void *)heur_kernel<P, R, 1>) == hipSuccess)
                   ? a.sharedSizeBytes : (size_t)160 * 1024;
    }();
    const size_t one = sizeof(float) * (size_t)d.n0 * d.n0;
    if (spl_integer && static_lds + one <= (size_t)160 * 1024)
        hipLaunchKernelGGL((heur_kernel<P, R, 2>), dim3(d.nodes), dim3(P * P), one, s, d, p);
    else if (static_lds + 2 * one <= (size_t)160 * 1024)
        hipLaunchKernelGGL((heur_kernel<P, R, 1>), dim3(d.nodes), dim3(P * P), 2 * one, s, d, p);
    else
        hipLaunchKernelGGL((heur_kernel<P, R, 0>), dim3(d.nodes), dim3(P * P), 0, s, d, p);
    return hipGetLastError();
}

}  // namespace bqb

// spl_integer: every entry of SP_L is an integer of magnitude <= 2^20
extern "C" hipError_t bqb_launch_heur(const BatchDims *d, const BatchPtrs *p, int tile, int spl_integer,
                                      hipStream_t s)
{
    using namespace bqb;
    switch (tile) {
    case 1604: return launch_heur<16, 4>(*d, *p, spl_integer, s);
    case 1605: return launch_heur<16, 5>(*d, *p, spl_integer, s);
    case 1606: return launch_heur<16, 6>(*d, *p, spl_integer, s);
    case 1607: return launch_heur<16, 7>(*d, *p, spl_integer, s);
    case 3204: return launch_heur<32, 4>(*d, *p, spl_integer, s);
    case 3205: return launch_heur<32, 5>(*d, *p, spl_integer, s);
    case 3206: return launch_heur<32, 6>(*d, *p, spl_integer, s);
    case 3207: return launch_heur<32, 7>(*d, *p, spl_integer, s);
    case 3208: return launch_heur<32, 8>(*d, *p, spl_integer, s);
    default: return hipErrorInvalidValue;
    }
}
