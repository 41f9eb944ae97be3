// bqb_dist.h -- workgroup-distributed dense FP64 matrices for gfx950.
//
// One workgroup of P*P threads owns one n x n matrix (n <= P*R).  Thread
// (tr, tc) = (tid % P, tid / P) holds the R x R register tile of elements
// (tr + P*a, tc + P*b), a, b < R (2-D cyclic layout).  Entries outside n are
// zero padding and stay zero through every operation below (the pivot loops
// stop at n, and a zero row/column is never touched by a rank-1 update).
//
// A wave holds 64 consecutive tids, i.e. 64/P grid columns; column sums are
// therefore intra-wave (shuffles) and row sums cross waves (LDS).
//
// The sequential algorithms (symmetric sweep inversion, Cholesky test)
// broadcast pivot column k through LDS; the register column index of the
// pivot is a compile-time constant (static_for over column blocks), so the
// register tiles are never indexed dynamically.  All loops that contain a barrier are uniform across the
// workgroup.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace bqb {

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

template <int P, int R>
struct Dist {
    static constexpr int NT = P * P;        // threads
    static constexpr int NMAX = P * R;      // largest n
    static constexpr int NW = NT / 64;      // waves
    static constexpr int CPW = 64 / P;      // grid columns per wave
};

// LDS scratch needed by one workgroup (doubles).
template <int P, int R, int KB>
struct Lds {
    static constexpr int NMAX = P * R;
    static constexpr int COL = 2 * NMAX;                 // double-buffered pivot column
    static constexpr int PANEL = 2 * 3 * KB * NMAX;      // double-buffered A/X/W GEMM panels
    static constexpr int TRANS = (P == 16 ? 2 : 1) * NMAX * (P + 1);   // symmetrize staging
    static constexpr int RED = (P * P / 64) * NMAX;      // row-reduction partials
    static constexpr int UNION = PANEL > TRANS ? (PANEL > RED ? PANEL : RED) : (TRANS > RED ? TRANS : RED);
};

struct Thr {
    int tid, tr, tc, lane, wave;
};

template <int P>
__device__ __forceinline__ Thr thr_id()
{
    Thr t;
    t.tid = threadIdx.x;
    t.tr = t.tid % P;
    t.tc = t.tid / P;
    t.lane = t.tid & 63;
    t.wave = t.tid >> 6;
    return t;
}

// ---------------------------------------------------------------- load/store
// Matrices live in zero-padded column-major slabs with leading dimension
// ld >= P*R, so every tile element is a valid address and padding loads as 0.
// Tiles are read/written through a buffer resource per slab: element
// (tr + P*a, tc + P*b) is at byte  vo + 8*P*a [VGPR + immediate]
// + b*8*P*ld [SGPR],  vo = 8*(tr + tc*ld), so one VGPR addresses the whole
// tile.  Accesses past the slab (num_records) read 0 and drop writes.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

struct Slab {
    __amdgpu_buffer_rsrc_t r;
    int vo;       // 8*(tr + tc*ld)
    int colstep;  // 8*P*ld
};

template <int P>
__device__ __forceinline__ Slab make_slab(const double *p, int ld, const Thr &t)
{
    Slab s;
    s.r = __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, ld * ld * 8, 0x00020000);
    s.vo = 8 * (t.tr + t.tc * ld);
    s.colstep = 8 * P * ld;
    return s;
}

template <int P>
__device__ __forceinline__ double sld(const Slab &s, int a, int b)
{
    u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(s.r, s.vo + 8 * P * a, b * s.colstep, 0);
    return __builtin_bit_cast(double, v);
}

template <int P>
__device__ __forceinline__ void sst(const Slab &s, int a, int b, double x)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), s.r, s.vo + 8 * P * a,
                                          b * s.colstep, 0);
}

template <int P, int R>
__device__ __forceinline__ void dload(double (&A)[R][R], const Slab &s)
{
#pragma unroll
    for (int b = 0; b < R; ++b)
#pragma unroll
        for (int a = 0; a < R; ++a) A[a][b] = sld<P>(s, a, b);
}

template <int P, int R>
__device__ __forceinline__ void dstore(const double (&A)[R][R], const Slab &s)
{
#pragma unroll
    for (int b = 0; b < R; ++b)
#pragma unroll
        for (int a = 0; a < R; ++a) sst<P>(s, a, b, A[a][b]);
}

// ---------------------------------------------------------------- reductions
// Sum of v over the whole workgroup, identical in every thread.  Fixed order.
template <int P>
__device__ __forceinline__ double wg_sum(double v, double *red, const Thr &t)
{
    constexpr int NW = P * P / 64;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if (t.lane == 0) red[t.wave] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += red[w];
    return s;
}

// Row sums: part[a] holds this thread's partial for row tr + P*a.  The total
// for every row I < NMAX is written to out[I] (LDS).  Fixed order.
template <int P, int R>
__device__ __forceinline__ void row_sums(const double (&part)[R], double *red, double *out,
                                         const Thr &t)
{
    constexpr int NW = P * P / 64;
    constexpr int CPW = 64 / P;
    constexpr int NMAX = P * R;
    double v[R];
#pragma unroll
    for (int a = 0; a < R; ++a) {
        v[a] = part[a];
#pragma unroll
        for (int o = P; o < 64; o <<= 1) v[a] += __shfl_xor(v[a], o, 64);
    }
    __syncthreads();
    if (t.lane < P) {
#pragma unroll
        for (int a = 0; a < R; ++a) red[t.wave * NMAX + t.tr + P * a] = v[a];
    }
    __syncthreads();
    for (int I = t.tid; I < NMAX; I += P * P) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) s += red[w * NMAX + I];
        out[I] = s;
    }
    __syncthreads();
    (void)CPW;
}

// ---------------------------------------------------------------- pivots
// 1/d for a pivot d > 0: the hardware reciprocal estimate refined by two
// Newton steps y <- y + y (1 - d y) (the error squares with each step), so
// the result is within an ulp of the rounded quotient at about half the
// instructions and latency of the IEEE division sequence.  A pivot that is
// not positive fails the factorization, so its reciprocal is never used.
__device__ __forceinline__ double pivot_recip(double d)
{
    double y = __builtin_amdgcn_rcp(d);
    double e = fma(-d, y, 1.0);
    y = fma(y, e, y);
    e = fma(-d, y, 1.0);
    return fma(y, e, y);
}

// ---------------------------------------------------------------- sweep inverse
// Copies the strictly lower register blocks onto the strictly upper ones:
// element (I, J) of a block with b > a takes element (J, I), staged through
// T (NMAX x (P+1) doubles).  Diagonal blocks are left as they are.
template <int P, int R>
__device__ __forceinline__ void mirror_lower(double (&A)[R][R], double *T, const Thr &t)
{
    constexpr int LD = P + 1;
    static_for<0, R - 1>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        if (r > 0) __syncthreads();   // T of the previous block column is read
#pragma unroll
        for (int a = r + 1; a < R; ++a) T[(t.tr + P * a) * LD + t.tc] = A[a][r];
        __syncthreads();
#pragma unroll
        for (int b = r + 1; b < R; ++b) A[r][b] = T[(t.tc + P * b) * LD + t.tr];
    });
    __syncthreads();
}

// In-place inverse of an SPD matrix by the symmetric sweep operator
// (Beaton / Goodnight): pivot k maps a_ij -> a_ij - a_ik (a_kj / a_kk),
// a_ik -> a_ik / a_kk, a_kk -> -1 / a_kk; after every pivot the matrix holds
// -inv(A), which is negated at the end.  The operator keeps the matrix
// symmetric, so only the register blocks with b <= a are swept: column k is
// assembled from the owners of its lower part (rows in blocks >= k's block)
// and of row k (columns in blocks < k's block), and the strictly upper blocks
// are mirrored from the lower ones at the end (T: NMAX x (P+1) doubles).
// Returns false (uniformly) if a pivot is not positive, i.e. A is not positive
// definite -- the same test as dpotrf.
template <int P, int R>
__device__ __forceinline__ bool sweep_inverse(double (&A)[R][R], int n, double *colbuf, double *T,
                                              const Thr &t)
{
    constexpr int NMAX = P * R;
    int buf = 0;
    __syncthreads();   // colbuf may still be read by the previous pivot loop
    bool ok = true;
    static_for<0, R>([&](auto kbc) {
        constexpr int kb = decltype(kbc)::value;
        for (int kt = 0; kt < P && ok; ++kt) {
            const int k = kb * P + kt;
            if (k >= n) break;
            double *cb = colbuf + buf * NMAX;
            if (t.tc == kt) {
#pragma unroll
                for (int a = kb; a < R; ++a) cb[t.tr + P * a] = A[a][kb];
            }
            if (t.tr == kt) {
#pragma unroll
                for (int b = 0; b < kb; ++b) cb[t.tc + P * b] = A[kb][b];
            }
            __syncthreads();
            const double d = cb[k];
            double cI[R], cJ[R];
#pragma unroll
            for (int a = 0; a < R; ++a) cI[a] = cb[t.tr + P * a];
#pragma unroll
            for (int b = 0; b < R; ++b) cJ[b] = cb[t.tc + P * b];
            // the pivot test ends the step, so the LDS reads and the
            // divide are not held back by it (a failed step's updates are
            // discarded with the matrix)
            const bool good = d > 0.0;
            const double dinv = pivot_recip(d);
            double tJ[R];
#pragma unroll
            for (int b = 0; b < R; ++b) tJ[b] = cJ[b] * dinv;
#pragma unroll
            for (int a = 0; a < R; ++a)
#pragma unroll
                for (int b = 0; b <= a; ++b) A[a][b] = fma(-cI[a], tJ[b], A[a][b]);
            if (t.tc == kt) {
#pragma unroll
                for (int a = kb; a < R; ++a) A[a][kb] = cI[a] * dinv;
            }
            if (t.tr == kt) {
#pragma unroll
                for (int b = 0; b <= kb; ++b) A[kb][b] = tJ[b];
            }
            if (t.tr == kt && t.tc == kt) A[kb][kb] = -dinv;
            if (!good) { ok = false; break; }
            buf ^= 1;
        }
    });
    if (!ok) return false;
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
        for (int b = 0; b <= a; ++b) A[a][b] = -A[a][b];
    mirror_lower<P, R>(A, T, t);
    return true;
}

// Positive-definiteness test by right-looking Cholesky (destroys A): the
// trailing matrix is updated with c_i (c_j / d).  Same decision rule as
// dpotrf: fails at the first pivot that is not > 0.
template <int P, int R>
__device__ __forceinline__ bool chol_test(double (&A)[R][R], int n, double *colbuf, const Thr &t)
{
    constexpr int NMAX = P * R;
    int buf = 0;
    __syncthreads();   // colbuf may still be read by the previous pivot loop
    bool ok = true;
    static_for<0, R>([&](auto kbc) {
        constexpr int kb = decltype(kbc)::value;
        for (int kt = 0; kt < P && ok; ++kt) {
            const int k = kb * P + kt;
            if (k >= n) break;
            double *cb = colbuf + buf * NMAX;
            if (t.tc == kt) {
#pragma unroll
                for (int a = kb; a < R; ++a) cb[t.tr + P * a] = A[a][kb];
            }
            __syncthreads();
            const double d = cb[k];
            double cI[R], cJ[R];
#pragma unroll
            for (int a = kb; a < R; ++a) cI[a] = cb[t.tr + P * a];
#pragma unroll
            for (int b = kb; b < R; ++b) cJ[b] = cb[t.tc + P * b];
            // the pivot test ends the step, so the LDS reads and the
            // divide are not held back by it (a failed step's updates are
            // discarded with the matrix)
            const bool good = d > 0.0;
            const double dinv = pivot_recip(d);
            double tJ[R];
#pragma unroll
            for (int b = kb; b < R; ++b) tJ[b] = cJ[b] * dinv;
            // lower blocks only: the pivots and the next columns are read
            // from blocks with a >= b, which see exactly the full update
#pragma unroll
            for (int a = kb; a < R; ++a)
#pragma unroll
                for (int b = kb; b <= a; ++b) A[a][b] = fma(-cI[a], tJ[b], A[a][b]);
            if (!good) { ok = false; break; }
            buf ^= 1;
        }
    });
    return ok;
}

// Cholesky factorisation A = L L' (lower) in place; the strictly upper part
// is set to zero.  Returns false (uniformly) at the first non-positive pivot.
template <int P, int R>
__device__ __forceinline__ bool chol_lower(double (&A)[R][R], int n, double *colbuf, const Thr &t)
{
    constexpr int NMAX = P * R;
    int buf = 0;
    __syncthreads();   // colbuf may still be read by the previous pivot loop
    bool ok = true;
    static_for<0, R>([&](auto kbc) {
        constexpr int kb = decltype(kbc)::value;
        for (int kt = 0; kt < P && ok; ++kt) {
            const int k = kb * P + kt;
            if (k >= n) break;
            double *cb = colbuf + buf * NMAX;
            if (t.tc == kt) {
#pragma unroll
                for (int a = kb; a < R; ++a) cb[t.tr + P * a] = A[a][kb];
            }
            __syncthreads();
            const double d = cb[k];
            double cI[R], cJ[R];
#pragma unroll
            for (int a = kb; a < R; ++a) cI[a] = cb[t.tr + P * a];
#pragma unroll
            for (int b = kb; b < R; ++b) cJ[b] = cb[t.tc + P * b];
            // the pivot test ends the step, so the LDS reads and the
            // divide are not held back by it (a failed step's updates are
            // discarded with the matrix)
            const bool good = d > 0.0;
            const double sq = sqrt(d);
            const double s = 1.0 / sq;
            double sJ[R];
#pragma unroll
            for (int b = kb; b < R; ++b) sJ[b] = cJ[b] * s;
#pragma unroll
            for (int a = kb; a < R; ++a) {
                const int I = t.tr + P * a;
                const double sI = cI[a] * s;
#pragma unroll
                for (int b = kb; b <= a; ++b) {   // the upper blocks end as zeros
                    const int J = t.tc + P * b;
                    if (I > k && J > k) A[a][b] = fma(-sI, sJ[b], A[a][b]);
                }
            }
            if (t.tc == kt) {
#pragma unroll
                for (int a = kb; a < R; ++a) {
                    const int I = t.tr + P * a;
                    if (I >= k) A[a][kb] = (I == k) ? sq : cI[a] * s;
                }
            }
            if (!good) { ok = false; break; }
            buf ^= 1;
        }
    });
    if (!ok) return false;
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
        for (int b = 0; b < R; ++b)
            if (t.tc + P * b > t.tr + P * a) A[a][b] = 0.0;
    return true;
}

// ---------------------------------------------------------------- symmetrize
// A <- (A + A')/2 using LDS staging of register column blocks: two, r and
// r + 1, per round on the 16 x 16 grid (T holds 2 NMAX (P+1) doubles), one
// on the 32 x 32 grid (NMAX (P+1)).  Element (r,b) with b < r takes the
// partner's value, which the partner already averaged in an earlier round;
// (r,b) with b >= r averages the two originals.  With two blocks per round,
// row block r + 1 does the same except at b = r: that partner is averaged
// in this round, so both sides average the originals (IEEE addition
// commutes: the same value on both sides).
template <int P, int R>
__device__ __forceinline__ void symmetrize(double (&A)[R][R], double *T, const Thr &t)
{
    constexpr int LD = P + 1;
    constexpr int TS = P * R * LD;
    constexpr int STEP = (P == 16) ? 2 : 1;
    static_for<0, (R + STEP - 1) / STEP>([&](auto rc) {
        constexpr int r0 = STEP * decltype(rc)::value;
        constexpr int r1 = (STEP == 2) ? r0 + 1 : R;
        __syncthreads();
#pragma unroll
        for (int a = 0; a < R; ++a) T[(t.tr + P * a) * LD + t.tc] = A[a][r0];
        if constexpr (r1 < R) {
#pragma unroll
            for (int a = 0; a < R; ++a) T[TS + (t.tr + P * a) * LD + t.tc] = A[a][r1];
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < R; ++b) {
            const double v = T[(t.tc + P * b) * LD + t.tr];
            if (b < r0) A[r0][b] = v;
            else A[r0][b] = 0.5 * (A[r0][b] + v);
        }
        if constexpr (r1 < R) {
#pragma unroll
            for (int b = 0; b < R; ++b) {
                const double v = T[TS + (t.tc + P * b) * LD + t.tr];
                if (b < r0) A[r1][b] = v;
                else A[r1][b] = 0.5 * (A[r1][b] + v);
            }
        }
    });
    __syncthreads();
}

}  // namespace bqb
