// bqb_learn.h -- the cut model: features of a cut and a 3 x 256 MLP on the
// bf16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation).
//
// The model maps the features of a cut (below) to two outputs: out[0], the
// warm-start multiplier in the form log1p(gamma / GAMMA_SCALE), and out[1],
// the logit of the cut being active (gamma > 1e-5) at the end of the node.
// Layers: 32 -> 256 -> 256 -> 256 (ReLU) -> 16 (outputs 0 and 1 used).
// Parameters in one flat FP32 array (LEARN_* offsets); the matrix cores read
// bf16 copies, W as [out][in] (the reduction contiguous) and its transpose
// for the backward pass.
//
// A workgroup of 256 threads (4 waves) handles 64 rows (cuts) at a time:
// activations stay in LDS as bf16 [64][K + 8], weights stream from global
// memory (L2-resident, 282 KB in bf16), wave w computes rows 16w..16w+15 of
// every output column block.
#pragma once
#include <hip/hip_runtime.h>

#include "bqb_batch.h"
#include "bqb_cutops.h"
#include "bqb_mixed.h"

namespace bqb {

constexpr int LEARN_NF = 32;       // input features (16 used, the rest zero)
constexpr int LEARN_H = 256;       // hidden width
constexpr int LEARN_NO = 16;       // output rows (2 used)
constexpr int LEARN_ROWS = 64;     // rows per workgroup tile
constexpr int LEARN_T = 256;       // threads per workgroup
constexpr double GAMMA_SCALE = 0.05;

// flat parameter layout
constexpr int LEARN_W1 = 0;
constexpr int LEARN_B1 = LEARN_W1 + LEARN_H * LEARN_NF;
constexpr int LEARN_W2 = LEARN_B1 + LEARN_H;
constexpr int LEARN_B2 = LEARN_W2 + LEARN_H * LEARN_H;
constexpr int LEARN_W3 = LEARN_B2 + LEARN_H;
constexpr int LEARN_B3 = LEARN_W3 + LEARN_H * LEARN_H;
constexpr int LEARN_W4 = LEARN_B3 + LEARN_H;
constexpr int LEARN_B4 = LEARN_W4 + LEARN_NO * LEARN_H;
constexpr int LEARN_NPARAM = LEARN_B4 + LEARN_NO;

// bf16 copies of the weights: W1 [256][32], W2, W3 [256][256], W4 [16][256],
// and transposes W2T, W3T [256][256], W4T [256][32] (rows 16..31 of the
// reduction zero) for the data gradients.
constexpr int LB_W1 = 0;
constexpr int LB_W2 = LB_W1 + LEARN_H * LEARN_NF;
constexpr int LB_W3 = LB_W2 + LEARN_H * LEARN_H;
constexpr int LB_W4 = LB_W3 + LEARN_H * LEARN_H;
constexpr int LB_W2T = LB_W4 + LEARN_NO * LEARN_H;
constexpr int LB_W3T = LB_W2T + LEARN_H * LEARN_H;
constexpr int LB_W4T = LB_W3T + LEARN_H * LEARN_H;
constexpr int LB_TOTAL = LB_W4T + LEARN_H * 32;

__device__ __forceinline__ unsigned short f2bf(float f)
{
    // round to nearest even
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

// out[64][N] = act(A[64][K] W'[K][N] + b): A in LDS (bf16, row stride lda),
// W global bf16 [N][K] (K contiguous), b global FP32 (or NULL).  The FP32
// result goes to outf (LDS or global, row stride ldf; may be NULL), the
// bf16 result to outb (row stride ldb; may be NULL).  relu: apply ReLU.
// K and N multiples of 32 and 16.
template <int K, int N>
__device__ __forceinline__ void mlp_layer(const unsigned short *A, int lda, const unsigned short *W, const float *b,
                                          bool relu, float *outf, int ldf, unsigned short *outb, int ldb)
{
    constexpr int NB = N / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, kq = 8 * (lane >> 4);
    f32x4 acc[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const unsigned short *arow = A + (size_t)(16 * wave + r16) * lda + kq;
#pragma unroll 2
    for (int k0 = 0; k0 < K; k0 += 32) {
        const bf16x8 a = __builtin_bit_cast(bf16x8, *(const uint4 *)(arow + k0));
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const bf16x8 w = __builtin_bit_cast(bf16x8, *(const uint4 *)(W + (size_t)(16 * j + r16) * K + kq + k0));
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, acc[j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int c = 16 * j + r16;
        const float bias = b ? b[c] : 0.f;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 16 * wave + 4 * (lane >> 4) + v;
            float x = acc[j][v] + bias;
            if (relu) x = fmaxf(x, 0.f);
            if (outf) outf[(size_t)r * ldf + c] = x;
            if (outb) outb[(size_t)r * ldb + c] = f2bf(x);
        }
    }
}

// ---------------------------------------------------------------- features
// Features of cut c of a node (c indexes [tri | pent | hept]), computed from
// the node's X and L (column-major, ld) -- the same formulas as
// parallel_biqbin_maxcut_amd/learn.py:cut_features:
//  0 B_c(X)            1 B_c(X) - 1        2-4 one-hot k = 3, 5, 7
//  5 (type - 1) / k    6 mean |X_ab|       7 min s_ab X_ab   8 max s_ab X_ab
//  9 mean s_ab L_ab / lscale               10 mean |L_ab| / lscale
// 11 n / 128          12 round / 20       13 cuts / (10 n)  14 relative gap
// 15 1, 16-31 zero.   s_ab = -e_a e_b (the cut's sign pattern), lscale the
// mean |L_ij| over the node's off-diagonal entries.
struct CutCtx {
    const double *X, *L;
    int ld, n, ntri, npent;
    const TriCut *tri;
    const KgonCut *pent, *hept;
    float lscale, round, cuts, gap;
};

__device__ __forceinline__ void cut_features(const CutCtx &cx, int c, float *f)
{
    int p[7], k, type;
    if (c < cx.ntri) {
        const TriCut t = cx.tri[c];
        p[0] = t.i; p[1] = t.j; p[2] = t.k; k = 3; type = t.type;
    } else if (c < cx.ntri + cx.npent) {
        const KgonCut t = cx.pent[c - cx.ntri];
        for (int a = 0; a < 5; ++a) p[a] = t.p[a];
        k = 5; type = t.type;
    } else {
        const KgonCut t = cx.hept[c - cx.ntri - cx.npent];
        for (int a = 0; a < 7; ++a) p[a] = t.p[a];
        k = 7; type = t.type;
    }
    double sx = 0.0, ax = 0.0, mn = 1e300, mx = -1e300, sl = 0.0, al = 0.0;
    int np = 0;
    for (int a = 0; a < k; ++a) {
        const int ea = (k == 3) ? tri_e(type, a) : kgon_e(type, a);
        for (int b = a + 1; b < k; ++b) {
            const int eb = (k == 3) ? tri_e(type, b) : kgon_e(type, b);
            const double s = (double)(-ea * eb);
            const size_t o = (size_t)p[a] + (size_t)p[b] * cx.ld;
            const double x = cx.X[o], l = cx.L[o];
            sx += s * x;
            ax += fabs(x);
            mn = fmin(mn, s * x);
            mx = fmax(mx, s * x);
            sl += s * l;
            al += fabs(l);
            ++np;
        }
    }
    const double coef = (k == 3) ? 1.0 : (k == 5) ? 0.5 : 1.0 / 3.0;
    const double bx = coef * sx;
    f[0] = (float)bx;
    f[1] = (float)(bx - 1.0);
    f[2] = k == 3; f[3] = k == 5; f[4] = k == 7;
    f[5] = (float)(type - 1) / (float)k;
    f[6] = (float)(ax / np);
    f[7] = (float)mn;
    f[8] = (float)mx;
    f[9] = (float)(sl / np) / cx.lscale;
    f[10] = (float)(al / np) / cx.lscale;
    f[11] = (float)cx.n / 128.f;
    f[12] = cx.round / 20.f;
    f[13] = cx.cuts / (10.f * (float)cx.n);
    f[14] = cx.gap;
    f[15] = 1.f;
    for (int q = 16; q < LEARN_NF; ++q) f[q] = 0.f;
}

}  // namespace bqb
