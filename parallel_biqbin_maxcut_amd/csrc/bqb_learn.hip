// bqb_learn.hip -- training and inference of the cut model (bqb_learn.h) on
// the bf16 matrix cores, and its use inside the node bound.
//
// Training step on a mini-batch of M rows (features -> targets):
//   mlp_fwd_kernel   rows in tiles of 64: H1, H2, H3 (bf16) and the outputs
//   mlp_bwd_kernel   rows in tiles of 64: loss, dOut, and the data gradients
//                    dY3 = (dOut W4) o [H3 > 0], dY2, dY1 (bf16)
//   mlp_wgrad_kernel dW = dY' H over the batch, 64 x 64 output blocks, the
//                    batch staged through LDS transposed (deterministic)
//   mlp_bgrad_kernel db = column sums of dY
//   mlp_adam_kernel  Adam on the FP32 master parameters
//   mlp_pack_kernel  bf16 copies and transposes for the next step
// Loss per row: (out0 - t0)^2 + BCE(sigmoid(out1), t1), averaged over M.
//
// In the node bound (learn_cuts_kernel, after every separation): the cuts a
// separation just appended (multiplier exactly 0; every survivor of a purge
// has y >= 1e-5, cutting_planes.c:151, 336, 529) get the model's warm-start
// multiplier -- the predicted magnitude times the predicted probability of
// the cut being active -- instead of 0 (cutting_planes.c:201, 392, 588), and
// the ones it predicts active with probability below a
// threshold are dropped in order (the ranking that stands beside the
// top-K-by-violation list, cutting_planes.c:74-116).  Any multiplier >= 0
// still gives a valid bound; the cold-start bundle judges the result.  In
// record mode the kernel instead logs the features of the new cuts (a
// quarter of them, chosen by a hash of the cut) for training.
#include <hip/hip_runtime.h>

#include <stdlib.h>
#include <string.h>

#include "bqb_learn.h"
#include "bqb_rng.h"
#include "bqb_runtime.h"

namespace bqb {

static_assert(LEARN_NPARAM == BQB_MLP_NPARAM, "parameter layout of the C ABI");

constexpr int LDA_F = LEARN_NF + 8;    // LDS row strides (bf16 elements)
constexpr int LDA_H = LEARN_H + 8;

// ---------------------------------------------------------------- forward
__global__ __launch_bounds__(LEARN_T) void mlp_fwd_kernel(int M, const int *rows, const unsigned short *feat,
                                                          const unsigned short *wb, const float *prm,
                                                          unsigned short *H1, unsigned short *H2,
                                                          unsigned short *H3, float *out)
{
    __shared__ __align__(16) unsigned short sF[LEARN_ROWS * LDA_F];
    __shared__ __align__(16) unsigned short sA[LEARN_ROWS * LDA_H];
    __shared__ __align__(16) unsigned short sB[LEARN_ROWS * LDA_H];
    __shared__ float sO[LEARN_ROWS * LEARN_NO];
    const int r0 = blockIdx.x * LEARN_ROWS;
    for (int e = threadIdx.x; e < LEARN_ROWS * LEARN_NF; e += LEARN_T) {
        const int r = e / LEARN_NF, c = e % LEARN_NF;
        const int g = r0 + r;
        sF[r * LDA_F + c] = (g < M) ? feat[(size_t)(rows ? rows[g] : g) * LEARN_NF + c] : 0;
    }
    __syncthreads();
    mlp_layer<LEARN_NF, LEARN_H>(sF, LDA_F, wb + LB_W1, prm + LEARN_B1, true, nullptr, 0, sA, LDA_H);
    __syncthreads();
    mlp_layer<LEARN_H, LEARN_H>(sA, LDA_H, wb + LB_W2, prm + LEARN_B2, true, nullptr, 0, sB, LDA_H);
    __syncthreads();
    if (H1)
        for (int e = threadIdx.x; e < LEARN_ROWS * LEARN_H; e += LEARN_T) {
            const int r = e / LEARN_H, c = e % LEARN_H;
            if (r0 + r < M) H1[(size_t)(r0 + r) * LEARN_H + c] = sA[r * LDA_H + c];
        }
    __syncthreads();
    mlp_layer<LEARN_H, LEARN_H>(sB, LDA_H, wb + LB_W3, prm + LEARN_B3, true, nullptr, 0, sA, LDA_H);
    __syncthreads();
    mlp_layer<LEARN_H, LEARN_NO>(sA, LDA_H, wb + LB_W4, prm + LEARN_B4, false, sO, LEARN_NO, nullptr, 0);
    __syncthreads();
    for (int e = threadIdx.x; e < LEARN_ROWS * LEARN_H; e += LEARN_T) {
        const int r = e / LEARN_H, c = e % LEARN_H;
        if (r0 + r >= M) continue;
        if (H2) H2[(size_t)(r0 + r) * LEARN_H + c] = sB[r * LDA_H + c];
        if (H3) H3[(size_t)(r0 + r) * LEARN_H + c] = sA[r * LDA_H + c];
    }
    for (int e = threadIdx.x; e < LEARN_ROWS * LEARN_NO; e += LEARN_T) {
        const int r = e / LEARN_NO;
        if (r0 + r < M) out[(size_t)(r0 + r) * LEARN_NO + e % LEARN_NO] = sO[e];
    }
}

// ---------------------------------------------------------------- backward
// dY4 [M][16], dY3/dY2/dY1 [M][256] (bf16, the pre-activation gradients);
// loss[blockIdx] receives the block's summed row losses.
__global__ __launch_bounds__(LEARN_T) void mlp_bwd_kernel(int M, const int *rows, const float *target,
                                                          const unsigned short *wb, const unsigned short *H1,
                                                          const unsigned short *H2, const unsigned short *H3,
                                                          const float *out, unsigned short *dY4,
                                                          unsigned short *dY3, unsigned short *dY2,
                                                          unsigned short *dY1, double *loss)
{
    __shared__ __align__(16) unsigned short sD[LEARN_ROWS * LDA_F];
    __shared__ __align__(16) unsigned short sA[LEARN_ROWS * LDA_H];
    __shared__ __align__(16) unsigned short sB[LEARN_ROWS * LDA_H];
    __shared__ double sl[LEARN_T / 64];
    const int r0 = blockIdx.x * LEARN_ROWS;
    const float invM = 1.f / (float)M;
    double lsum = 0.0;
    for (int e = threadIdx.x; e < LEARN_ROWS * LEARN_NF; e += LEARN_T) {
        const int r = e / LEARN_NF, c = e % LEARN_NF;
        const int g = r0 + r;
        float d = 0.f;
        if (g < M && c < 2) {
            const int src = rows ? rows[g] : g;
            const float o = out[(size_t)g * LEARN_NO + c];
            const float t = target[(size_t)src * 2 + c];
            if (c == 0) {
                d = 2.f * (o - t) * invM;
                lsum += (double)(o - t) * (o - t);
            } else {
                const float s = 1.f / (1.f + expf(-o));
                d = (s - t) * invM;
                // BCE in a stable form: max(o, 0) - o t + log(1 + exp(-|o|))
                lsum += (double)(fmaxf(o, 0.f) - o * t + log1pf(expf(-fabsf(o))));
            }
        }
        const unsigned short db = f2bf(d);
        sD[r * LDA_F + c] = db;
        if (g < M && c < LEARN_NO) dY4[(size_t)g * LEARN_NO + c] = db;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
    if ((threadIdx.x & 63) == 0) sl[threadIdx.x >> 6] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) loss[blockIdx.x] = sl[0] + sl[1] + sl[2] + sl[3];
    // dY3 = (dOut W4) o [H3 > 0]
    mlp_layer<32, LEARN_H>(sD, LDA_F, wb + LB_W4T, nullptr, false, nullptr, 0, sA, LDA_H);
    __syncthreads();
    for (int e = threadIdx.x; e < LEARN_ROWS * LEARN_H; e += LEARN_T) {
        const int r = e / LEARN_H, c = e % LEARN_H;
        const int g = r0 + r;
        unsigned short v = 0;
        if (g < M && H3[(size_t)g * LEARN_H + c] != 0) v = sA[r * LDA_H + c];
        sA[r * LDA_H + c] = v;
        if (g < M) dY3[(size_t)g * LEARN_H + c] = v;
    }
    __syncthreads();
    mlp_layer<LEARN_H, LEARN_H>(sA, LDA_H, wb + LB_W3T, nullptr, false, nullptr, 0, sB, LDA_H);
    __syncthreads();
    for (int e = threadIdx.x; e < LEARN_ROWS * LEARN_H; e += LEARN_T) {
        const int r = e / LEARN_H, c = e % LEARN_H;
        const int g = r0 + r;
        unsigned short v = 0;
        if (g < M && H2[(size_t)g * LEARN_H + c] != 0) v = sB[r * LDA_H + c];
        sB[r * LDA_H + c] = v;
        if (g < M) dY2[(size_t)g * LEARN_H + c] = v;
    }
    __syncthreads();
    mlp_layer<LEARN_H, LEARN_H>(sB, LDA_H, wb + LB_W2T, nullptr, false, nullptr, 0, sA, LDA_H);
    __syncthreads();
    for (int e = threadIdx.x; e < LEARN_ROWS * LEARN_H; e += LEARN_T) {
        const int r = e / LEARN_H, c = e % LEARN_H;
        const int g = r0 + r;
        if (g < M) dY1[(size_t)g * LEARN_H + c] = (H1[(size_t)g * LEARN_H + c] != 0) ? sA[r * LDA_H + c] : 0;
    }
}

// dW[N][K] = sum_m dY[m][j] X[m][k] over the batch (X rows through `rows`
// when X is the feature matrix).  Block (bj, bk): 64 x 64 outputs; 32 rows
// of the batch staged at a time, transposed into LDS.
__global__ __launch_bounds__(LEARN_T) void mlp_wgrad_kernel(int M, int N, int K, const unsigned short *dY, int ldy,
                                                            const unsigned short *X, int ldx, const int *xrows,
                                                            float *dW)
{
    __shared__ __align__(16) unsigned short sY[64 * 40];   // [j][m]
    __shared__ __align__(16) unsigned short sX[64 * 40];   // [k][m]
    const int j0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, kq = 8 * (lane >> 4);
    f32x4 acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int m0 = 0; m0 < M; m0 += 32) {
        for (int e = threadIdx.x; e < 32 * 64; e += LEARN_T) {
            const int m = e / 64, c = e % 64;
            const int g = m0 + m;
            unsigned short y = 0, x = 0;
            if (g < M) {
                if (j0 + c < N) y = dY[(size_t)g * ldy + j0 + c];
                if (k0 + c < K) x = X[(size_t)(xrows ? xrows[g] : g) * ldx + k0 + c];
            }
            sY[c * 40 + m] = y;
            sX[c * 40 + m] = x;
        }
        __syncthreads();
        const bf16x8 a = __builtin_bit_cast(bf16x8, *(const uint4 *)(sY + (16 * wave + r16) * 40 + kq));
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bf16x8 b = __builtin_bit_cast(bf16x8, *(const uint4 *)(sX + (16 * q + r16) * 40 + kq));
            acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[q], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = k0 + 16 * q + r16;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = j0 + 16 * wave + 4 * (lane >> 4) + v;
            if (j < N && k < K) dW[(size_t)j * K + k] = acc[q][v];
        }
    }
}

__global__ void mlp_bgrad_kernel(int M, int N, const unsigned short *dY, int ldy, float *db)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; ++m) s += bf2f(dY[(size_t)m * ldy + j]);
    db[j] = s;
}

// Adam (Kingma & Ba) on the FP32 parameters; the output rows 2..15 of W4 and
// b4 are not used and stay zero.
__global__ void mlp_adam_kernel(float *prm, const float *g, float *m1, float *m2, int count, float lr, float b1,
                                float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    if (i >= LEARN_W4 && i < LEARN_B4 && (i - LEARN_W4) / LEARN_H >= 2) return;
    if (i >= LEARN_B4 + 2) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    prm[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__global__ void mlp_pack_kernel(const float *prm, unsigned short *wb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < LEARN_H * LEARN_NF) wb[LB_W1 + i] = f2bf(prm[LEARN_W1 + i]);
    if (i < LEARN_H * LEARN_H) {
        const int j = i / LEARN_H, k = i % LEARN_H;
        const unsigned short w2 = f2bf(prm[LEARN_W2 + i]), w3 = f2bf(prm[LEARN_W3 + i]);
        wb[LB_W2 + i] = w2;
        wb[LB_W3 + i] = w3;
        wb[LB_W2T + k * LEARN_H + j] = w2;
        wb[LB_W3T + k * LEARN_H + j] = w3;
    }
    if (i < LEARN_NO * LEARN_H) {
        const int j = i / LEARN_H, k = i % LEARN_H;
        const unsigned short w4 = f2bf(prm[LEARN_W4 + i]);
        wb[LB_W4 + i] = w4;
        wb[LB_W4T + k * 32 + j] = w4;
    }
    if (i < LEARN_H * 16) {          // columns 16..31 of W4T: zero
        const int k = i / 16, j = 16 + i % 16;
        wb[LB_W4T + k * 32 + j] = 0;
    }
}

// ---------------------------------------------------------------- features
// Diagnostic: the features of explicit cuts (m of them) of one node.
__global__ __launch_bounds__(256) void cut_features_kernel(CutCtx cx, int m, float *out)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < m; c += gridDim.x * blockDim.x)
        cut_features(cx, c, out + (size_t)c * LEARN_NF);
}

// ---------------------------------------------------------------- in the bound
// One workgroup per node.  rec (record mode): features and keys of a quarter
// of the new cuts are appended to the log; else the model predicts and acts.
struct LearnArgs {
    const unsigned short *wb;
    const float *prm;
    float drop;                  // drop new cuts with P(active) below this
    int *rec_count;              // record mode: log size counter (NULL: predict)
    int rec_cap;
    float *rec_feat;             // [rec_cap][16]
    int *rec_key;                // [rec_cap][12]: k, type, p[7], round, seed lo, seed hi
    int *idx;                    // [nodes][max_m] scratch: new-cut indices
    unsigned char *keep;         // [nodes][max_m] scratch
    int after_round;             // 0: nodes of the phase mask (first separation);
                                 // 1: nodes at the end of an outer round that separated
};

__device__ __forceinline__ uint32_t cut_hash(int k, int type, const int *p)
{
    uint32_t h = 2166136261u ^ (uint32_t)(k * 16 + type);
    for (int a = 0; a < k; ++a) h = (h ^ (uint32_t)p[a]) * 16777619u;
    return h;
}

__global__ __launch_bounds__(LEARN_T) void learn_cuts_kernel(BatchDims d, BatchPtrs p, LearnArgs la)
{
    __shared__ __align__(16) unsigned short sF[LEARN_ROWS * LDA_F];
    __shared__ __align__(16) unsigned short sA[LEARN_ROWS * LDA_H];
    __shared__ __align__(16) unsigned short sB[LEARN_ROWS * LDA_H];
    __shared__ float sO[LEARN_ROWS * LEARN_NO];
    __shared__ double red[LEARN_T / 64];
    __shared__ int s_cnt[LEARN_T], s_nnew;
    const int node = blockIdx.x;
    if (node >= d.nodes) return;
    NodeState *sp = p.st + node;
    const NodeState s = *sp;
    if (la.after_round) {
        if (s.stage != BQB_STAGE_LOOP || s.rphase != 2 || !s.sep_tri) return;
    } else if (!p.mask[node]) {
        return;
    }
    if (!s.active || s.status) return;
    const int n = s.n;
    const int mt = s.NIneq, mp = s.NPentIneq, mh = s.NHeptaIneq, m = mt + mp + mh;
    const size_t slab = (size_t)d.ld * d.ld;
    double *ty = p.tri_y + (size_t)node * d.max_tri;
    double *py = p.pent_y + (size_t)node * d.max_pent;
    double *hy = p.hept_y + (size_t)node * d.max_hept;
    // mean |L_ij| over the off-diagonal
    double al = 0.0;
    for (int e = threadIdx.x; e < n * n; e += LEARN_T) {
        const int i = e % n, j = e / n;
        if (i != j) al += fabs(p.L[node * slab + i + (size_t)j * d.ld]);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) al += __shfl_xor(al, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = al;
    __syncthreads();
    const double lsum = red[0] + red[1] + red[2] + red[3];
    CutCtx cx;
    cx.X = p.X + node * slab;
    cx.L = p.L + node * slab;
    cx.ld = d.ld;
    cx.n = n;
    cx.ntri = mt;
    cx.npent = mp;
    cx.tri = p.tri + (size_t)node * d.max_tri;
    cx.pent = p.pent + (size_t)node * d.max_pent;
    cx.hept = p.hept + (size_t)node * d.max_hept;
    cx.lscale = (float)(n > 1 ? lsum / ((double)n * (n - 1)) : 1.0) + 1e-12f;
    cx.round = (float)s.count;
    cx.cuts = (float)m;
    // the uncertified bound (certification changes only the prune tests)
    const double plain_bound = s.f + s.fixedvalue;
    cx.gap = (float)((plain_bound - s.LB) / (fabs(plain_bound) + 1.0));
    // the new cuts, in cut order (ordered compaction of y == 0)
    int *idx = la.idx + (size_t)node * d.max_m;
    unsigned char *keep = la.keep + (size_t)node * d.max_m;
    int nnew = 0;
    for (int base = 0; base < m; base += LEARN_T) {
        const int c = base + threadIdx.x;
        int isnew = 0;
        if (c < m) {
            const double y = (c < mt) ? ty[c] : (c < mt + mp) ? py[c - mt] : hy[c - mt - mp];
            isnew = (y == 0.0);
        }
        s_cnt[threadIdx.x] = isnew;
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int i = 0; i < LEARN_T; ++i) { const int v = s_cnt[i]; s_cnt[i] = acc; acc += v; }
            s_nnew = acc;
        }
        __syncthreads();
        if (isnew) idx[nnew + s_cnt[threadIdx.x]] = c;
        nnew += s_nnew;
        __syncthreads();
    }
    // the new cuts follow the survivors of each list's purge (and stay
    // behind them when some are dropped below): the rebuild that evaluates
    // the warm start reads where they begin
    if (threadIdx.x == 0) {
        int ft = mt, fp = mt + mp, fh = m;
        for (int q = nnew - 1; q >= 0; --q) {
            const int c = idx[q];
            if (c < mt) ft = c;
            else if (c < mt + mp) fp = c;
            else fh = c;
        }
        sp->nsurv_tri = ft;
        sp->nsurv_pent = fp - mt;
        sp->nsurv_hept = fh - mt - mp;
    }
    if (la.rec_count) {
        // record a quarter of the new cuts (by hash of the cut)
        for (int q = threadIdx.x; q < nnew; q += LEARN_T) {
            const int c = idx[q];
            int k, type, pp[7] = {0, 0, 0, 0, 0, 0, 0};
            if (c < mt) { const TriCut t = cx.tri[c]; k = 3; type = t.type; pp[0] = t.i; pp[1] = t.j; pp[2] = t.k; }
            else if (c < mt + mp) { const KgonCut t = cx.pent[c - mt]; k = 5; type = t.type; for (int a = 0; a < 5; ++a) pp[a] = t.p[a]; }
            else { const KgonCut t = cx.hept[c - mt - mp]; k = 7; type = t.type; for (int a = 0; a < 7; ++a) pp[a] = t.p[a]; }
            if ((cut_hash(k, type, pp) ^ (uint32_t)s.seed) & 3u) continue;
            const int pos = atomicAdd(la.rec_count, 1);
            if (pos >= la.rec_cap) continue;
            float f[LEARN_NF];
            cut_features(cx, c, f);
            for (int a = 0; a < 16; ++a) la.rec_feat[(size_t)pos * 16 + a] = f[a];
            int *kk = la.rec_key + (size_t)pos * 12;
            kk[0] = k; kk[1] = type;
            for (int a = 0; a < 7; ++a) kk[2 + a] = pp[a];
            kk[9] = s.count;
            kk[10] = (int)(uint32_t)s.seed;
            kk[11] = (int)(uint32_t)(s.seed >> 32);
        }
        return;
    }
    // predict in tiles of 64 new cuts
    for (int t0 = 0; t0 < nnew; t0 += LEARN_ROWS) {
        if (threadIdx.x < LEARN_ROWS) {
            float f[LEARN_NF];
            const int q = t0 + threadIdx.x;
            if (q < nnew) cut_features(cx, idx[q], f);
            else for (int a = 0; a < LEARN_NF; ++a) f[a] = 0.f;
            for (int a = 0; a < LEARN_NF; ++a) sF[threadIdx.x * LDA_F + a] = f2bf(f[a]);
        }
        __syncthreads();
        mlp_layer<LEARN_NF, LEARN_H>(sF, LDA_F, la.wb + LB_W1, la.prm + LEARN_B1, true, nullptr, 0, sA, LDA_H);
        __syncthreads();
        mlp_layer<LEARN_H, LEARN_H>(sA, LDA_H, la.wb + LB_W2, la.prm + LEARN_B2, true, nullptr, 0, sB, LDA_H);
        __syncthreads();
        mlp_layer<LEARN_H, LEARN_H>(sB, LDA_H, la.wb + LB_W3, la.prm + LEARN_B3, true, nullptr, 0, sA, LDA_H);
        __syncthreads();
        mlp_layer<LEARN_H, LEARN_NO>(sA, LDA_H, la.wb + LB_W4, la.prm + LEARN_B4, false, sO, LEARN_NO, nullptr, 0);
        __syncthreads();
        if (threadIdx.x < LEARN_ROWS && t0 + threadIdx.x < nnew) {
            const int c = idx[t0 + threadIdx.x];
            const float o0 = sO[threadIdx.x * LEARN_NO], o1 = sO[threadIdx.x * LEARN_NO + 1];
            const float pa = 1.f / (1.f + expf(-o1));
            // expected multiplier: P(active) times the predicted magnitude
            const double g = (double)pa * GAMMA_SCALE * expm1((double)fminf(fmaxf(o0, 0.f), 6.f));
            if (c < mt) ty[c] = g;
            else if (c < mt + mp) py[c - mt] = g;
            else hy[c - mt - mp] = g;
            keep[c] = (pa >= la.drop) ? 1 : 0;
        }
        __syncthreads();
    }
    // drop the new cuts ranked out, in order, per cut type
    if (la.drop > 0.f && nnew > 0) {
        if (threadIdx.x == 0) {
            int w = 0, q = 0;
            TriCut *tc = p.tri + (size_t)node * d.max_tri;
            for (int c = 0; c < mt; ++c) {
                const bool isnew = q < nnew && idx[q] == c;
                if (isnew) ++q;
                if (isnew && !keep[c]) continue;
                tc[w] = tc[c]; ty[w] = ty[c]; ++w;
            }
            const int nt = w;
            KgonCut *pc = p.pent + (size_t)node * d.max_pent;
            w = 0;
            for (int c = mt; c < mt + mp; ++c) {
                const bool isnew = q < nnew && idx[q] == c;
                if (isnew) ++q;
                if (isnew && !keep[c]) continue;
                pc[w] = pc[c - mt]; py[w] = py[c - mt]; ++w;
            }
            const int np = w;
            KgonCut *hc = p.hept + (size_t)node * d.max_hept;
            w = 0;
            for (int c = mt + mp; c < m; ++c) {
                const bool isnew = q < nnew && idx[q] == c;
                if (isnew) ++q;
                if (isnew && !keep[c]) continue;
                hc[w] = hc[c - mt - mp]; hy[w] = hy[c - mt - mp]; ++w;
            }
            sp->NIneq = nt;
            sp->NPentIneq = np;
            sp->NHeptaIneq = w;
            sp->cuts_changed = 1;
        }
    }
}

}  // namespace bqb

using namespace bqb;

// ================================================================ host API
struct bqb_mlp {
    float *prm, *grad, *m1, *m2;        // device FP32 [NPARAM]
    unsigned short *wb;                 // device bf16 copies [LB_TOTAL]
    int step;
    // training buffers, grown on demand
    int cap;                            // rows of the activation buffers (one batch)
    unsigned short *H1, *H2, *H3, *dY4, *dY3, *dY2, *dY1;
    float *out;
    double *loss;
    int *rows;                          // [M] batch order over the data set
    // data set
    int M;
    unsigned short *feat;               // [M][32] bf16
    float *target;                      // [M][2]
};

static int ck(hipError_t e, const char *w) { return bqb_check(e, w); }

static int mlp_pack(bqb_mlp *m)
{
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((LEARN_H * LEARN_H + 255) / 256), dim3(256), 0, 0, m->prm, m->wb);
    return ck(hipGetLastError(), "mlp pack");
}

extern "C" void bqb_mlp_destroy(bqb_mlp *m)
{
    if (!m) return;
    void *ptrs[] = {m->prm, m->grad, m->m1, m->m2, m->wb, m->H1, m->H2, m->H3, m->dY4, m->dY3, m->dY2, m->dY1,
                    m->out, m->loss, m->rows, m->feat, m->target};
    for (void *q : ptrs)
        if (q) hipFree(q);
    free(m);
}

extern "C" bqb_mlp *bqb_mlp_create(const float *params, uint64_t seed)
{
    if (bqb_ensure_init()) return nullptr;
    bqb_mlp *m = (bqb_mlp *)calloc(1, sizeof(bqb_mlp));
    if (!m) { bqb_set_error("calloc"); return nullptr; }
    const size_t pb = sizeof(float) * LEARN_NPARAM;
    if (ck(hipMalloc(&m->prm, pb), "hipMalloc") || ck(hipMalloc(&m->grad, pb), "hipMalloc") ||
        ck(hipMalloc(&m->m1, pb), "hipMalloc") || ck(hipMalloc(&m->m2, pb), "hipMalloc") ||
        ck(hipMalloc(&m->wb, sizeof(unsigned short) * LB_TOTAL), "hipMalloc")) {
        bqb_mlp_destroy(m);
        return nullptr;
    }
    float *h = (float *)calloc(LEARN_NPARAM, sizeof(float));
    if (!h) { bqb_set_error("calloc"); bqb_mlp_destroy(m); return nullptr; }
    if (params) memcpy(h, params, pb);
    else {
        // He-uniform initial weights from the counter RNG, zero biases
        const int off[4] = {LEARN_W1, LEARN_W2, LEARN_W3, LEARN_W4};
        const int fan[4] = {LEARN_NF, LEARN_H, LEARN_H, LEARN_H};
        const int cnt[4] = {LEARN_H * LEARN_NF, LEARN_H * LEARN_H, LEARN_H * LEARN_H, 2 * LEARN_H};
        for (int l = 0; l < 4; ++l) {
            const float a = sqrtf(6.f / (float)(l == 0 ? 16 : fan[l]));
            for (int i = 0; i < cnt[l]; ++i) {
                const double u = (double)(bqb_draw(bqb_key(seed, (uint64_t)l), (uint64_t)i) >> 11) * 0x1.0p-53;
                h[off[l] + i] = (l == 0 && i % LEARN_NF >= 16) ? 0.f : (float)((2.0 * u - 1.0) * a) * (l == 3 ? 0.1f : 1.f);
            }
        }
    }
    int rc = ck(hipMemcpy(m->prm, h, pb, hipMemcpyHostToDevice), "H2D params");
    free(h);
    if (rc || ck(hipMemset(m->m1, 0, pb), "memset") || ck(hipMemset(m->m2, 0, pb), "memset") || mlp_pack(m) ||
        ck(hipDeviceSynchronize(), "mlp init")) {
        bqb_mlp_destroy(m);
        return nullptr;
    }
    return m;
}

extern "C" int bqb_mlp_get_params(bqb_mlp *m, float *out)
{
    return ck(hipMemcpy(out, m->prm, sizeof(float) * LEARN_NPARAM, hipMemcpyDeviceToHost), "D2H params");
}

extern "C" int bqb_mlp_set_params(bqb_mlp *m, const float *in)
{
    if (ck(hipMemcpy(m->prm, in, sizeof(float) * LEARN_NPARAM, hipMemcpyHostToDevice), "H2D params")) return -1;
    return mlp_pack(m) || ck(hipDeviceSynchronize(), "pack");
}

static int grow(bqb_mlp *m, int rows)
{
    if (rows <= m->cap) return 0;
    int cap = 1024;
    while (cap < rows) cap *= 2;
    void **bufs[] = {(void **)&m->H1, (void **)&m->H2, (void **)&m->H3, (void **)&m->dY3, (void **)&m->dY2,
                     (void **)&m->dY1};
    for (void **b : bufs) {
        if (*b) hipFree(*b);
        *b = nullptr;
        if (ck(hipMalloc(b, sizeof(unsigned short) * (size_t)cap * LEARN_H), "hipMalloc")) return -1;
    }
    if (m->dY4) hipFree(m->dY4);
    if (m->out) hipFree(m->out);
    if (m->loss) hipFree(m->loss);
    m->dY4 = nullptr; m->out = nullptr; m->loss = nullptr;
    if (ck(hipMalloc(&m->dY4, sizeof(unsigned short) * (size_t)cap * LEARN_NO), "hipMalloc") ||
        ck(hipMalloc(&m->out, sizeof(float) * (size_t)cap * LEARN_NO), "hipMalloc") ||
        ck(hipMalloc(&m->loss, sizeof(double) * (size_t)(cap / LEARN_ROWS + 1)), "hipMalloc"))
        return -1;
    m->cap = cap;
    return 0;
}

// Uploads a data set: feat [M][32] (FP32, rounded to bf16 here), target [M][2].
extern "C" int bqb_mlp_set_data(bqb_mlp *m, int M, const float *feat, const float *target)
{
    if (M < 1) { bqb_set_error("mlp data: M = %d", M); return -1; }
    if (m->feat) hipFree(m->feat);
    if (m->target) hipFree(m->target);
    if (m->rows) hipFree(m->rows);
    m->feat = nullptr; m->target = nullptr; m->rows = nullptr; m->M = 0;
    unsigned short *hb = (unsigned short *)malloc(sizeof(unsigned short) * (size_t)M * LEARN_NF);
    if (!hb) { bqb_set_error("malloc"); return -1; }
    for (size_t i = 0; i < (size_t)M * LEARN_NF; ++i) {
        unsigned int u;
        memcpy(&u, &feat[i], 4);
        if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
        hb[i] = (unsigned short)(u >> 16);
    }
    int rc = ck(hipMalloc(&m->feat, sizeof(unsigned short) * (size_t)M * LEARN_NF), "hipMalloc") ||
             ck(hipMalloc(&m->target, sizeof(float) * (size_t)M * 2), "hipMalloc") ||
             ck(hipMalloc(&m->rows, sizeof(int) * (size_t)M), "hipMalloc") ||
             ck(hipMemcpy(m->feat, hb, sizeof(unsigned short) * (size_t)M * LEARN_NF, hipMemcpyHostToDevice), "H2D") ||
             (target && ck(hipMemcpy(m->target, target, sizeof(float) * (size_t)M * 2, hipMemcpyHostToDevice), "H2D"));
    free(hb);
    if (rc) return -1;
    m->M = M;
    return 0;
}

// Forward and backward on rows[0..B) of the data set (rows NULL: the first
// B rows); gradients into m->grad; returns the mean loss in *loss.
static int mlp_fwd_bwd(bqb_mlp *m, int B, const int *rows_dev, double *loss)
{
    if (grow(m, B)) return -1;
    const int blocks = (B + LEARN_ROWS - 1) / LEARN_ROWS;
    hipLaunchKernelGGL(mlp_fwd_kernel, dim3(blocks), dim3(LEARN_T), 0, 0, B, rows_dev, m->feat, m->wb, m->prm,
                       m->H1, m->H2, m->H3, m->out);
    hipLaunchKernelGGL(mlp_bwd_kernel, dim3(blocks), dim3(LEARN_T), 0, 0, B, rows_dev, m->target, m->wb, m->H1,
                       m->H2, m->H3, m->out, m->dY4, m->dY3, m->dY2, m->dY1, m->loss);
    struct { const unsigned short *dy; int ldy, N; const unsigned short *x; int ldx, K; const int *xr; int woff, boff; } L[4] = {
        {m->dY1, LEARN_H, LEARN_H, m->feat, LEARN_NF, LEARN_NF, rows_dev, LEARN_W1, LEARN_B1},
        {m->dY2, LEARN_H, LEARN_H, m->H1, LEARN_H, LEARN_H, nullptr, LEARN_W2, LEARN_B2},
        {m->dY3, LEARN_H, LEARN_H, m->H2, LEARN_H, LEARN_H, nullptr, LEARN_W3, LEARN_B3},
        {m->dY4, LEARN_NO, LEARN_NO, m->H3, LEARN_H, LEARN_H, nullptr, LEARN_W4, LEARN_B4}};
    for (int l = 0; l < 4; ++l) {
        dim3 g((L[l].N + 63) / 64, (L[l].K + 63) / 64);
        hipLaunchKernelGGL(mlp_wgrad_kernel, g, dim3(LEARN_T), 0, 0, B, L[l].N, L[l].K, L[l].dy, L[l].ldy, L[l].x,
                           L[l].ldx, L[l].xr, m->grad + L[l].woff);
        hipLaunchKernelGGL(mlp_bgrad_kernel, dim3((L[l].N + 255) / 256), dim3(256), 0, 0, B, L[l].N, L[l].dy,
                           L[l].ldy, m->grad + L[l].boff);
    }
    if (ck(hipGetLastError(), "mlp fwd/bwd launch")) return -1;
    if (loss) {
        double *hl = (double *)malloc(sizeof(double) * blocks);
        if (!hl) { bqb_set_error("malloc"); return -1; }
        if (ck(hipMemcpy(hl, m->loss, sizeof(double) * blocks, hipMemcpyDeviceToHost), "D2H loss")) { free(hl); return -1; }
        double s = 0.0;
        for (int i = 0; i < blocks; ++i) s += hl[i];
        free(hl);
        *loss = s / B;
    }
    return 0;
}

extern "C" int bqb_mlp_gradient(bqb_mlp *m, int M, const float *feat, const float *target, float *grad,
                                double *loss)
{
    if (bqb_mlp_set_data(m, M, feat, target)) return -1;
    if (hipMemset(m->grad, 0, sizeof(float) * LEARN_NPARAM) != hipSuccess) { bqb_set_error("memset"); return -1; }
    if (mlp_fwd_bwd(m, M, nullptr, loss)) return -1;
    return ck(hipMemcpy(grad, m->grad, sizeof(float) * LEARN_NPARAM, hipMemcpyDeviceToHost), "D2H grad");
}

extern "C" int bqb_mlp_forward(bqb_mlp *m, int M, const float *feat, float *out)
{
    if (bqb_mlp_set_data(m, M, feat, nullptr) || grow(m, M)) return -1;
    const int blocks = (M + LEARN_ROWS - 1) / LEARN_ROWS;
    hipLaunchKernelGGL(mlp_fwd_kernel, dim3(blocks), dim3(LEARN_T), 0, 0, M, nullptr, m->feat, m->wb, m->prm,
                       nullptr, nullptr, nullptr, m->out);
    if (ck(hipGetLastError(), "mlp forward")) return -1;
    float *h = (float *)malloc(sizeof(float) * (size_t)M * LEARN_NO);
    if (!h) { bqb_set_error("malloc"); return -1; }
    int rc = ck(hipMemcpy(h, m->out, sizeof(float) * (size_t)M * LEARN_NO, hipMemcpyDeviceToHost), "D2H out");
    if (!rc)
        for (int i = 0; i < M; ++i) { out[2 * i] = h[(size_t)i * LEARN_NO]; out[2 * i + 1] = h[(size_t)i * LEARN_NO + 1]; }
    free(h);
    return rc;
}

// Adam over `epochs` passes of the data set in shuffled mini-batches of
// `batch` rows (the order from the counter RNG with `seed`); loss_out
// (epochs entries, may be NULL) receives each epoch's mean loss.
extern "C" int bqb_mlp_train(bqb_mlp *m, int M, const float *feat, const float *target, int epochs, int batch,
                             float lr, uint64_t seed, double *loss_out)
{
    if (batch < LEARN_ROWS || epochs < 1) { bqb_set_error("mlp train: batch >= 64, epochs >= 1"); return -1; }
    if (bqb_mlp_set_data(m, M, feat, target) || grow(m, batch < M ? batch : M)) return -1;
    int *perm = (int *)malloc(sizeof(int) * (size_t)M);
    if (!perm) { bqb_set_error("malloc"); return -1; }
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int rc = 0;
    for (int ep = 0; ep < epochs && !rc; ++ep) {
        for (int i = 0; i < M; ++i) perm[i] = i;
        for (int i = M - 1; i > 0; --i) {           // Fisher-Yates
            const int j = (int)(bqb_draw(bqb_key(seed, (uint64_t)ep), (uint64_t)i) % (uint64_t)(i + 1));
            const int t = perm[i]; perm[i] = perm[j]; perm[j] = t;
        }
        if ((rc = ck(hipMemcpy(m->rows, perm, sizeof(int) * (size_t)M, hipMemcpyHostToDevice), "H2D rows"))) break;
        double sum = 0.0;
        int nb = 0;
        for (int s0 = 0; s0 + batch <= M || (s0 < M && nb == 0); s0 += batch) {
            const int B = (s0 + batch <= M) ? batch : M - s0;
            double l = 0.0;
            if ((rc = mlp_fwd_bwd(m, B, m->rows + s0, loss_out ? &l : nullptr))) break;
            ++m->step;
            const float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
            hipLaunchKernelGGL(mlp_adam_kernel, dim3((LEARN_NPARAM + 255) / 256), dim3(256), 0, 0, m->prm, m->grad,
                               m->m1, m->m2, LEARN_NPARAM, lr, b1, b2, eps, c1, c2);
            if ((rc = mlp_pack(m))) break;
            sum += l;
            ++nb;
        }
        if (loss_out) loss_out[ep] = nb ? sum / nb : 0.0;
    }
    free(perm);
    if (!rc) rc = ck(hipDeviceSynchronize(), "mlp train");
    return rc;
}

// Diagnostic: the device features of explicit cuts of one node (X and L
// n x n column-major; cut records as bqb_fct_eval).
extern "C" int bqb_cut_features(int n, const double *X, const double *L, int ntri, const int *tri, int npent,
                                const int *pent, int nhept, const int *hept, float round, float gap, float *out)
{
    if (bqb_ensure_init()) return -1;
    const int m = ntri + npent + nhept;
    if (n < 2 || m < 1) { bqb_set_error("cut features: n >= 2 and at least one cut"); return -1; }
    const size_t nn = (size_t)n * n;
    size_t bytes = 2 * nn * 8 + (size_t)ntri * sizeof(TriCut) + (size_t)(npent + nhept) * sizeof(KgonCut) +
                   (size_t)m * LEARN_NF * 4 + 1024;
    char *d = nullptr;
    if (ck(hipMalloc(&d, bytes), "hipMalloc")) return -1;
    TriCut *ht = (TriCut *)calloc(ntri + 1, sizeof(TriCut));
    KgonCut *hk = (KgonCut *)calloc(npent + nhept + 1, sizeof(KgonCut));
    int rc = -1;
    double lsum = 0.0;
    CutCtx cx;
    char *q = d;
    if (!ht || !hk) { bqb_set_error("calloc"); goto out; }
    for (int c = 0; c < ntri; ++c) { ht[c].i = tri[4 * c]; ht[c].j = tri[4 * c + 1]; ht[c].k = tri[4 * c + 2]; ht[c].type = tri[4 * c + 3]; }
    for (int c = 0; c < npent; ++c) { hk[c].type = pent[6 * c]; for (int a = 0; a < 5; ++a) hk[c].p[a] = pent[6 * c + 1 + a]; }
    for (int c = 0; c < nhept; ++c) { hk[npent + c].type = hept[8 * c]; for (int a = 0; a < 7; ++a) hk[npent + c].p[a] = hept[8 * c + 1 + a]; }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            if (i != j) lsum += fabs(L[i + (size_t)j * n]);
    cx.X = (const double *)q; q += nn * 8;
    cx.L = (const double *)q; q += nn * 8;
    cx.tri = (const TriCut *)q; q += ((size_t)ntri * sizeof(TriCut) + 15) & ~(size_t)15;
    cx.pent = (const KgonCut *)q; cx.hept = cx.pent + npent; q += ((size_t)(npent + nhept) * sizeof(KgonCut) + 15) & ~(size_t)15;
    cx.ld = n; cx.n = n; cx.ntri = ntri; cx.npent = npent;
    cx.lscale = (float)(lsum / ((double)n * (n - 1))) + 1e-12f;
    cx.round = round;
    cx.cuts = (float)m;
    cx.gap = gap;
    if (ck(hipMemcpy((void *)cx.X, X, nn * 8, hipMemcpyHostToDevice), "H2D") ||
        ck(hipMemcpy((void *)cx.L, L, nn * 8, hipMemcpyHostToDevice), "H2D") ||
        (ntri && ck(hipMemcpy((void *)cx.tri, ht, ntri * sizeof(TriCut), hipMemcpyHostToDevice), "H2D")) ||
        (npent + nhept && ck(hipMemcpy((void *)cx.pent, hk, (npent + nhept) * sizeof(KgonCut), hipMemcpyHostToDevice), "H2D")))
        goto out;
    hipLaunchKernelGGL(cut_features_kernel, dim3((m + 255) / 256), dim3(256), 0, 0, cx, m, (float *)q);
    if (ck(hipGetLastError(), "features launch") || ck(hipDeviceSynchronize(), "features")) goto out;
    rc = ck(hipMemcpy(out, q, (size_t)m * LEARN_NF * 4, hipMemcpyDeviceToHost), "D2H");
out:
    free(ht);
    free(hk);
    hipFree(d);
    return rc;
}

// ---------------------------------------------------------------- the bound
extern "C" hipError_t bqb_launch_learn(const BatchDims *d, const BatchPtrs *p, const bqb_mlp *m, float drop,
                                       int *rec_count, int rec_cap, float *rec_feat, int *rec_key, int *idx,
                                       unsigned char *keep, int after_round, hipStream_t s)
{
    LearnArgs la;
    la.after_round = after_round;
    la.wb = m ? m->wb : nullptr;
    la.prm = m ? m->prm : nullptr;
    la.drop = drop;
    la.rec_count = rec_count;
    la.rec_cap = rec_cap;
    la.rec_feat = rec_feat;
    la.rec_key = rec_key;
    la.idx = idx;
    la.keep = keep;
    hipLaunchKernelGGL(learn_cuts_kernel, dim3(d->nodes), dim3(LEARN_T), 0, s, *d, *p, la);
    return hipGetLastError();
}
