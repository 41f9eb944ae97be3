// bqb_mixed.h -- n x n products of the interior-point iteration on the bf16
// matrix cores (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.
//
// The FP64 operands are split into s bf16 slices (Ozaki's scheme): every
// line (a row of the left operand, a column of the right one) is scaled by a
// power of two 2^-e so that its largest entry lies in [1/2, 1), and its
// entries a are written as  a = sum_p d_p 2^(-7-8(p-1)) + r  with integer
// digits |d_p| <= 128 (exact in bf16's 8-bit significand) and
// |r| <= 2^(-8s).  A product of two slices sums at most 256 terms
// d d' 2^(-14-8(p+q-2)) with |d d'| <= 2^14 -- integers of at most 22 bits in
// the units of the pair -- so the FP32 accumulator of the matrix core holds
// every partial sum exactly, whatever its order.  The slice products with
// p + q <= s + 1 are combined in FP64 (smallest first) and scaled back by
// 2^(e_i + e_j): each entry of C = A B is then exact up to the remainders and
// the dropped slice terms (a slice entry is at most 2^(-8(p-1)) 2^e and
// 2^e < 2 max|line|), |error| <= (4s + 1) n 2^(-8s) max|A(i,:)| max|B(:,j)|
// (plus the final FP64 roundings).
//
// Operands and result live in the node's global slabs: the slices in a
// line-major buffer (k contiguous, lines padded to 16, k padded to 32, zero
// padding), the result as a column-major ld x ld slab.  One workgroup of NT
// threads computes one product; every wave takes 16 x 16 output tiles.
#pragma once
#include <hip/hip_runtime.h>

namespace bqb {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MX_MAX_SLICES = 4;

__host__ __device__ constexpr int mx_pad16(int n) { return (n + 15) & ~15; }
__host__ __device__ constexpr int mx_pad32(int n) { return (n + 31) & ~31; }
// bytes of one operand's slices for lines up to nmax
__host__ __device__ constexpr size_t mx_operand_bytes(int nmax, int s)
{
    return (size_t)2 * s * mx_pad16(nmax) * mx_pad32(nmax);
}

// Splits the line value v (already scaled into (-1, 1)) into s bf16 slices
// written to dst[p * pstride] (bf16 bit patterns).
__device__ __forceinline__ void mx_split(double v, int s, unsigned short *dst, size_t pstride)
{
    double r = v;
#pragma unroll
    for (int p = 0; p < MX_MAX_SLICES; ++p) {
        if (p < s) {
            const int sh = 7 + 8 * p;
            const double d = rint(ldexp(r, sh));            // |d| <= 128
            const double q = ldexp(d, -sh);                  // exact
            r -= q;                                          // exact
            const float f = (float)q;                        // exact: 8 significant bits
            dst[p * pstride] = (unsigned short)(__float_as_uint(f) >> 16);
        }
    }
}

// Slices of the n lines of one operand.  line(c, k) returns entry k of line
// c (0 <= c, k < n).  Lines are padded to mx_pad16(n) and k to mx_pad32(n)
// with zero slices; e[c] (LDS) receives the line's scale exponent.  One wave
// per line at a time.
template <int NT, class Line>
__device__ __forceinline__ void mx_slice(const Line &line, int n, int s, unsigned short *S, int *e)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lines = mx_pad16(n), kp = mx_pad32(n);
    const size_t pstride = (size_t)lines * kp;
    for (int c = wave; c < lines; c += NW) {
        double mx = 0.0;
        for (int k = lane; k < n && c < n; k += 64) mx = fmax(mx, fabs(line(c, k)));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
        const int ex = (mx > 0.0) ? ilogb(mx) + 1 : 0;      // mx 2^-ex in [1/2, 1)
        if (lane == 0) e[c] = ex;
        for (int k = lane; k < kp; k += 64) {
            const double v = (c < n && k < n) ? ldexp(line(c, k), -ex) : 0.0;
            mx_split(v, s, S + (size_t)c * kp + k, pstride);
        }
    }
}

// C(i, j) = sum_k A(i, k) B(k, j) for i, j < n from the slices SA (lines =
// rows of A) and SB (lines = columns of B) and their exponents eA, eB (LDS);
// C column-major with leading dimension ldc.
template <int NT>
__device__ __forceinline__ void mx_tiles(const unsigned short *SA, const unsigned short *SB, const int *eA,
                                         const int *eB, int n, int s, double *C, int ldc)
{
    constexpr int NW = NT / 64;
    constexpr int NP = MX_MAX_SLICES * (MX_MAX_SLICES + 1) / 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lines = mx_pad16(n), kp = mx_pad32(n);
    const size_t pstride = (size_t)lines * kp;
    const int T = lines / 16;
    const int r16 = lane & 15, kq = 8 * (lane >> 4);
    for (int tile = wave; tile < T * T; tile += NW) {
        const int bj = tile % T, bi = tile / T;
        f32x4 acc[NP];
#pragma unroll
        for (int q = 0; q < NP; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
        const unsigned short *pa = SA + (size_t)(bi * 16 + r16) * kp + kq;
        const unsigned short *pb = SB + (size_t)(bj * 16 + r16) * kp + kq;
        for (int k0 = 0; k0 < kp; k0 += 32) {
            bf16x8 a[MX_MAX_SLICES], b[MX_MAX_SLICES];
#pragma unroll
            for (int p = 0; p < MX_MAX_SLICES; ++p) {
                if (p < s) {
                    a[p] = __builtin_bit_cast(bf16x8, *(const uint4 *)(pa + p * pstride + k0));
                    b[p] = __builtin_bit_cast(bf16x8, *(const uint4 *)(pb + p * pstride + k0));
                }
            }
            int q = 0;
#pragma unroll
            for (int p = 0; p < MX_MAX_SLICES; ++p)
#pragma unroll
                for (int r = 0; r < MX_MAX_SLICES; ++r) {
                    if (p < s && r < s && p + r <= s - 1)
                        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], b[r], acc[q], 0, 0, 0);
                    if (p + r <= MX_MAX_SLICES - 1) ++q;
                }
        }
        // lane holds rows 4 (lane / 16) + v of column lane % 16 of the tile
        const int j = bj * 16 + r16;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = bi * 16 + 4 * (lane >> 4) + v;
            double sum = 0.0;
            // smallest pairs first: p + r = s - 1 down to 0
#pragma unroll
            for (int g = 2 * MX_MAX_SLICES - 2; g >= 0; --g) {
                int q2 = 0;
#pragma unroll
                for (int p = 0; p < MX_MAX_SLICES; ++p)
#pragma unroll
                    for (int r = 0; r < MX_MAX_SLICES; ++r) {
                        if (p + r == g && p < s && r < s && p + r <= s - 1) sum += (double)acc[q2][v];
                        if (p + r <= MX_MAX_SLICES - 1) ++q2;
                    }
            }
            if (i < n && j < n) C[i + (size_t)j * ldc] = ldexp(sum, eA[i] + eB[j]);
        }
    }
}

// The whole product: slices of both operands, then the tiles.  Lines of A
// are given by la(i, k) = A(i, k), lines of B by lb(j, k) = B(k, j).  S
// holds 2 * mx_operand_bytes(n, s) bytes, e 2 * mx_pad16(n) ints of LDS.
template <int NT, class LA, class LB>
__device__ __forceinline__ void mx_product(const LA &la, const LB &lb, int n, int s, unsigned short *S, int *e,
                                           double *C, int ldc)
{
    unsigned short *SA = S;
    unsigned short *SB = S + mx_operand_bytes(n, s) / 2;
    int *eA = e, *eB = e + mx_pad16(n);
    __syncthreads();      // operands written by the caller, e still read by a previous product
    mx_slice<NT>(la, n, s, SA, eA);
    mx_slice<NT>(lb, n, s, SB, eB);
    __syncthreads();
    mx_tiles<NT>(SA, SB, eA, eB, n, s, C, ldc);
    __syncthreads();
}

}  // namespace bqb
