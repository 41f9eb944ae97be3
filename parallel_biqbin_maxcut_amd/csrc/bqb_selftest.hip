// bqb_selftest.hip -- unit entry for the workgroup-distributed primitives of
// bqb_dist.h: one workgroup loads an n x n matrix, applies one primitive and
// stores the result.  Used by the tests to check each primitive against a
// dense reference independently of the solvers built on top of them.
#include <hip/hip_runtime.h>

#include "bqb_dist.h"
#include "bqb_kernels.h"
#include "bqb_mixed.h"

namespace bqb {

template <int P, int R>
__global__ __launch_bounds__(P * P) void dist_selftest_kernel(int op, int n, int ld, const double *in,
                                                              double *out, int *flag)
{
    constexpr int NMAX = P * R;
    __shared__ double s_col[2 * NMAX];
    __shared__ double s_T[Lds<P, R, 8>::TRANS];
    const Thr t = thr_id<P>();
    const Slab si = make_slab<P>(in, ld, t), so = make_slab<P>(out, ld, t);
    double A[R][R];
    dload<P, R>(A, si);
    bool ok = true;
    if (op == 0) ok = sweep_inverse<P, R>(A, n, s_col, s_T, t);
    else if (op == 1) ok = chol_lower<P, R>(A, n, s_col, t);
    else if (op == 2) ok = chol_test<P, R>(A, n, s_col, t);
    else if (op == 3) symmetrize<P, R>(A, s_T, t);
    dstore<P, R>(A, so);
    if (t.tid == 0) *flag = ok ? 1 : 0;
}

// C = A B (n x n, column-major, n <= 256) through the sliced bf16
// matrix-core product of the mixed-precision IPM; S holds the slices.
__global__ __launch_bounds__(256) void mx_selftest_kernel(int n, int s, const double *A, const double *B, double *C,
                                                          unsigned short *S)
{
    __shared__ int e[2 * 256];
    const auto la = [&](int i, int k) { return A[i + (size_t)k * n]; };
    const auto lb = [&](int j, int k) { return B[k + (size_t)j * n]; };
    mx_product<256>(la, lb, n, s, S, e, C, n);
}

// Plain streaming copy (16-byte vectors, grid-stride): the calibration kernel
// for the HBM byte counters (its algorithmic traffic is exactly 2 x bytes).
__global__ __launch_bounds__(256) void copy_probe_kernel(const double2 *__restrict__ src,
                                                         double2 *__restrict__ dst, size_t n2)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}

}  // namespace bqb

extern "C" hipError_t bqb_launch_copy_probe(const void *src, void *dst, size_t bytes, int reps)
{
    const size_t n2 = bytes / sizeof(double2);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(bqb::copy_probe_kernel, dim3(4096), dim3(256), 0, 0, (const double2 *)src,
                           (double2 *)dst, n2);
    return hipGetLastError();
}

extern "C" hipError_t bqb_launch_mx_selftest(int n, int s, const double *A, const double *B, double *C,
                                             unsigned short *S)
{
    hipLaunchKernelGGL(bqb::mx_selftest_kernel, dim3(1), dim3(256), 0, 0, n, s, A, B, C, S);
    return hipGetLastError();
}

extern "C" hipError_t bqb_launch_dist_selftest(int tile, int op, int n, int ld, const double *in,
                                               double *out, int *flag)
{
    using namespace bqb;
    switch (tile) {
    case 1604: hipLaunchKernelGGL((dist_selftest_kernel<16, 4>), dim3(1), dim3(256), 0, 0, op, n, ld, in, out, flag); break;
    case 1607: hipLaunchKernelGGL((dist_selftest_kernel<16, 7>), dim3(1), dim3(256), 0, 0, op, n, ld, in, out, flag); break;
    case 3204: hipLaunchKernelGGL((dist_selftest_kernel<32, 4>), dim3(1), dim3(1024), 0, 0, op, n, ld, in, out, flag); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
