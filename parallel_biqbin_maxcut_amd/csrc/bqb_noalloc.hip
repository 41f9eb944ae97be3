// bqb_noalloc.hip -- the batched IPM for callers that cannot allocate.
//
// bqb_ipm_batch_noalloc() runs the same kernel as bqb_ipm_batch() on inputs
// and outputs that are already in device memory, and takes its working
// memory (zero-padded L and X slabs, the IPM's three work slabs, trial
// counters) from one file-scope device array instead of hipMalloc.  The
// array is sized once (BQB_NOALLOC_SCRATCH_BYTES) and lives as long as the
// process on every GPU the library runs on.  Its device address is taken
// with hipGetSymbolAddress once per device (at bqb_init, or at the first
// call on a device that bqb_init did not select) and only that address is
// ever handed to kernels.
//
// Every call on one GPU uses that GPU's array: calls that overlap on one GPU
// (two host threads, or two streams without an ordering between them) race.
// Each call rewrites the slabs it reads (L is packed, X is set to I by the
// kernel, the work slabs are written before they are read), so nothing left
// by an earlier call is used.
#include <hip/hip_runtime.h>

#include <mutex>

#include "bqb_kernels.h"
#include "bqb_runtime.h"

#define BQB_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) double scratch[BQB_NOALLOC_SCRATCH_BYTES / sizeof(double)];

static void *g_scratch_ptr[BQB_NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

namespace {

constexpr int PT = 256;

// Offset (in doubles) of node b's matrix in the caller's concatenated
// column-major layout: the sum of n[c]^2 over c < b.  Every workgroup of a
// launch computes its own (nodes is small next to the matrix work).
__device__ size_t compact_offset(const int *n, int b)
{
    __shared__ unsigned long long s_part[PT / 64];
    unsigned long long acc = 0;
    for (int c = threadIdx.x; c < b; c += PT) {
        const long long v = n[c] > 0 ? n[c] : 0;
        acc += (unsigned long long)(v * v);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    unsigned long long s = 0;
    for (int w = 0; w < PT / 64; ++w) s += s_part[w];
    __syncthreads();
    return (size_t)s;
}

// L (compact, caller's memory) -> zero-padded ld x ld slab of node b.  A
// node whose n does not fit the slab is left to the IPM kernel, which
// reports it as ERR_SIZE without reading its slab.
__global__ __launch_bounds__(PT) void noalloc_pack_kernel(const int *n, const double *L, double *Ls, int ld)
{
    const int b = blockIdx.x;
    const size_t off = compact_offset(n, b);
    const int nb = n[b];
    if (nb < 1 || nb > ld) return;
    double *dst = Ls + (size_t)b * ld * ld;
    const double *src = L + off;
    for (int e = threadIdx.x; e < ld * ld; e += PT) {
        const int i = e % ld, j = e / ld;
        dst[e] = (i < nb && j < nb) ? src[i + (size_t)j * nb] : 0.0;
    }
}

// X slab of node b -> the caller's compact layout (nodes the IPM rejected
// keep their X untouched).
__global__ __launch_bounds__(PT) void noalloc_unpack_kernel(const int *n, const double *Xs, double *X, int ld)
{
    const int b = blockIdx.x;
    const size_t off = compact_offset(n, b);
    const int nb = n[b];
    if (nb < 1 || nb > ld) return;
    const double *src = Xs + (size_t)b * ld * ld;
    double *dst = X + off;
    for (int e = threadIdx.x; e < nb * nb; e += PT) {
        const int i = e % nb, j = e / nb;
        dst[e] = src[i + (size_t)j * ld];
    }
}

}  // namespace

// Device address of the scratch array on the current device, looked up once
// per device.  0 on success.
extern "C" int bqb_noalloc_setup(void)
{
    int dev = 0;
    if (bqb_check(hipGetDevice(&dev), "hipGetDevice")) return -1;
    if (dev < 0 || dev >= BQB_NOALLOC_MAX_DEVICES) {
        bqb_set_error("device %d: at most %d devices are supported", dev, BQB_NOALLOC_MAX_DEVICES);
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return 0;
    void *ptr = nullptr;
    if (bqb_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return -1;
    if (!ptr) { bqb_set_error("hipGetSymbolAddress(scratch) returned no address"); return -1; }
    g_scratch_ptr[dev] = ptr;
    return 0;
}

extern "C" size_t bqb_ipm_noalloc_bytes(int nodes, int nmax)
{
    const int tile = bqb_ipm_tile_reg_for(nmax);
    const int ld = (tile / 100) * (tile % 100);
    if (nodes < 1 || !ld) return 0;
    const size_t slab = sizeof(double) * (size_t)ld * ld;
    return 5 * slab * (size_t)nodes + ((sizeof(int) * (size_t)nodes + 255) & ~(size_t)255);
}

extern "C" int bqb_ipm_batch_noalloc(int nodes, int nmax, const int *n, const double *L, double *X,
                                     double *phi, int *iters, int *status, void *stream)
{
    if (nodes < 1) { bqb_set_error("noalloc IPM: nodes = %d", nodes); return -1; }
    if (!n || !L || !X || !phi || !iters || !status) { bqb_set_error("noalloc IPM: null device pointer"); return -1; }
    // register tiles: their three work slabs per node are what the scratch
    // layout holds
    const int tile = bqb_ipm_tile_reg_for(nmax);
    if (nmax < 1 || !tile) { bqb_set_error("noalloc IPM: nmax = %d outside 1..256", nmax); return -1; }
    const size_t need = bqb_ipm_noalloc_bytes(nodes, nmax);
    if (need > (size_t)BQB_NOALLOC_SCRATCH_BYTES) {
        bqb_set_error("noalloc IPM: %d nodes of size <= %d need %zu bytes of scratch, more than its %zu",
                      nodes, nmax, need, (size_t)BQB_NOALLOC_SCRATCH_BYTES);
        return -1;
    }
    if (bqb_noalloc_setup()) return -1;
    int dev = 0;
    if (bqb_check(hipGetDevice(&dev), "hipGetDevice")) return -1;
    char *base = (char *)g_scratch_ptr[dev];
    const int ld = (tile / 100) * (tile % 100);
    const size_t slab = (size_t)ld * ld;
    double *Ls = (double *)base;
    double *Xs = Ls + slab * nodes;
    double *W = Xs + slab * nodes;
    int *trials = (int *)(W + 3 * slab * nodes);
    hipStream_t s = (hipStream_t)stream;

    hipLaunchKernelGGL(noalloc_pack_kernel, dim3(nodes), dim3(PT), 0, s, n, L, Ls, ld);
    if (bqb_check(hipGetLastError(), "noalloc pack launch")) return -1;
    IpmBatchArgs a = {};
    a.n = n;
    a.L = Ls;
    a.X = Xs;
    a.work = W;
    a.ld = ld;
    a.phi = phi;
    a.iters = iters;
    a.trials = trials;
    a.status = status;
    a.wslabs = 3;
    if (bqb_check(bqb_launch_ipm_batch(&a, nodes, tile, s), "noalloc IPM launch")) return -1;
    hipLaunchKernelGGL(noalloc_unpack_kernel, dim3(nodes), dim3(PT), 0, s, n, (const double *)Xs, X, ld);
    return bqb_check(hipGetLastError(), "noalloc unpack launch");
}
