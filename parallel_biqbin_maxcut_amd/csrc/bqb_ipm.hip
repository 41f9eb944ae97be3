// bqb_ipm.hip -- launches of the batched interior-point kernel with FP64
// products (the kernel itself is in bqb_ipm_kernel.h; the mixed-precision
// instantiations are compiled separately in bqb_ipm_mx.hip), and the rule
// that picks the register tile.
#include <hip/hip_runtime.h>

#include "bqb_batch.h"
#include "bqb_ipm_kernel.h"

extern "C" hipError_t bqb_launch_ipm_mixed(const IpmBatchArgs *a, int nodes, int tile, hipStream_t s);

namespace bqb {

template <int P, int R>
static hipError_t launch_ipm(const IpmBatchArgs &a, int nodes, hipStream_t s)
{
    hipLaunchKernelGGL((ipm_batch_kernel<P, R, false>), dim3(nodes), dim3(P * P), 0, s, a);
    return hipGetLastError();
}

}  // namespace bqb


using namespace bqb;

// Picks the register-tile shape for the largest n in the batch.
static int g_ipm_tile = 0;   // 0: by size; else a forced tile (tuning)

// Forces the IPM register tile (P*100 + R: 1604 ... 3208) for later
// contexts and batches whose size it covers; 0 restores the size rule.
extern "C" int bqb_set_ipm_tile(int tile)
{
    switch (tile) {
    case 0: case 1604: case 1605: case 1606: case 1607: case 3204: case 3205:
    case 3206: case 3207: case 3208:
    case 6401: case 8001: case 9601: case 11201: case 12801: case 16001: case 19201: case 22401: case 25601:
        g_ipm_tile = tile;
        return 0;
    default:
        return -1;
    }
}

extern "C" int bqb_ipm_reg_tile_for_ld(int ld)
{
    switch (ld) {
    case 64: return 1604;
    case 80: return 1605;
    case 96: return 1606;
    case 112: return 1607;
    case 128: return 3204;
    case 160: return 3205;
    case 192: return 3206;
    case 224: return 3207;
    case 256: return 3208;
    default: return 0;
    }
}

// Slab leading dimensions of the size classes, and the IPM tile chosen per
// class: the register tile of that size, or the blocked solve with the same
// slabs (tile ld*100 + 1).  Defaults: register tiles up to 192, the blocked
// solve above (where the 32 x 7 and 32 x 8 register tiles spill); the tile
// autotuner (bqb_autotune_ipm) or bqb_set_ipm_tile_table replace them.
static const int k_class_ld[BQB_TILE_CLASSES] = {64, 80, 96, 112, 128, 160, 192, 224, 256};
static int g_tile_table[BQB_TILE_CLASSES] = {1604, 1605, 1606, 1607, 3204, 3205, 3206, 22401, 25601};

static int tile_class(int nmax)
{
    for (int c = 0; c < BQB_TILE_CLASSES; ++c)
        if (nmax <= k_class_ld[c]) return c;
    return -1;
}

extern "C" int bqb_ipm_tile_for(int nmax)
{
    if (g_ipm_tile && (g_ipm_tile / 100) * (g_ipm_tile % 100) >= nmax) return g_ipm_tile;
    const int c = tile_class(nmax);
    return c < 0 || nmax < 1 ? 0 : g_tile_table[c];
}

extern "C" int bqb_ipm_tile_table(int *tiles)
{
    for (int c = 0; c < BQB_TILE_CLASSES; ++c) tiles[c] = g_tile_table[c];
    return BQB_TILE_CLASSES;
}

extern "C" int bqb_set_ipm_tile_table(const int *tiles)
{
    for (int c = 0; c < BQB_TILE_CLASSES; ++c) {
        const int ld = k_class_ld[c], t = tiles[c];
        if (t != bqb_ipm_reg_tile_for_ld(ld) && t != ld * 100 + 1) return -1;
    }
    for (int c = 0; c < BQB_TILE_CLASSES; ++c) g_tile_table[c] = tiles[c];
    return 0;
}

extern "C" int bqb_ipm_class_ld(int c) { return c >= 0 && c < BQB_TILE_CLASSES ? k_class_ld[c] : 0; }

// Tile and product precision per class for contexts whose precision is
// automatic (bqb_ctx_set_precision with slices = -1): the fastest of the
// register / blocked kernels with FP64 or bf16 products.  Defaults from a
// measurement of 128-256 basic solves per class on MI355X: bf16 products
// lose below 193 (the products are a small share of a solve there, and
// slicing costs more than it saves) and at 225-256 (the blocked FP64 solve
// is fastest); at 193-224 the register tile with bf16 products wins (27.0
// vs 34.7 ms, FP64 register tile, and 36.5 ms, blocked).
// bqb_autotune_ipm_auto measures and replaces them.
static int g_auto_tile[BQB_TILE_CLASSES] = {1604, 1605, 1606, 1607, 3204, 3205, 3206, 3207, 25601};
static int g_auto_slices[BQB_TILE_CLASSES] = {0, 0, 0, 0, 0, 0, 0, 4, 0};

extern "C" int bqb_ipm_auto_table(int *tiles, int *slices)
{
    for (int c = 0; c < BQB_TILE_CLASSES; ++c) {
        if (tiles) tiles[c] = g_auto_tile[c];
        if (slices) slices[c] = g_auto_slices[c];
    }
    return BQB_TILE_CLASSES;
}

extern "C" int bqb_set_ipm_auto_table(const int *tiles, const int *slices)
{
    for (int c = 0; c < BQB_TILE_CLASSES; ++c) {
        const int ld = k_class_ld[c], t = tiles[c], s = slices[c];
        if (t != bqb_ipm_reg_tile_for_ld(ld) && t != ld * 100 + 1) return -1;
        if (s != 0 && (s < 2 || s > BQB_MIX_MAX_SLICES)) return -1;
    }
    for (int c = 0; c < BQB_TILE_CLASSES; ++c) {
        g_auto_tile[c] = tiles[c];
        g_auto_slices[c] = slices[c];
    }
    return 0;
}

extern "C" int bqb_ipm_auto_for(int nmax, int *slices)
{
    const int c = tile_class(nmax);
    if (c < 0 || nmax < 1) return 0;
    if (slices) *slices = g_auto_slices[c];
    return g_auto_tile[c];
}

extern "C" int bqb_ipm_tile_reg_for(int nmax)
{
    if (nmax <= 64) return 1604;
    if (nmax <= 80) return 1605;
    if (nmax <= 96) return 1606;
    if (nmax <= 112) return 1607;
    if (nmax <= 128) return 3204;
    if (nmax <= 160) return 3205;
    // above 160 the 32 x 32 tile holds more registers than a 1024-thread
    // block has per lane: the compiler spills part of the tile
    if (nmax <= 192) return 3206;
    if (nmax <= 224) return 3207;
    if (nmax <= 256) return 3208;
    return 0;
}

namespace {

// Nodes whose solve failed (a factorization or a line search gave up, or
// the iteration cap was hit): they are solved again with FP64 products and
// another tile.  flags (optional, stride in ints) receive `flag`.
__global__ void ipm_retry_mask_kernel(int nodes, const int *active, const int *status, int *mask2, int *flags,
                                      int stride, int flag)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nodes) return;
    const int st = status[b];
    const int retry = (!active || active[b]) &&
                      (st == ERR_CHOL_Z || st == ERR_SCHUR || st == ERR_IPM_ITERS || st == ERR_LINESEARCH);
    mask2[b] = retry;
    if (retry && flags) flags[(size_t)b * stride] |= flag;
}

}  // namespace

extern "C" hipError_t bqb_launch_ipm_retry(const IpmBatchArgs *a, int nodes, int tile, int *mask2, int *flags,
                                           int stride, int flag, hipStream_t s)
{
    hipLaunchKernelGGL(ipm_retry_mask_kernel, dim3((nodes + 255) / 256), dim3(256), 0, s, nodes, a->active,
                       (const int *)a->status, mask2, flags, stride, flag);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    IpmBatchArgs b = *a;
    b.active = mask2;
    b.mix_slices = 0;
    b.flops_mx = nullptr;
    return bqb_launch_ipm_batch(&b, nodes, tile, s);
}

extern "C" hipError_t bqb_launch_ipm_batch(const IpmBatchArgs *a, int nodes, int tile, hipStream_t s)
{
    const int blocked = tile % 100 == 1;
    if (blocked) return bqb_launch_ipm_blk(a, nodes, s);
    if (a->mix_slices > 0) {
        if (a->mix_slices < 2 || a->mix_slices > BQB_MIX_MAX_SLICES) return hipErrorInvalidValue;
        return bqb_launch_ipm_mixed(a, nodes, tile, s);
    }
    switch (tile) {
    case 1604: return launch_ipm<16, 4>(*a, nodes, s);
    case 1605: return launch_ipm<16, 5>(*a, nodes, s);
    case 1606: return launch_ipm<16, 6>(*a, nodes, s);
    case 1607: return launch_ipm<16, 7>(*a, nodes, s);
    case 3204: return launch_ipm<32, 4>(*a, nodes, s);
    case 3205: return launch_ipm<32, 5>(*a, nodes, s);
    case 3206: return launch_ipm<32, 6>(*a, nodes, s);
    case 3207: return launch_ipm<32, 7>(*a, nodes, s);
    case 3208: return launch_ipm<32, 8>(*a, nodes, s);
    default: return hipErrorInvalidValue;
    }
}
