// Wave-level reductions through DPP lane permutations and v_readlane.
//
// A reduction over a 16-lane DPP row takes four permute+op steps (swap
// neighbours, swap pairs, mirror halves of 8, mirror halves of 16); the
// four row results are then combined from SGPRs, so the whole-wave result
// is uniform.  No LDS traffic and no waits on the LDS queue, unlike
// ds_bpermute based shuffles.  Every lane of a row ends with the same bits
// (each step combines a value with its partner's, and both partners form
// a+b and b+a, which are equal in IEEE arithmetic).
//
// All 64 lanes of the wave must be active: a permute reading a disabled
// lane yields 0 instead of that lane's value.
#pragma once

#include <hip/hip_runtime.h>

namespace bqb {

// GFX9 DPP control words
constexpr int DPP_SWAP1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int DPP_SWAP2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int DPP_HALF_MIRROR = 0x141; // lane i <- lane 7-i within 8
constexpr int DPP_MIRROR = 0x140;      // lane i <- lane 15-i within 16

template <int CTRL, class T>
__device__ __forceinline__ T dpp_move(T v)
{
    return __builtin_amdgcn_update_dpp(T(0), v, CTRL, 0xF, 0xF, false);
}

// v_readlane of a double (the builtin itself is 32-bit: read both halves)
__device__ __forceinline__ double lane_value(double v, int lane)
{
    typedef int i32x2 __attribute__((ext_vector_type(2)));
    i32x2 h = __builtin_bit_cast(i32x2, v);
    h.x = __builtin_amdgcn_readlane(h.x, lane);
    h.y = __builtin_amdgcn_readlane(h.y, lane);
    return __builtin_bit_cast(double, h);
}

__device__ __forceinline__ int lane_value(int v, int lane)
{
    return __builtin_amdgcn_readlane(v, lane);
}

// sum over each 16-lane row, in every lane of the row
__device__ __forceinline__ double row16_sum(double v)
{
    v += dpp_move<DPP_SWAP1>(v);
    v += dpp_move<DPP_SWAP2>(v);
    v += dpp_move<DPP_HALF_MIRROR>(v);
    v += dpp_move<DPP_MIRROR>(v);
    return v;
}

__device__ __forceinline__ double row16_max(double v)
{
    v = fmax(v, dpp_move<DPP_SWAP1>(v));
    v = fmax(v, dpp_move<DPP_SWAP2>(v));
    v = fmax(v, dpp_move<DPP_HALF_MIRROR>(v));
    v = fmax(v, dpp_move<DPP_MIRROR>(v));
    return v;
}

__device__ __forceinline__ double row16_min(double v)
{
    v = fmin(v, dpp_move<DPP_SWAP1>(v));
    v = fmin(v, dpp_move<DPP_SWAP2>(v));
    v = fmin(v, dpp_move<DPP_HALF_MIRROR>(v));
    v = fmin(v, dpp_move<DPP_MIRROR>(v));
    return v;
}

// whole-wave sum / max / min, uniform result
__device__ __forceinline__ double wave_sum64(double v)
{
    v = row16_sum(v);
    return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

__device__ __forceinline__ double wave_max64(double v)
{
    v = row16_max(v);
    return fmax(fmax(lane_value(v, 0), lane_value(v, 16)), fmax(lane_value(v, 32), lane_value(v, 48)));
}

__device__ __forceinline__ double wave_min64(double v)
{
    v = row16_min(v);
    return fmin(fmin(lane_value(v, 0), lane_value(v, 16)), fmin(lane_value(v, 32), lane_value(v, 48)));
}

// the same over lanes 0..31 only (lanes 32..63 must still be active)
__device__ __forceinline__ double wave_sum32(double v)
{
    v = row16_sum(v);
    return lane_value(v, 0) + lane_value(v, 16);
}

__device__ __forceinline__ double wave_max32(double v)
{
    v = row16_max(v);
    return fmax(lane_value(v, 0), lane_value(v, 16));
}

__device__ __forceinline__ double wave_min32(double v)
{
    v = row16_min(v);
    return fmin(lane_value(v, 0), lane_value(v, 16));
}

// Whole-wave (value, index) selections, uniform result.  The winner under a
// total order does not depend on the combining order, so these give the
// same pair as any other reduction tree.
template <int CTRL>
__device__ __forceinline__ void argmax_step(double &v, int &i)
{
    const double ov = dpp_move<CTRL>(v);
    const int oi = dpp_move<CTRL>(i);
    if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
}

template <int CTRL>
__device__ __forceinline__ void argmin_step(double &v, int &i)
{
    const double ov = dpp_move<CTRL>(v);
    const int oi = dpp_move<CTRL>(i);
    if (ov < v || (ov == v && oi < i)) { v = ov; i = oi; }
}

// largest value, smallest index among equal values
__device__ __forceinline__ void wave_argmax64(double &v, int &i)
{
    argmax_step<DPP_SWAP1>(v, i);
    argmax_step<DPP_SWAP2>(v, i);
    argmax_step<DPP_HALF_MIRROR>(v, i);
    argmax_step<DPP_MIRROR>(v, i);
    double bv = lane_value(v, 0);
    int bi = lane_value(i, 0);
#pragma unroll
    for (int r = 16; r < 64; r += 16) {
        const double rv = lane_value(v, r);
        const int ri = lane_value(i, r);
        if (rv > bv || (rv == bv && ri < bi)) { bv = rv; bi = ri; }
    }
    v = bv;
    i = bi;
}

// smallest value, smallest index among equal values
__device__ __forceinline__ void wave_argmin64(double &v, int &i)
{
    argmin_step<DPP_SWAP1>(v, i);
    argmin_step<DPP_SWAP2>(v, i);
    argmin_step<DPP_HALF_MIRROR>(v, i);
    argmin_step<DPP_MIRROR>(v, i);
    double bv = lane_value(v, 0);
    int bi = lane_value(i, 0);
#pragma unroll
    for (int r = 16; r < 64; r += 16) {
        const double rv = lane_value(v, r);
        const int ri = lane_value(i, r);
        if (rv < bv || (rv == bv && ri < bi)) { bv = rv; bi = ri; }
    }
    v = bv;
    i = bi;
}

}  // namespace bqb
