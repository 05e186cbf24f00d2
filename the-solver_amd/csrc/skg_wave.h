// Wavefront-level helpers shared by the HIP sources (gfx950, wave64): DPP
// lane swaps, v_readlane, and fixed-order wave reductions built on them.
#ifndef SKG_WAVE_H
#define SKG_WAVE_H

#include <hip/hip_runtime.h>

namespace {

// A lane's neighbour under a DPP pattern (all 64 lanes active).
template <int Ctrl> __device__ __forceinline__ double dpp(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), Ctrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), Ctrl, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
template <int Ctrl> __device__ __forceinline__ float dpp(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), Ctrl, 0xf, 0xf, false));
}

// The value lane `l` holds (l uniform; v_readlane).
__device__ __forceinline__ double lane_value(double v, int l)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane(__double2loint(v), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double((int)hi, (int)lo);
}
__device__ __forceinline__ float lane_value(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Sum over the lanes of one wavefront, the same in every lane and in a fixed
// order: pairs, quads, eights and rows of 16 through DPP lane swaps (each
// step adds two partial sums of the same shape, so every lane of a row holds
// the same bits), then the four row sums from v_readlane.  All 64 lanes must
// be active.
template <typename T> __device__ __forceinline__ T wave_sum(T v)
{
    v += dpp<0xB1>(v);    // quad_perm [1,0,3,2]: lane ^ 1
    v += dpp<0x4E>(v);    // quad_perm [2,3,0,1]: lane ^ 2
    v += dpp<0x141>(v);   // row_half_mirror: lane i <-> 7 - i of its eight
    v += dpp<0x140>(v);   // row_mirror: lane i <-> 15 - i of its row
    return ((lane_value(v, 0) + lane_value(v, 16)) + lane_value(v, 32)) + lane_value(v, 48);
}

__device__ __forceinline__ double wave_max(double v)
{
    v = fmax(v, dpp<0xB1>(v));
    v = fmax(v, dpp<0x4E>(v));
    v = fmax(v, dpp<0x141>(v));
    v = fmax(v, dpp<0x140>(v));
    return fmax(fmax(lane_value(v, 0), lane_value(v, 16)), fmax(lane_value(v, 32), lane_value(v, 48)));
}

}  // namespace

#endif
