// Values of the neighbouring lane of a wavefront through DPP (a VALU move,
// no LDS round trip as for __shfl_up / __shfl_down):
//   lane_prev(x): x of lane i-1 (DPP wave_shr:1), 0 on lane 0;
//   lane_next(x): x of lane i+1 (DPP wave_shl:1), 0 on lane 63.
// Equal to __shfl_up(x, 1) / __shfl_down(x, 1) on every lane whose
// neighbour lies in its group; callers replace the values of the lanes at a
// group's edge, as they do for the shuffles.
#pragma once

#include <hip/hip_runtime.h>

namespace chy {

constexpr int kDppWaveShl1 = 0x130;  // lane i reads lane i+1
constexpr int kDppWaveShr1 = 0x138;  // lane i reads lane i-1

template <int CTRL>
__device__ __forceinline__ int dpp_move(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ double dpp_move(double x)
{
    const long long v = __double_as_longlong(x);
    const int lo = dpp_move<CTRL>((int)(v & 0xffffffffll)), hi = dpp_move<CTRL>((int)(v >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

template <int CTRL>
__device__ __forceinline__ float dpp_move(float x)
{
    return __int_as_float(dpp_move<CTRL>(__float_as_int(x)));
}

template <typename T>
__device__ __forceinline__ T lane_prev(T x)
{
    return dpp_move<kDppWaveShr1>(x);
}

template <typename T>
__device__ __forceinline__ T lane_next(T x)
{
    return dpp_move<kDppWaveShl1>(x);
}

}  // namespace chy
