// Wave-wide FP64 sums for 64-lane wavefronts.  A butterfly of xor shuffles:
// every lane ends with the same bits (each step adds a pair of values in
// both orders, and FP addition is commutative), and the order depends only
// on the lane layout, so results are deterministic.
//
// Several independent sums are reduced together, one butterfly step for all
// of them at a time: the shuffles of one step are issued back to back and
// waited for once, instead of one LDS-permute round trip per value per step.
#pragma once
#include <hip/hip_runtime.h>

namespace glued {

template <int NV>
__device__ inline void wave_sum_n(double (&v)[NV])
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double t[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) t[k] = __shfl_xor(v[k], o, 64);
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] += t[k];
    }
}

__device__ inline double wave_sum1(double x)
{
    double v[1] = {x};
    wave_sum_n<1>(v);
    return v[0];
}

}  // namespace glued
