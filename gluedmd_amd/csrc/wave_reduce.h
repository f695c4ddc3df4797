// Wave-wide FP64 sums for 64-lane wavefronts.  A butterfly: every lane ends
// with the same bits (each step adds a pair of values in both orders, and FP
// addition is commutative), and the order depends only on the lane layout,
// so results are deterministic.
//
// Several independent sums are reduced together, one butterfly step for all
// of them at a time, so the cross-row shuffles of one step are issued back to
// back and waited for once.
#pragma once
#include <hip/hip_runtime.h>

namespace glued {

// One DPP step: every lane receives lane CTRL(i)'s value (both halves of
// the FP64 word moved by the same lane permutation).
template <int CTRL>
__device__ inline double dpp_f64(double x)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

constexpr int DPP_QUAD_XOR1 = 0xB1;     // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_XOR2 = 0x4E;     // quad_perm [2,3,0,1]
constexpr int DPP_ROW_HALF_MIRROR = 0x141;  // lane i <-> 7 - i within 8
constexpr int DPP_ROW_MIRROR = 0x140;       // lane i <-> 15 - i within 16

// Steps inside a 16-lane row use DPP lane moves (plain VALU moves); the two
// steps across rows use xor shuffles.  Each step pairs lane i with lane p(i),
// p an involution, so both lanes of a pair add the same two values: after
// the six steps every lane holds the same bits (the whole-wave sum).
template <int NV>
__device__ inline void wave_sum_n(double (&v)[NV])
{
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] += dpp_f64<DPP_QUAD_XOR1>(v[k]);
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] += dpp_f64<DPP_QUAD_XOR2>(v[k]);
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] += dpp_f64<DPP_ROW_HALF_MIRROR>(v[k]);
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] += dpp_f64<DPP_ROW_MIRROR>(v[k]);
#pragma unroll
    for (int o = 16; o <= 32; o <<= 1) {
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
