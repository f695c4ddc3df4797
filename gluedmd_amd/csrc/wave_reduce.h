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

// Reduce-scatter of up to 16 FP64 sums over a wave.  Four DPP steps inside
// each 16-lane row halve the values a lane carries (row mirror, half
// mirror, quad xor 2, quad xor 1: the partners of every step agree on the
// lane bits earlier steps split on, so they carry the same value subset),
// then two xor shuffles add across rows.  Afterwards every lane l holds the
// wave total of value (l & 15); values NV..15 are zero padding.  Each total
// is formed by the same additions whatever the data.
template <int CTRL, int H>
__device__ inline void rs_step(double (&v)[16], bool upper)
{
#pragma unroll
    for (int j = 0; j < H; j++) {
        const double send = upper ? v[j] : v[j + H];
        const double keep = upper ? v[j + H] : v[j];
        v[j] = keep + dpp_f64<CTRL>(send);
    }
}

template <int NV>
__device__ inline double wave_reduce_scatter(const double (&x)[NV])
{
    static_assert(NV >= 1 && NV <= 16, "at most 16 values");
    const int lane = threadIdx.x & 63;
    double v[16];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = k < NV ? x[k < NV ? k : 0] : 0.0;
    rs_step<DPP_ROW_MIRROR, 8>(v, lane & 8);
    rs_step<DPP_ROW_HALF_MIRROR, 4>(v, lane & 4);
    rs_step<DPP_QUAD_XOR2, 2>(v, lane & 2);
    rs_step<DPP_QUAD_XOR1, 1>(v, lane & 1);
    double t = v[0];
    t += __shfl_xor(t, 16, 64);
    t += __shfl_xor(t, 32, 64);
    return t;
}

// Reduce-scatter of up to 8 FP64 sums over a wave, for few values: lane l
// (l < NV) ends with the wave total of value l (other lanes hold partial
// sums of no use).  Within each group of W = 4 or 8 lanes the DPP steps of
// wave_reduce_scatter halve the values a lane carries, so lane l holds value
// (l mod W) summed over its group; the groups of a 16-lane row are then
// added by row rotations (by 4 and 8 lanes), and the rows by two xor
// shuffles.  Each
// total is formed by the same additions whatever the data.
constexpr int DPP_ROW_ROR4 = 0x124;  // row rotation by 4: lane i receives lane (i - 4) mod 16 of its row
constexpr int DPP_ROW_ROR8 = 0x128;  // row rotation by 8: lane i receives lane i ^ 8

template <int NV>
__device__ inline double wave_reduce_scatter_small(const double (&x)[NV])
{
    static_assert(NV >= 1 && NV <= 8, "at most 8 values");
    constexpr int W = NV <= 4 ? 4 : 8;
    const int lane = threadIdx.x & 63;
    double v[16];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = k < NV ? x[k < NV ? k : 0] : 0.0;
    if constexpr (W == 8) rs_step<DPP_ROW_HALF_MIRROR, 4>(v, lane & 4);
    rs_step<DPP_QUAD_XOR2, 2>(v, lane & 2);
    rs_step<DPP_QUAD_XOR1, 1>(v, lane & 1);
    double t = v[0];
    if constexpr (W == 4) t += dpp_f64<DPP_ROW_ROR4>(t);
    t += dpp_f64<DPP_ROW_ROR8>(t);
    t += __shfl_xor(t, 16, 64);
    t += __shfl_xor(t, 32, 64);
    return t;
}

__device__ inline double wave_sum1(double x)
{
    double v[1] = {x};
    wave_sum_n<1>(v);
    return v[0];
}

// Ticket of a cross-workgroup reduction: true in every thread of the
// workgroup that finished last.  Only thread 0 wrote this workgroup's
// partial, so only it fences (an agent-scope fence writes back the XCD's L2
// and is expensive on this part: one per workgroup, never one per wave).
__device__ inline bool last_block(unsigned int* counter, unsigned int total)
{
    __shared__ int shLast;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        shLast = atomicAdd(counter, 1u) == total - 1;
        if (shLast) __threadfence();
    }
    __syncthreads();
    return shLast != 0;
}

// Sum of n partials written by other workgroups of the same launch (read
// with device-scope loads after the ticket), by one wave: lane l adds
// entries l, l + 64, ... in index order, then the butterfly; the order
// depends only on n.
__device__ inline double wave_sum_partials(const double* src, int n)
{
    double a = 0.0;
    for (int j = threadIdx.x & 63; j < n; j += 64) a += __hip_atomic_load(src + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return wave_sum1(a);
}

}  // namespace glued
