// Device side of the chain-rule force scatter, shared by the scatter kernel
// and the fused bias + scatter kernel (both need it in their translation
// unit).  CommonGluedKernels.cpp:2430-2456, re-organised: Jacobian rows are
// stored grouped by atom slot (the CV kernels write row r at rowPos[r]), so
// the rows of one atom are contiguous; each thread owns one atom, sums
// F = -sum_rows dU/dCV * J in FP64 and issues one 64-bit fixed-point atomic
// per component.
#pragma once
#include "glued_internal.h"

namespace glued {

__device__ inline void atomic_add_fixed(long long* p, double f)
{
    const long long q = (long long)(f * FIXED_SCALE);
    atomicAdd((unsigned long long*)p, (unsigned long long)q);
}

// dU/dCV of walker w into sU (LDS, first nLds values) and, from the threads
// of workgroup `writer`, into global dUdCV.  Every thread of the workgroup
// must call it (it ends with a barrier).
__device__ inline void combine_dudcv(const StepView& v, const ScatterPlan& sp, int w, double* sU, int nLds, bool writer)
{
    for (int c = threadIdx.x; c < v.V; c += blockDim.x) {
        double u = 0.0;
        if (sp.testGrad) {
            u = sp.testGrad[c];
        } else {
            for (int k = sp.combStart[c]; k < sp.combStart[c + 1]; k++) u += v.bgrad[(size_t)w * v.G + sp.combSlot[k]];
        }
        if (c < nLds) sU[c] = u;
        if (writer) v.dUdCV[(size_t)w * v.V + c] = u;
    }
    __syncthreads();
}

// Force of one atom slot s of walker w (sU holds dU/dCV for c < nLds; larger
// CV indices are read from global dUdCV, written by an earlier launch stage).
__device__ inline void scatter_slot(const StepView& v, const ScatterPlan& sp, int w, int s, const double* sU, int nLds)
{
    const int b = sp.segStart[s], e = sp.segStart[s + 1];
    if (b == e) return;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const size_t base = (size_t)w * v.R;
    for (int k = b; k < e; k++) {
        const int c = sp.entryCv[k];
        const double u = c < nLds ? sU[c] : v.dUdCV[(size_t)w * v.V + c];
        const float4 j = v.jrow[base + k];
        fx -= u * (double)j.x;
        fy -= u * (double)j.y;
        fz -= u * (double)j.z;
    }
    const WalkerInfo wi = v.walkers[w];
    const int g = v.gidx[(size_t)w * v.S + s];
    long long* F = v.force + wi.forceOff;
    atomic_add_fixed(F + g, fx);
    atomic_add_fixed(F + g + wi.padded, fy);
    atomic_add_fixed(F + g + 2 * (size_t)wi.padded, fz);
}

// Whole-walker scatter by one workgroup (the fused stage; requires V <= nLds).
__device__ inline void scatter_walker_block(const StepView& v, const ScatterPlan& sp, int w, double* sU, int nLds)
{
    combine_dudcv(v, sp, w, sU, nLds, true);
    for (int s = threadIdx.x; s < v.S; s += blockDim.x) scatter_slot(v, sp, w, s, sU, nLds);
}

}  // namespace glued
