// Chain-rule force scatter, CommonGluedKernels.cpp:2430-2456, re-organised
// for gfx950: Jacobian rows are grouped by atom at plan time (CSR over atom
// slots), so each thread owns one atom, sums F = -sum_rows dU/dCV * J in FP64
// and issues exactly one 64-bit fixed-point atomic per component.  Rows that
// hit the same atom (path s/z rows, shared torsion atoms, an atom in several
// CVs) are therefore pre-reduced before the atomic; the reference issues one
// atomic per row and truncates per row, so results differ by at most
// (rows per atom - 1) * 2^-32 kJ/mol/nm per component.
//
// The first threads of every workgroup also form dU/dCV for their walker from
// the per-bias gradient slots (the combine the reference does through +=
// into an auto-cleared buffer), or copy the injected test gradients.
#include "glued_internal.h"
#include <hip/hip_runtime.h>

namespace glued {



constexpr int SCATTER_BLOCK = 256;
constexpr int MAX_LDS_VALUES = 1024;

__device__ inline void atomic_add_fixed(long long* p, double f)
{
    const long long q = (long long)(f * FIXED_SCALE);
    atomicAdd((unsigned long long*)p, (unsigned long long)q);
}

__global__ __launch_bounds__(SCATTER_BLOCK) void k_scatter(StepView v, ScatterPlan sp, int w0, int includeForces)
{
    __shared__ double sU[MAX_LDS_VALUES];
    const int w = w0 + blockIdx.y;
    for (int c = threadIdx.x; c < v.V; c += blockDim.x) {
        double u = 0.0;
        if (sp.testGrad) {
            u = sp.testGrad[c];
        } else {
            for (int k = sp.combStart[c]; k < sp.combStart[c + 1]; k++) u += v.bgrad[(size_t)w * v.G + sp.combSlot[k]];
        }
        if (c < MAX_LDS_VALUES) sU[c] = u;
        if (blockIdx.x == 0) v.dUdCV[(size_t)w * v.V + c] = u;
    }
    __syncthreads();
    if (!includeForces) return;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= v.S) return;
    const int b = sp.segStart[s], e = sp.segStart[s + 1];
    if (b == e) return;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const size_t base = (size_t)w * v.R;
    for (int k = b; k < e; k++) {
        const int r = sp.segRow[k];
        const int c = sp.rowCv[r];
        const double u = c < MAX_LDS_VALUES ? sU[c] : v.dUdCV[(size_t)w * v.V + c];
        fx -= u * (double)v.jx[base + r];
        fy -= u * (double)v.jy[base + r];
        fz -= u * (double)v.jz[base + r];
    }
    const WalkerInfo wi = v.walkers[w];
    const int g = v.gidx[(size_t)w * v.S + s];
    long long* F = v.force + wi.forceOff;
    atomic_add_fixed(F + g, fx);
    atomic_add_fixed(F + g + wi.padded, fy);
    atomic_add_fixed(F + g + 2 * (size_t)wi.padded, fz);
}

// Dense scatter of an energy CV whose Jacobian is -F_ext over every atom:
// F += dV/dU * F_ext, one thread per (atom, component), coalesced, HBM-bound.
__global__ void k_scatter_dense(StepView v, ScatterPlan sp, int w0)
{
    const int w = w0 + blockIdx.y;
    const WalkerInfo wi = v.walkers[w];
    const double u = v.dUdCV[(size_t)w * v.V + sp.energyValue];
    const float* Fe = sp.energyF + wi.forceOff;  // same field-major layout as the force block
    long long* F = v.force + wi.forceOff;
    const long long n = 3LL * wi.padded;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double f = u * (double)Fe[i];
        F[i] += (long long)(f * FIXED_SCALE);
    }
}

void launch_scatter(const StepView& v, const ScatterPlan& sp, int w0, int nw, int includeForces, hipStream_t s)
{
    int nb = (v.S + SCATTER_BLOCK - 1) / SCATTER_BLOCK;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_scatter, dim3(nb, nw), dim3(SCATTER_BLOCK), 0, s, v, sp, w0, includeForces);
}

void launch_scatter_dense(const StepView& v, const ScatterPlan& sp, int maxPadded, int w0, int nw, hipStream_t s)
{
    long long n = 3LL * maxPadded;
    int nb = (int)((n + 255) / 256);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_scatter_dense, dim3(nb, nw), dim3(256), 0, s, v, sp, w0);
}

}  // namespace glued
