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
#include "scatter_dev.h"

namespace glued {



constexpr int SCATTER_BLOCK = 256;
constexpr int MAX_LDS_VALUES = 1024;

__global__ __launch_bounds__(SCATTER_BLOCK) void k_scatter(StepView v, ScatterPlan sp, int w0, int includeForces)
{
    __shared__ double sU[MAX_LDS_VALUES];
    const int w = w0 + blockIdx.y;
    combine_dudcv(v, sp, w, sU, MAX_LDS_VALUES, blockIdx.x == 0);
    if (!includeForces) return;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= v.S) return;
    scatter_slot(v, sp, w, s, sU, MAX_LDS_VALUES);
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
