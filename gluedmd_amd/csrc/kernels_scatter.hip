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

// Dense scatter of an energy CV whose Jacobian is -F_ext over every atom
// (the CV_ENERGY rows, CommonGluedKernels.cpp:4227-4230 / 2471-2503):
// F += dV/dU * F_ext.  Streams 12 B of F_ext and read-modify-writes 24 B of
// fixed-point force per atom; 4 consecutive field entries per thread with
// 16-byte loads and stores.  One y-row of blocks per walker; the walker's
// F_ext pointer comes from a device table (graph-safe).
__global__ __launch_bounds__(256) void k_scatter_dense(StepView v, const float* const* __restrict__ fext, int energyValue,
                                                     int w0)
{
    const int w = w0 + blockIdx.y;
    const float* Fe = fext[w];
    if (!Fe) return;
    const WalkerInfo wi = v.walkers[w];
    const double u = v.dUdCV[(size_t)w * v.V + energyValue];
    long long* F = v.force + wi.forceOff;
    const long long n4 = (3LL * wi.padded) >> 2;  // padded is a multiple of 4
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (long long)gridDim.x * blockDim.x) {
        const float4 e = reinterpret_cast<const float4*>(Fe)[i];
        longlong2* p = reinterpret_cast<longlong2*>(F + 4 * i);
        longlong2 a = p[0], b = p[1];
        a.x += (long long)(u * (double)e.x * FIXED_SCALE);
        a.y += (long long)(u * (double)e.y * FIXED_SCALE);
        b.x += (long long)(u * (double)e.z * FIXED_SCALE);
        b.y += (long long)(u * (double)e.w * FIXED_SCALE);
        p[0] = a;
        p[1] = b;
    }
}

void launch_scatter(const StepView& v, const ScatterPlan& sp, int w0, int nw, int includeForces, hipStream_t s)
{
    int nb = (v.S + SCATTER_BLOCK - 1) / SCATTER_BLOCK;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_scatter, dim3(nb, nw), dim3(SCATTER_BLOCK), 0, s, v, sp, w0, includeForces);
}

void launch_scatter_dense(const StepView& v, const float* const* fext, int energyValue, int maxPadded, int w0, int nw,
                          hipStream_t s)
{
    const long long n4 = (3LL * maxPadded) / 4;
    int nb = (int)((n4 + 255) / 256);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(k_scatter_dense, dim3(nb, nw), dim3(256), 0, s, v, fext, energyValue, w0);
}

}  // namespace glued
