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

// The dense energy-CV scatter and the atom-slot scatter in one pass: each
// thread owns 4 consecutive entries of one force component, adds the dense
// term and, for the atoms that have Jacobian rows (one bit per GPU atom),
// the rows' sum in storage order, each rounded to fixed point on its own as
// the separate kernels round them: the forces equal those of k_scatter +
// k_scatter_dense bit for bit, with one read-modify-write per entry instead
// of an atomic per component and slot.  dU/dCV comes from the finalize.
__global__ __launch_bounds__(256) void k_scatter_dense_rows(StepView v, ScatterPlan sp,
                                                            const float* const* __restrict__ fext, int energyValue,
                                                            const int* __restrict__ slotOf,
                                                            const unsigned* __restrict__ mask, int slotStride,
                                                            int maskStride, int w0)
{
    const int w = w0 + blockIdx.y;
    const float* Fe = fext[w];
    const WalkerInfo wi = v.walkers[w];
    long long* F = v.force + wi.forceOff;
    const double* U = v.dUdCV + (size_t)w * v.V;
    const size_t rbase = (size_t)w * v.R;
    if (!Fe) {  // no energy input for this walker: the slots alone, with atomics
        for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < v.S; s += gridDim.x * blockDim.x) {
            const int b = sp.segStart[s], e = sp.segStart[s + 1];
            if (b == e) continue;
            double fx = 0.0, fy = 0.0, fz = 0.0;
            for (int k = b; k < e; k++) {
                const double u = U[sp.entryCv[k]];
                fx -= u * (double)v.jx[rbase + k];
                fy -= u * (double)v.jy[rbase + k];
                fz -= u * (double)v.jz[rbase + k];
            }
            const int g = v.gidx[(size_t)w * v.S + s];
            atomic_add_fixed(F + g, fx);
            atomic_add_fixed(F + g + wi.padded, fy);
            atomic_add_fixed(F + g + 2 * (size_t)wi.padded, fz);
        }
        return;
    }
    const double u = U[energyValue];
    const int* sOf = slotOf + (size_t)w * slotStride;
    const unsigned* msk = mask + (size_t)w * maskStride;
    const long long P = wi.padded;
    const long long n4 = (3LL * P) >> 2;  // padded is a multiple of 4
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (long long)gridDim.x * blockDim.x) {
        const float4 ef = reinterpret_cast<const float4*>(Fe)[i];
        const long long e0 = 4 * i;
        const int c = (int)(e0 / P);                   // component of all four entries
        const int g0 = (int)(e0 - (long long)c * P);   // first GPU atom (a multiple of 4)
        const unsigned bits = (msk[g0 >> 5] >> (g0 & 31)) & 0xFu;
        longlong2* p = reinterpret_cast<longlong2*>(F + e0);
        longlong2 a = p[0], b = p[1];
        long long q[4] = {(long long)(u * (double)ef.x * FIXED_SCALE), (long long)(u * (double)ef.y * FIXED_SCALE),
                          (long long)(u * (double)ef.z * FIXED_SCALE), (long long)(u * (double)ef.w * FIXED_SCALE)};
        if (bits) {
            const float* J = c == 0 ? v.jx : (c == 1 ? v.jy : v.jz);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                if (!((bits >> t) & 1u)) continue;
                const int s = sOf[g0 + t];
                double f = 0.0;
                for (int k = sp.segStart[s]; k < sp.segStart[s + 1]; k++) f -= U[sp.entryCv[k]] * (double)J[rbase + k];
                q[t] += (long long)(f * FIXED_SCALE);
            }
        }
        a.x += q[0];
        a.y += q[1];
        b.x += q[2];
        b.y += q[3];
        p[0] = a;
        p[1] = b;
    }
}

void launch_scatter_dense_rows(const StepView& v, const ScatterPlan& sp, const float* const* fext, int energyValue,
                               const int* slotOf, const unsigned* mask, int slotStride, int maskStride, int maxPadded,
                               int w0, int nw, hipStream_t s)
{
    const long long n4 = (3LL * maxPadded) / 4;
    int nb = (int)((n4 + 255) / 256);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(k_scatter_dense_rows, dim3(nb, nw), dim3(256), 0, s, v, sp, fext, energyValue, slotOf, mask,
                       slotStride, maskStride, w0);
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
