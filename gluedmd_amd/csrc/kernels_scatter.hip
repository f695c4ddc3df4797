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

// The dense energy-CV scatter and the atom-slot scatter in one pass,
// atom-major: each thread owns 4 consecutive GPU atoms of one walker and all
// three components (16-byte loads and stores per component array), adds the
// dense term and, for the atoms that have Jacobian rows (one bit per GPU
// atom), the rows' sum in storage order, each rounded to fixed point on its
// own as the separate kernels round them: the forces equal those of
// k_scatter + k_scatter_dense bit for bit, with one read-modify-write per
// entry instead of an atomic per component and slot, and every row of a
// masked atom read once (one 16-byte row).  dU/dCV comes from the finalize.
//
// Dependent loads of a thread with rows: {mask, running count} of its
// 32-atom word (issued with the field loads), the row ranges of its atoms
// (a list in GPU order: a wave reads it nearly in sequence), then the rows
// and their CV indices; dU/dCV is staged in LDS behind the one barrier,
// which comes after every first-round load is in flight.
constexpr int DENSE_ROWS_BLOCK = 256;
constexpr int DENSE_ROWS_LDS_U = 256;  // dU/dCV values staged in LDS (more are read from memory)

__global__ __launch_bounds__(DENSE_ROWS_BLOCK) void k_scatter_dense_rows(StepView v, ScatterPlan sp,
                                                                         const float* const* __restrict__ fext,
                                                                         int energyValue,
                                                                         const int2* __restrict__ rowList,
                                                                         const uint2* __restrict__ mask,
                                                                         int listStride, int maskStride, int w0)
{
    __shared__ double sU[DENSE_ROWS_LDS_U];
    const int w = w0 + blockIdx.y;
    const float* Fe = fext[w];
    const WalkerInfo wi = v.walkers[w];
    long long* F = v.force + wi.forceOff;
    const double* U = v.dUdCV + (size_t)w * v.V;
    const size_t rbase = (size_t)w * v.R;
    if (!Fe) {  // no energy input for this walker (uniform over the workgroup): the slots alone, with atomics
        for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < v.S; s += gridDim.x * blockDim.x) {
            const int b = sp.segStart[s], e = sp.segStart[s + 1];
            if (b == e) continue;
            double fx = 0.0, fy = 0.0, fz = 0.0;
            for (int k = b; k < e; k++) {
                const double u = U[sp.entryCv[k]];
                const float4 j = v.jrow[rbase + k];
                fx -= u * (double)j.x;
                fy -= u * (double)j.y;
                fz -= u * (double)j.z;
            }
            const int g = v.gidx[(size_t)w * v.S + s];
            atomic_add_fixed(F + g, fx);
            atomic_add_fixed(F + g + wi.padded, fy);
            atomic_add_fixed(F + g + 2 * (size_t)wi.padded, fz);
        }
        return;
    }
    const long long P = wi.padded;
    const long long n4 = P >> 2;  // groups of 4 atoms (padded is a multiple of 4)
    const long long i = (long long)blockIdx.x * DENSE_ROWS_BLOCK + threadIdx.x;
    const bool own = i < n4;
    const long long g0 = 4 * (own ? i : n4 - 1);  // clamped: every thread loads, none waits on a branch
    float4 ef[3];
    longlong2 fa[3], fb[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        ef[c] = reinterpret_cast<const float4*>(Fe + c * P + g0)[0];
        const longlong2* p = reinterpret_cast<const longlong2*>(F + c * P + g0);
        fa[c] = p[0];
        fb[c] = p[1];
    }
    const uint2 mw = mask[(size_t)w * maskStride + (g0 >> 5)];
    const double u = U[energyValue];
    const bool ldsU = v.V <= DENSE_ROWS_LDS_U;
    if (ldsU)
        for (int c = threadIdx.x; c < v.V; c += DENSE_ROWS_BLOCK) sU[c] = U[c];
    __syncthreads();
    if (!own) return;
    const unsigned bits = (mw.x >> (g0 & 31)) & 0xFu;
    long long q[3][4];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        q[c][0] = (long long)(u * (double)ef[c].x * FIXED_SCALE);
        q[c][1] = (long long)(u * (double)ef[c].y * FIXED_SCALE);
        q[c][2] = (long long)(u * (double)ef[c].z * FIXED_SCALE);
        q[c][3] = (long long)(u * (double)ef[c].w * FIXED_SCALE);
    }
    if (bits) {
        // the row ranges of this word's atoms with rows follow each other in
        // the list
        // the list (clamped to this group's entries: unconditional loads)
        const int2* rl = rowList + (size_t)w * listStride + mw.y + __popc(mw.x & ((1u << (g0 & 31)) - 1u));
        const int nr = __popc(bits);
        // four named entries and selects (an indexed array here was kept on
        // the scratch stack: a private-memory round trip per atom)
        const int2 rr0 = rl[0], rr1 = rl[min(1, nr - 1)], rr2 = rl[min(2, nr - 1)], rr3 = rl[min(3, nr - 1)];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int e = __popc(bits & ((1u << t) - 1u));  // list entry of atom t
            const int2 lo = (e & 1) ? rr1 : rr0, hi = (e & 1) ? rr3 : rr2;
            const int2 r = (e & 2) ? hi : lo;
            const int kb = (bits >> t) & 1u ? r.x : 0, ke = (bits >> t) & 1u ? r.y : 0;
            double fx = 0.0, fy = 0.0, fz = 0.0;
            for (int k = kb; k < ke; k++) {
                const int cv = sp.entryCv[k];
                const float4 j = v.jrow[rbase + k];
                const double uk = ldsU ? sU[cv] : U[cv];
                fx -= uk * (double)j.x;
                fy -= uk * (double)j.y;
                fz -= uk * (double)j.z;
            }
            if ((bits >> t) & 1u) {
                q[0][t] += (long long)(fx * FIXED_SCALE);
                q[1][t] += (long long)(fy * FIXED_SCALE);
                q[2][t] += (long long)(fz * FIXED_SCALE);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        longlong2* p = reinterpret_cast<longlong2*>(F + c * P + g0);
        fa[c].x += q[c][0];
        fa[c].y += q[c][1];
        fb[c].x += q[c][2];
        fb[c].y += q[c][3];
        p[0] = fa[c];
        p[1] = fb[c];
    }
}

void launch_scatter_dense_rows(const StepView& v, const ScatterPlan& sp, const float* const* fext, int energyValue,
                               const int2* rowList, const uint2* mask, int listStride, int maskStride, int maxPadded,
                               int w0, int nw, hipStream_t s)
{
    // one thread per 4 atoms of the largest walker (total atoms < 2^31)
    const long long n4 = maxPadded / 4;
    const int nb = (int)((n4 + DENSE_ROWS_BLOCK - 1) / DENSE_ROWS_BLOCK);
    hipLaunchKernelGGL(k_scatter_dense_rows, dim3(nb, nw), dim3(DENSE_ROWS_BLOCK), 0, s, v, sp, fext, energyValue,
                       rowList, mask, listStride, maskStride, w0);
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
