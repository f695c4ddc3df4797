// CV kernels for gfx950.  Every kernel takes a walker range: walker = w0 +
// blockIdx.y, so one launch covers one walker (plain form, the reference's
// structure) or every walker (batched form).
//
// Arithmetic: positions are FP32 in HBM.  Differences of two FP32 coordinates
// are formed in FP64, where they are exact, and the minimum image is applied
// in FP64 with explicitly unfused operations, so distance predicates (the
// coordination cutoff) are bit-identical to the FP64 oracle.  Bulk pair
// arithmetic (switching functions) runs in FP32; reductions accumulate in
// FP64.  Few-atom CVs are evaluated entirely in FP64 (a handful of threads).
#include "glued_internal.h"
#include <hip/hip_runtime.h>
#include <cmath>
#include "wave_reduce.h"
#include "cv_small_dev.h"

namespace glued {


__device__ inline double wave_sum(double v) { return wave_sum1(v); }

// Block-wide FP64 sum of NV values per thread over NW waves (the block
// size, fixed at compile time); every thread receives the result.  Thread k
// (k < NV) adds the NW wave partials of value k in wave order, so each
// thread reads NV values from LDS rather than NV * NW.  sh holds >= NV * 16
// doubles.
template <int NV, int NW = 4>
__device__ inline void block_sum(double (&v)[NV], double* sh)
{
    static_assert(NW >= 1 && NW <= 16 && NV <= 16, "1..16 waves, at most 16 values");
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if constexpr (NV <= 2) {
        wave_sum_n<NV>(v);
        __syncthreads();
        if (lane == 0)
#pragma unroll
            for (int k = 0; k < NV; k++) sh[k * 16 + wid] = v[k];
    } else {
        // lane k of every wave holds the wave total of value k
        const double t = wave_reduce_scatter<NV>(v);
        __syncthreads();
        if (lane < NV) sh[lane * 16 + wid] = t;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < NW; i++) t += sh[threadIdx.x * 16 + i];
        sh[threadIdx.x * 16] = t;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = sh[k * 16];
    __syncthreads();
}


// Position of entry e of walker w's gather table ([W][stride] posq indices).
__device__ inline float4 gather_pos(const StepView& v, const int* __restrict__ table, int stride, int w, int e)
{
    return v.posq[table[(size_t)w * stride + e]];
}



__global__ void k_cv_small(StepView v, const SmallCVItem* __restrict__ items, int n, int w0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    small_item(v, items[i], w0 + blockIdx.y);
}

void launch_cv_small(const StepView& v, const SmallCVItem* items, int n, int w0, int nw, hipStream_t s)
{
    if (n <= 0) return;
    dim3 g((n + 63) / 64, nw);
    hipLaunchKernelGGL(k_cv_small, g, dim3(64), 0, s, v, items, n, w0);
}

// ---------------------------------------------------------------------------
// Largest eigenpair of Horn's 4x4 key matrix by Newton on its characteristic
// polynomial, eigenvector from the adjugate of (K - l I).  FP64, one thread.
// S[a*3+b] = sum_i y_ia x_ib (centred reference y, centred walker x).
// P(l) = l^4 + e2 l^2 - e3 l + e4 with e2 = -tr(K^2)/2, e3 = tr(K^3)/3,
// e4 = det K (tr K = 0); Newton from (Gx+Gy)/2 >= lmax descends monotonically.
// R (row-major) maps the reference onto the walker: R y ~ x.
// ---------------------------------------------------------------------------
__device__ inline double det3(double a, double b, double c, double d, double e, double f, double g, double h, double i)
{
    return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

// Minor of the 4x4 matrix A with row RR and column CC deleted (static indices).
template <int RR, int CC>
__device__ inline double minor4(const double (&A)[4][4])
{
    constexpr int r0 = RR == 0 ? 1 : 0, r1 = RR <= 1 ? 2 : 1, r2 = RR <= 2 ? 3 : 2;
    constexpr int c0 = CC == 0 ? 1 : 0, c1 = CC <= 1 ? 2 : 1, c2 = CC <= 2 ? 3 : 2;
    return det3(A[r0][c0], A[r0][c1], A[r0][c2], A[r1][c0], A[r1][c1], A[r1][c2], A[r2][c0], A[r2][c1], A[r2][c2]);
}

// Column CC of adj(A): adj(A)[i][CC] = (-1)^(i+CC) minor(CC, i).
template <int CC>
__device__ inline void adj_column(const double (&A)[4][4], double* q)
{
    q[0] = ((0 + CC) & 1 ? -1.0 : 1.0) * minor4<CC, 0>(A);
    q[1] = ((1 + CC) & 1 ? -1.0 : 1.0) * minor4<CC, 1>(A);
    q[2] = ((2 + CC) & 1 ? -1.0 : 1.0) * minor4<CC, 2>(A);
    q[3] = ((3 + CC) & 1 ? -1.0 : 1.0) * minor4<CC, 3>(A);
}

__device__ void horn_rotation(const double* S, double Gx, double Gy, double* R, double* lmax)
{
    const double Sxx = S[0], Sxy = S[1], Sxz = S[2], Syx = S[3], Syy = S[4], Syz = S[5], Szx = S[6], Szy = S[7],
                 Szz = S[8];
    double K[4][4];
    K[0][0] = Sxx + Syy + Szz;
    K[0][1] = K[1][0] = Syz - Szy;
    K[0][2] = K[2][0] = Szx - Sxz;
    K[0][3] = K[3][0] = Sxy - Syx;
    K[1][1] = Sxx - Syy - Szz;
    K[1][2] = K[2][1] = Sxy + Syx;
    K[1][3] = K[3][1] = Szx + Sxz;
    K[2][2] = -Sxx + Syy - Szz;
    K[2][3] = K[3][2] = Syz + Szy;
    K[3][3] = -Sxx - Syy + Szz;
    // K symmetric: tr K^2 = sum_ij K_ij^2, tr K^3 = sum_ij (K^2)_ij K_ij
    double trK2 = 0.0, trK3 = 0.0;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) trK2 += K[i][j] * K[i][j];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = i; j < 4; j++) {
            double k2 = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) k2 += K[i][k] * K[k][j];
            trK3 += (i == j ? 1.0 : 2.0) * k2 * K[i][j];
        }
    const double detK = K[0][0] * minor4<0, 0>(K) - K[0][1] * minor4<0, 1>(K) + K[0][2] * minor4<0, 2>(K) -
                        K[0][3] * minor4<0, 3>(K);
    const double e2 = -0.5 * trK2, e3 = trK3 / 3.0, e4 = detK;
    double l = 0.5 * (Gx + Gy);
    for (int it = 0; it < 60; it++) {
        double l2 = l * l;
        double f = (l2 + e2) * l2 - e3 * l + e4;
        double df = 4.0 * l2 * l + 2.0 * e2 * l - e3;
        if (!(fabs(df) > 0.0)) break;
        // approximate reciprocal (v_rcp_f64): Newton converges to the same
        // root with any slope estimate this close; only the rate of the last
        // step changes, from quadratic to a factor ~1e-7 per step
        double dl = f * __builtin_amdgcn_rcp(df);
        l -= dl;
        // quadratic convergence: after a step of relative size 1e-10 the
        // remaining error is ~1e-20 relative for separated eigenvalues
        if (fabs(dl) <= 1e-10 * fabs(l) + 1e-300) break;
    }
    *lmax = l;
    double A[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) A[i][j] = K[i][j] - (i == j ? l : 0.0);
    // adj(A) of the rank-3 symmetric A is c q q^T: its diagonal is c q_i^2, so
    // the column through the largest |diagonal cofactor| is the best-conditioned
    // multiple of the eigenvector (first maximum on ties)
    const double d0 = fabs(minor4<0, 0>(A)), d1 = fabs(minor4<1, 1>(A)), d2 = fabs(minor4<2, 2>(A)),
                 d3 = fabs(minor4<3, 3>(A));
    int ib = 0;
    double db = d0;
    if (d1 > db) { ib = 1; db = d1; }
    if (d2 > db) { ib = 2; db = d2; }
    if (d3 > db) { ib = 3; db = d3; }
    double q[4];
    switch (ib) {
        case 0: adj_column<0>(A, q); break;
        case 1: adj_column<1>(A, q); break;
        case 2: adj_column<2>(A, q); break;
        default: adj_column<3>(A, q); break;
    }
    double best = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    double qn = sqrt(best > 0.0 ? best : 1.0);
    if (!(best > 1e-300)) {
        q[0] = 1.0;
        q[1] = q[2] = q[3] = 0.0;
        qn = 1.0;
    }
    const double iq = 1.0 / qn;
    const double q0 = q[0] * iq, qx = q[1] * iq, qy = q[2] * iq, qz = q[3] * iq;
    R[0] = q0 * q0 + qx * qx - qy * qy - qz * qz;
    R[1] = 2 * (qx * qy - q0 * qz);
    R[2] = 2 * (qx * qz + q0 * qy);
    R[3] = 2 * (qy * qx + q0 * qz);
    R[4] = q0 * q0 - qx * qx + qy * qy - qz * qz;
    R[5] = 2 * (qy * qz - q0 * qx);
    R[6] = 2 * (qz * qx - q0 * qy);
    R[7] = 2 * (qz * qy + q0 * qx);
    R[8] = q0 * q0 - qx * qx - qy * qy + qz * qz;
}

// ---------------------------------------------------------------------------
// Group CVs, one 256-thread workgroup per (item, walker).
//   gyration :556-639 (COM without PBC, displacement with PBC)
//   COM distance :456-539
//   RMSD :810-881 (no fit) and the Kabsch extension (fit: split == 1)
//   DRMSD :1130-1224, contact map :1236-1323
// groupSlots: atom slots (pairs: 2 per pair); groupMass: mass per listed atom;
// groupRef: RMSD reference xyz / DRMSD d0 / contact-map [r0, n, m, w].
// ---------------------------------------------------------------------------
constexpr int GROUP_BLOCK = 256;
constexpr int GROUP_REG = 4;  // atoms per thread held in registers by the group CVs

__device__ inline double ipowd(double x, int n)
{
    double r = 1.0;
    for (int i = 0; i < n; i++) r *= x;
    return r;
}

// x^n for a small non-negative integer n by repeated squaring.
__device__ inline double ipow_sq(double x, int n)
{
    double r = 1.0;
    while (n > 0) {
        if (n & 1) r *= x;
        x *= x;
        n >>= 1;
    }
    return r;
}

// One contact-map pair, CommonGluedKernels.cpp:1236-1323 (in FP64: the
// reference's FP32 loses digits in 1 - x^m near x = 1), for the minimum-image
// difference dr = b - a and the pair's parameters {r0, n, m, w}.  Returns
// w s(|dr|) and sets c, the factor of the pair's rows: row(b) = c dr,
// row(a) = -c dr.  |1 - x^m| < 1e-8 gives s = n/m and no derivative.
__device__ inline double contact_pair(double dx, double dy, double dz, const float4& par, double& c)
{
    const double r0 = par.x, wgt = par.w;
    const int nn = (int)par.y, mm = (int)par.z;
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    const double invd = d > 1e-10 ? 1.0 / d : 0.0;
    const double ir0 = 1.0 / r0, x = d * ir0;
    // x^(n-1), x^(m-1) by squaring; x^n, x^m one product further
    const double xn1 = nn > 0 ? ipow_sq(x, nn - 1) : 0.0, xm1 = mm > 0 ? ipow_sq(x, mm - 1) : 0.0;
    const double xn = nn > 0 ? xn1 * x : 1.0, xm = mm > 0 ? xm1 * x : 1.0;
    const double num = 1.0 - xn, den = 1.0 - xm;
    double sw, dsdd;
    if (fabs(den) < 1e-8) {
        sw = (double)nn / (double)mm;
        dsdd = 0.0;
    } else {
        const double iden = 1.0 / den;
        sw = num * iden;
        dsdd = (-(double)nn * xn1 * den + (double)mm * xm1 * num) * iden * iden * ir0;
    }
    c = wgt * dsdd * invd;
    return wgt * sw;
}

// Kabsch RMSD of up to 256 * KABSCH_PER_THREAD atoms by one 256-thread
// workgroup: every thread keeps its atoms and their reference in registers,
// a wave reduce-scatter and one LDS exchange give every thread the raw moments
//   m[3a+b] = sum y'_a x_b, m[9..11] = sum x, m[12] = sum |x|^2   (y' = y - ybar),
// from which the centred covariance sum y'(x - xbar)^T = m[3a+b] - (sum y'_a) xbar_b
// and sum |x - xbar|^2 = sum |x|^2 - n |xbar|^2 (FP64); sum y' and sum |y'|^2
// are constants of the reference, formed on the host.  Every thread then
// runs the same Horn solve on the same bits (no broadcast) and writes its rows.
constexpr int KABSCH_PER_THREAD = 4;
constexpr int KABSCH_MOMENTS = 13;  // sum y'x^T (9), sum x (3), sum |x|^2; the reference sums come from the host

template <int NT>
__device__ void kabsch_block(const StepView& v, const GroupCVItem& it, const int* __restrict__ slots,
                             const float* __restrict__ ref, int w)
{
    __shared__ double shK[NT / 64][16];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int n = it.end - it.begin;
    const float* Y = ref + 3 * (size_t)it.refOff;
    const double* yc = it.refCentroid;
    float qx[KABSCH_PER_THREAD][3], qy[KABSCH_PER_THREAD][3];  // walker atom, raw FP32 reference
    int rp[KABSCH_PER_THREAD];
    // threads past the end load a valid entry (index clamped) and drop it: no
    // divergent branch, so every gather of the wave is in flight at once
#pragma unroll
    for (int r = 0; r < KABSCH_PER_THREAD; r++) {
        const int j = min(tid + NT * r, n - 1);
        const float4 p = gather_pos(v, slots, v.gStride, w, it.begin + j);
        qx[r][0] = p.x;
        qx[r][1] = p.y;
        qx[r][2] = p.z;
#pragma unroll
        for (int k = 0; k < 3; k++) qy[r][k] = Y[3 * j + k];
        rp[r] = v.rowPos[it.row0 + j];
    }
    double m[KABSCH_MOMENTS];
#pragma unroll
    for (int k = 0; k < KABSCH_MOMENTS; k++) m[k] = 0.0;
#pragma unroll
    for (int r = 0; r < KABSCH_PER_THREAD; r++) {
        if (tid + NT * r >= n) break;
        const double x[3] = {(double)qx[r][0], (double)qx[r][1], (double)qx[r][2]};
        const double y[3] = {(double)qy[r][0] - yc[0], (double)qy[r][1] - yc[1], (double)qy[r][2] - yc[2]};
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) m[3 * a + b] += y[a] * x[b];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            m[9 + k] += x[k];
            m[12] += x[k] * x[k];
        }
    }
    // lane k of each wave holds the wave total of moment k; the waves' totals
    // meet in LDS and every thread adds them in wave order
    const double tw = wave_reduce_scatter<KABSCH_MOMENTS>(m);
    if (lane < KABSCH_MOMENTS) shK[wid][lane] = tw;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KABSCH_MOMENTS; k++) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < NT / 64; q++) t += shK[q][k];
        m[k] = t;
    }
    const double xc[3] = {m[9] / n, m[10] / n, m[11] / n};
    double S[9], R[9], l;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) S[3 * a + b] = m[3 * a + b] - it.refSum[a] * xc[b];
    const double Gx = m[12] - n * (xc[0] * xc[0] + xc[1] * xc[1] + xc[2] * xc[2]);
    horn_rotation(S, Gx, it.refSq, R, &l);
    double msd = (Gx + it.refSq - 2.0 * l) / n;
    if (msd < 0.0) msd = 0.0;
    const double rv = sqrt(msd);
    const double sc = (rv > 1e-10 ? 1.0 / rv : 0.0) / n;
    if (tid == 0) v.cvVal[(size_t)w * v.V + it.value] = rv;
    const size_t base = (size_t)w * v.R;
#pragma unroll
    for (int r = 0; r < KABSCH_PER_THREAD; r++) {
        if (tid + NT * r >= n) break;
        const double x0 = qx[r][0] - xc[0], x1 = qx[r][1] - xc[1], x2 = qx[r][2] - xc[2];
        const double y[3] = {(double)qy[r][0] - yc[0], (double)qy[r][1] - yc[1], (double)qy[r][2] - yc[2]};
        const double r0 = R[0] * y[0] + R[1] * y[1] + R[2] * y[2];
        const double r1 = R[3] * y[0] + R[4] * y[1] + R[5] * y[2];
        const double r2 = R[6] * y[0] + R[7] * y[1] + R[8] * y[2];
        store_row_at(v, base + rp[r], (x0 - r0) * sc, (x1 - r1) * sc, (x2 - r2) * sc);
    }
}

template <int NT, int REG>
__device__ void group_item(const StepView& v, const GroupCVItem& it, const int* __restrict__ slots,
                           const float* __restrict__ mass, const float* __restrict__ ref, int w)
{
    __shared__ double sh[17 * 16];
    __shared__ double shR[9];
    const WalkerBox wb = load_box(v, w);
    const double* bx = wb.b;
    double* cvv = v.cvVal + (size_t)w * v.V;
    const int tid = threadIdx.x;
    constexpr int nt = NT;

    if (it.type == CV_GYRATION || it.type == CV_COM_DISTANCE) {
        const int n = it.end - it.begin;
        const int n1 = it.type == CV_COM_DISTANCE ? it.split : n;
        // atoms of this thread: j = tid + r * nt; the first REG stay in
        // registers (one gather round for every pass when n <= REG * nt)
        float4 q[REG];
        double mj[REG];
        size_t rr[REG];
#pragma unroll
        for (int r = 0; r < REG; r++) {  // clamped index: no divergent loads
            const int j = min(tid + r * nt, n - 1);
            q[r] = gather_pos(v, slots, v.gStride, w, it.begin + j);
            mj[r] = mass[it.begin + j];
            rr[r] = row_index(v, w, it.row0 + j);
        }
        // every atom of this thread: registers first (static indices), then
        // the tail past REG * nt with fresh gathers
        auto visit = [&](auto&& f) {
#pragma unroll
            for (int r = 0; r < REG; r++) {
                const int j = tid + r * nt;
                if (j < n) f(j, q[r], mj[r], rr[r]);
            }
            for (int j = tid + REG * nt; j < n; j += nt)
                f(j, gather_pos(v, slots, v.gStride, w, it.begin + j), (double)mass[it.begin + j],
                  row_index(v, w, it.row0 + j));
        };
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // m*x (g1), M1, m*x (g2), M2
        visit([&](int j, const float4& qq, double m, size_t) {
            const bool g1 = j < n1;  // static indices only (a runtime index would spill acc)
            const double m1 = g1 ? m : 0.0, m2 = g1 ? 0.0 : m;
            acc[0] += m1 * qq.x;
            acc[1] += m1 * qq.y;
            acc[2] += m1 * qq.z;
            acc[3] += m1;
            acc[4] += m2 * qq.x;
            acc[5] += m2 * qq.y;
            acc[6] += m2 * qq.z;
            acc[7] += m2;
        });
        block_sum<8, NT / 64>(acc, sh);
        if (it.type == CV_GYRATION) {
            const double M = acc[3];
            const double cx = acc[0] / M, cy = acc[1] / M, cz = acc[2] / M;
            double r2[1] = {0.0};
            visit([&](int j, const float4& qq, double m, size_t) {
                double dx = qq.x - cx, dy = qq.y - cy, dz = qq.z - cz;
                if (v.pbc) pbc_exact(dx, dy, dz, bx);
                r2[0] += m * (dx * dx + dy * dy + dz * dz);
            });
            block_sum<1, NT / 64>(r2, sh);
            const double rg = sqrt(r2[0] / M);
            const double inv = rg > 1e-10 ? 1.0 / rg : 0.0;
            visit([&](int j, const float4& qq, double m, size_t ri) {
                double dx = qq.x - cx, dy = qq.y - cy, dz = qq.z - cz;
                if (v.pbc) pbc_exact(dx, dy, dz, bx);
                const double sc = m / M * inv;
                store_row_at(v, ri, sc * dx, sc * dy, sc * dz);
            });
            if (tid == 0) cvv[it.value] = rg;
        } else {
            double dx = acc[4] / acc[7] - acc[0] / acc[3];
            double dy = acc[5] / acc[7] - acc[1] / acc[3];
            double dz = acc[6] / acc[7] - acc[2] / acc[3];
            if (v.pbc) pbc_exact(dx, dy, dz, bx);
            double d = sqrt(dx * dx + dy * dy + dz * dz);
            double inv = d > 1e-10 ? 1.0 / d : 0.0;
            visit([&](int j, const float4&, double m, size_t ri) {
                double sc = (j < n1 ? -m / acc[3] : m / acc[7]) * inv;
                store_row_at(v, ri, sc * dx, sc * dy, sc * dz);
            });
            if (tid == 0) cvv[it.value] = d;
        }
        return;
    }

    if constexpr (NT > GROUP_BLOCK) {
        // the wide kernel takes no Kabsch item (the planner keeps every fitted
        // RMSD on the 256-thread kernels, whose registers hold the Horn solve)
        if (it.type == CV_RMSD && it.split) return;
    } else {
        if (it.type == CV_RMSD && it.split && it.end - it.begin <= NT * KABSCH_PER_THREAD) {
            kabsch_block<NT>(v, it, slots, ref, w);
            return;
        }
    }
    if (it.type == CV_RMSD) {
        const int n = it.end - it.begin;
        const float* Y = ref + 3 * (size_t)it.refOff;
        const double* yc = it.refCentroid;
        const bool fit = it.split != 0;
        // one gather round: the first REG atoms of this thread stay in
        // registers with their reference (centred for the fit)
        float4 q[REG];
        float yr[REG][3];  // raw FP32 reference; centred on use
        size_t rr[REG];
#pragma unroll
        for (int r = 0; r < REG; r++) {  // clamped index: no divergent loads
            const int j = min(tid + r * nt, n - 1);
            q[r] = gather_pos(v, slots, v.gStride, w, it.begin + j);
#pragma unroll
            for (int k = 0; k < 3; k++) yr[r][k] = Y[3 * j + k];
            rr[r] = row_index(v, w, it.row0 + j);
        }
        // every atom of this thread with its (centred, for the fit) reference:
        // registers first (static indices), then the tail with fresh gathers
        auto visit = [&](auto&& f) {
#pragma unroll
            for (int r = 0; r < REG; r++) {
                const int j = tid + r * nt;
                if (j < n) {
                    const double yy[3] = {(double)yr[r][0] - (fit ? yc[0] : 0.0), (double)yr[r][1] - (fit ? yc[1] : 0.0),
                                          (double)yr[r][2] - (fit ? yc[2] : 0.0)};
                    f(j, q[r], yy, rr[r]);
                }
            }
            for (int j = tid + REG * nt; j < n; j += nt) {
                const double yy[3] = {(double)Y[3 * j] - (fit ? yc[0] : 0.0), (double)Y[3 * j + 1] - (fit ? yc[1] : 0.0),
                                      (double)Y[3 * j + 2] - (fit ? yc[2] : 0.0)};
                f(j, gather_pos(v, slots, v.gStride, w, it.begin + j), yy, row_index(v, w, it.row0 + j));
            }
        };
        if (!fit) {
            double s1[1] = {0.0};
            visit([&](int j, const float4& qq, const double* yy, size_t) {
                const double dx = (double)qq.x - yy[0], dy = (double)qq.y - yy[1], dz = (double)qq.z - yy[2];
                s1[0] += dx * dx + dy * dy + dz * dz;
            });
            block_sum<1, NT / 64>(s1, sh);
            const double rv = sqrt(s1[0] / n);
            const double sc = (rv > 1e-10 ? 1.0 / rv : 0.0) / n;
            visit([&](int j, const float4& qq, const double* yy, size_t ri) {
                const double dx = (double)qq.x - yy[0], dy = (double)qq.y - yy[1], dz = (double)qq.z - yy[2];
                store_row_at(v, ri, dx * sc, dy * sc, dz * sc);
            });
            if (tid == 0) cvv[it.value] = rv;
            return;
        }
        if constexpr (NT > GROUP_BLOCK) return;
        // Kabsch: one block reduction of the raw moments
        //   m[3a+b] = sum y'_a x_b, m[9..11] = sum x, m[12] = sum |x|^2   (y' = y - ybar),
        // then the centred covariance sum y'(x - xbar)^T = m[3a+b] - (sum y'_a) xbar_b
        // and sum |x - xbar|^2 = sum |x|^2 - n |xbar|^2, both in FP64; sum y' and
        // sum |y'|^2 are the host's (the reference does not move).
        double m[KABSCH_MOMENTS];
#pragma unroll
        for (int k = 0; k < KABSCH_MOMENTS; k++) m[k] = 0.0;
        visit([&](int j, const float4& qq, const double* y, size_t) {
            const double x[3] = {(double)qq.x, (double)qq.y, (double)qq.z};
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) m[3 * a + b] += y[a] * x[b];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                m[9 + k] += x[k];
                m[12] += x[k] * x[k];
            }
        });
        block_sum<KABSCH_MOMENTS, NT / 64>(m, sh);
        const double xc[3] = {m[9] / n, m[10] / n, m[11] / n};
        __shared__ double shScale;
        if (tid == 0) {
            double S[9], R[9], l;
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) S[3 * a + b] = m[3 * a + b] - it.refSum[a] * xc[b];
            const double Gx = m[12] - n * (xc[0] * xc[0] + xc[1] * xc[1] + xc[2] * xc[2]);
            horn_rotation(S, Gx, it.refSq, R, &l);
            double msd = (Gx + it.refSq - 2.0 * l) / n;
            if (msd < 0.0) msd = 0.0;
            const double rv = sqrt(msd);
            for (int k = 0; k < 9; k++) shR[k] = R[k];
            shScale = (rv > 1e-10 ? 1.0 / rv : 0.0) / n;
            cvv[it.value] = rv;
        }
        __syncthreads();
        const double sc = shScale;
        visit([&](int j, const float4& qq, const double* y, size_t ri) {
            const double x0 = qq.x - xc[0], x1 = qq.y - xc[1], x2 = qq.z - xc[2];
            const double r0 = shR[0] * y[0] + shR[1] * y[1] + shR[2] * y[2];
            const double r1 = shR[3] * y[0] + shR[4] * y[1] + shR[5] * y[2];
            const double r2 = shR[6] * y[0] + shR[7] * y[1] + shR[8] * y[2];
            store_row_at(v, ri, (x0 - r0) * sc, (x1 - r1) * sc, (x2 - r2) * sc);
        });
        return;
    }

    // Pair CVs (DRMSD, contact map): the first PR pairs of this thread, with
    // their per-pair parameters, are gathered in one round before any
    // arithmetic, so a thread waits for memory once rather than once per
    // pair; pairs past PR * nt are gathered on use.  Pairs are visited in the
    // same order either way (p = tid, tid + nt, ...), so the sums do not
    // depend on PR.
    if (it.type != CV_DRMSD && it.type != CV_CONTACTMAP) return;
    constexpr int PR = REG / 2 > 0 ? REG / 2 : 1;
    const int np = (it.end - it.begin) / 2;  // >= 1 (validated)
    const float* prm = ref + it.refOff;       // DRMSD: d0 per pair; contact map: r0, n, m, w
    const bool cm = it.type == CV_CONTACTMAP;
    float4 pa[PR], pb[PR], pp[PR];
    size_t ra[PR], rb[PR];
#pragma unroll
    for (int r = 0; r < PR; r++) {  // clamped index: no divergent loads
        const int p = min(tid + r * nt, np - 1);
        pa[r] = gather_pos(v, slots, v.gStride, w, it.begin + 2 * p);
        pb[r] = gather_pos(v, slots, v.gStride, w, it.begin + 2 * p + 1);
        ra[r] = row_index(v, w, it.row0 + 2 * p);
        rb[r] = row_index(v, w, it.row0 + 2 * p + 1);
        const float* pq = prm + (cm ? 4 * p : p);  // DRMSD reads its one value four times
        pp[r] = make_float4(pq[0], pq[cm ? 1 : 0], pq[cm ? 2 : 0], pq[cm ? 3 : 0]);
    }
    auto visit_pairs = [&](auto&& f) {
#pragma unroll
        for (int r = 0; r < PR; r++) {
            const int p = tid + r * nt;
            if (p < np) f(p, pa[r], pb[r], pp[r], ra[r], rb[r]);
        }
        for (int p = tid + PR * nt; p < np; p += nt) {
            const float4 par = cm ? make_float4(prm[4 * p], prm[4 * p + 1], prm[4 * p + 2], prm[4 * p + 3])
                                  : make_float4(prm[p], 0.f, 0.f, 0.f);
            f(p, gather_pos(v, slots, v.gStride, w, it.begin + 2 * p),
              gather_pos(v, slots, v.gStride, w, it.begin + 2 * p + 1), par, row_index(v, w, it.row0 + 2 * p),
              row_index(v, w, it.row0 + 2 * p + 1));
        }
    };

    if (it.type == CV_DRMSD) {
        double s[1] = {0.0};
        visit_pairs([&](int p, const float4& a, const float4& b, const float4& par, size_t, size_t) {
            double dx = (double)b.x - a.x, dy = (double)b.y - a.y, dz = (double)b.z - a.z;
            if (v.pbc) pbc_exact(dx, dy, dz, bx);
            double dev = sqrt(dx * dx + dy * dy + dz * dz) - (double)par.x;
            s[0] += dev * dev;
        });
        block_sum<1, NT / 64>(s, sh);
        const double val = sqrt(s[0] / np);
        const double inv = val > 1e-10 ? 1.0 / val : 0.0;
        visit_pairs([&](int p, const float4& a, const float4& b, const float4& par, size_t ia, size_t ib) {
            double dx = (double)b.x - a.x, dy = (double)b.y - a.y, dz = (double)b.z - a.z;
            if (v.pbc) pbc_exact(dx, dy, dz, bx);
            double d = sqrt(dx * dx + dy * dy + dz * dz);
            double coeff = (d - (double)par.x) / np * inv * (d > 1e-10 ? 1.0 / d : 0.0);
            store_row_at(v, ia, -coeff * dx, -coeff * dy, -coeff * dz);
            store_row_at(v, ib, coeff * dx, coeff * dy, coeff * dz);
        });
        if (tid == 0) cvv[it.value] = val;
        return;
    }

    if (it.type == CV_CONTACTMAP) {
        double s[1] = {0.0};
        visit_pairs([&](int p, const float4& a, const float4& b, const float4& par, size_t ia, size_t ib) {
            double dx = (double)b.x - a.x, dy = (double)b.y - a.y, dz = (double)b.z - a.z;
            if (v.pbc) pbc_exact(dx, dy, dz, bx);
            double coeff;
            s[0] += contact_pair(dx, dy, dz, par, coeff);
            store_row_at(v, ia, -coeff * dx, -coeff * dy, -coeff * dz);
            store_row_at(v, ib, coeff * dx, coeff * dy, coeff * dz);
        });
        block_sum<1, NT / 64>(s, sh);
        if (tid == 0) cvv[it.value] = s[0];
        return;
    }
}

__global__ __launch_bounds__(GROUP_BLOCK) void k_cv_group(StepView v, const GroupCVItem* __restrict__ items,
                                                          const int* __restrict__ slots,
                                                          const float* __restrict__ mass,
                                                          const float* __restrict__ ref, int w0)
{
    group_item<GROUP_BLOCK, GROUP_REG>(v, items[blockIdx.x], slots, mass, ref, w0 + blockIdx.y);
}

void launch_cv_group(const StepView& v, const GroupCVItem* items, int n, const int* slots, const float* mass,
                     const float* ref, int w0, int nw, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cv_group, dim3(n, nw), dim3(GROUP_BLOCK), 0, s, v, items, slots, mass, ref, w0);
}

// Group CVs over many atoms or pairs (more than GROUP_BLOCK * GROUP_REG): one
// workgroup each, 8 atoms per thread in registers.  Its size is fixed per
// plan (1,024 threads, or 512 when the batched step carries these items in
// the 512-thread path-row launch), so that every form runs the same code and
// gives the same bits.
constexpr int GROUP_BIG_BLOCK = 1024;
constexpr int GROUP_BIG_BLOCK_PATH = 512;
constexpr int GROUP_BIG_REG = 8;

template <int NT>
__global__ __launch_bounds__(NT) void k_cv_group_big(StepView v, const GroupCVItem* __restrict__ items,
                                                     const int* __restrict__ slots, const float* __restrict__ mass,
                                                     const float* __restrict__ ref, int w0)
{
    group_item<NT, GROUP_BIG_REG>(v, items[blockIdx.x], slots, mass, ref, w0 + blockIdx.y);
}

void launch_cv_group_big(const StepView& v, const GroupCVItem* items, int n, const int* slots, const float* mass,
                         const float* ref, int w0, int nw, int nt, hipStream_t s)
{
    if (n <= 0) return;
    if (nt == GROUP_BIG_BLOCK_PATH)
        hipLaunchKernelGGL(k_cv_group_big<GROUP_BIG_BLOCK_PATH>, dim3(n, nw), dim3(GROUP_BIG_BLOCK_PATH), 0, s, v,
                           items, slots, mass, ref, w0);
    else
        hipLaunchKernelGGL(k_cv_group_big<GROUP_BIG_BLOCK>, dim3(n, nw), dim3(GROUP_BIG_BLOCK), 0, s, v, items,
                           slots, mass, ref, w0);
}

// ---------------------------------------------------------------------------
// Coordination, CommonGluedKernels.cpp:899-1031, with an optional cutoff.
//
// Cell list per group (cells of edge >= cutoff, at least 3 per dimension,
// orthorhombic box), built by a counting sort of both groups into cell order
// (k_coord_rank / k_coord_place; the zero / bin / scan /
// fill kernels are the fallback for cell grids too large for LDS).
// Without a usable cell list the partner group is scanned in chunks.
//
// Pair passes: pass A (home = group A atom) forms CN and the A rows, pass B
// (home = group B atom) the B rows; every in-cutoff pair is visited once per
// pass, so each atom's row is a sum owned by one wave (no atomics on rows).
// Within a wave, lanes scan their own neighbour cell with an FP32 pre-check
// (1e-5 relative margin); the candidates are queued in LDS and evaluated
// densely, 64 at a time, on the exact FP64 predicate r^2 < rc^2 (as the
// oracle) and the FP64 switching function.  CN, pair count and gradients are
// accumulated as 2^32 fixed-point integers, so the result does not depend on
// the order atoms were placed in a cell, on the queue order or on the wave
// layout: the sums are exact integer sums of rounded terms, bit-reproducible.
// ---------------------------------------------------------------------------

__device__ inline int cell_of(double x, double y, double z, const double* bx, const int* nc)
{
    double f[3] = {x * bx[9], y * bx[10], z * bx[11]};
    int c[3];
    for (int k = 0; k < 3; k++) {
        double t = f[k] - floor(f[k]);
        int ci = (int)(t * nc[k]);
        c[k] = ci < 0 ? 0 : (ci >= nc[k] ? nc[k] - 1 : ci);
    }
    return (c[2] * nc[1] + c[1]) * nc[0] + c[0];
}

__global__ void k_coord_zero(CoordBuffers cb, int base, int n, int w0)
{
    const int w = w0 + blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        cb.cellCount[(size_t)w * cb.cellStride + base + i] = 0;
}

__global__ void k_coord_bin(StepView v, CoordBuffers cb, const CoordItem* items, int item, int w0)
{
    const int w = w0 + blockIdx.y;
    const CoordItem it = items[item];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = it.nA + it.nB;
    if (i >= n) return;
    const int slot = cb.slots[(i < it.nA) ? it.aBegin + i : it.bBegin + (i - it.nA)];
    float4 q = load_pos(v, w, slot);
    const int ncell = it.nc[0] * it.nc[1] * it.nc[2];
    int c = cell_of(q.x, q.y, q.z, v.box + (size_t)w * 12, it.nc) + (i < it.nA ? 0 : ncell);
    cb.cellOf[(size_t)w * cb.cellOfStride + it.sortBase + i] = c;
    atomicAdd(&cb.cellCount[(size_t)w * cb.cellStride + it.cellBase + c], 1);
}

// exclusive scan of the 2*ncell counts of one item; one workgroup per walker
__global__ void k_coord_scan(CoordBuffers cb, const CoordItem* items, int item, int w0)
{
    __shared__ int sh[1024];
    const int w = w0 + blockIdx.y;
    const CoordItem it = items[item];
    const int n = 2 * it.nc[0] * it.nc[1] * it.nc[2];
    int* cnt = cb.cellCount + (size_t)w * cb.cellStride + it.cellBase;
    int* st = cb.cellStart + (size_t)w * (cb.cellStride + cb.nItems) + it.cellBase + item;
    int carry = 0;
    for (int base = 0; base < n; base += blockDim.x) {
        int i = base + threadIdx.x;
        int x = i < n ? cnt[i] : 0;
        sh[threadIdx.x] = x;
        __syncthreads();
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {
            int t = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < n) {
            st[i] = carry + sh[threadIdx.x] - x;
            cnt[i] = 0;  // reused as the fill cursor
        }
        int tot = sh[blockDim.x - 1];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) st[n] = carry;
}

__global__ void k_coord_fill(StepView v, CoordBuffers cb, const CoordItem* items, int item, int w0)
{
    const int w = w0 + blockIdx.y;
    const CoordItem it = items[item];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = it.nA + it.nB;
    if (i >= n) return;
    const int slot = cb.slots[(i < it.nA) ? it.aBegin + i : it.bBegin + (i - it.nA)];
    float4 q = load_pos(v, w, slot);
    const int c = cb.cellOf[(size_t)w * cb.cellOfStride + it.sortBase + i];
    const int* st = cb.cellStart + (size_t)w * (cb.cellStride + cb.nItems) + it.cellBase + item;
    int pos = st[c] + atomicAdd(&cb.cellCount[(size_t)w * cb.cellStride + it.cellBase + c], 1);
    const int local = i < it.nA ? i : i - it.nA;
    const int rp = v.rowPos[(i < it.nA ? it.rowA : it.rowB) + local];
    cb.sorted[(size_t)w * cb.sortStride + it.sortBase + pos] = make_float4(q.x, q.y, q.z, __int_as_float(rp));
}

constexpr int CELL_BLOCK = 1024;
constexpr int CELL_LDS = 8192;  // cells (both groups) counted in LDS by the sort

// Two-kernel counting sort (no global atomics, no single-CU scatter), stable:
// within a cell the atoms keep their list order, so the sorted array -- and
// with it the order in which the pair kernels meet the pairs -- is the same
// on every evaluation.
//   rank    one 1,024-thread workgroup per 1,024 atoms: gather, cell, rank of
//           the atom among the workgroup's earlier atoms of that cell (within
//           a wave from a ballot over the lanes of the same cell; across
//           waves from LDS counters that the waves advance in wave order),
//           and the workgroup's count per cell (cell-major table);
//   place   every workgroup scans the table (cell-major, workgroup-minor) for
//           where its atoms of each cell start, and places its atoms.
int coord_blocks(int n) { return (n + CELL_BLOCK - 1) / CELL_BLOCK; }
int coord_sort_fits(int ncellBoth, int n) { return ncellBoth <= CELL_LDS && ncellBoth * coord_blocks(n) <= 65536; }

constexpr int CELL_WAVE_TAB = 2048;  // cells (both groups) up to which the waves' counts go to a per-wave table

// The rank launch may carry the wide group CVs of the step as ge.n extra
// workgroups after the ranking ones (they share no data with the sort, and so
// take no launch of their own in the batched step).
static_assert(CELL_BLOCK == GROUP_BIG_BLOCK, "wide group CVs ride in the rank launch");

__global__ __launch_bounds__(CELL_BLOCK) void k_coord_rank(StepView v, CoordBuffers cb, CoordItem it, int item,
                                                           int w0, GroupExtra ge)
{
    constexpr int WAVES = CELL_BLOCK / 64;
    __shared__ int lcnt[CELL_LDS];
    __shared__ unsigned short wcnt[WAVES][CELL_WAVE_TAB];  // per wave, its atoms per cell
    const int nb = gridDim.x - ge.n;
    if ((int)blockIdx.x >= nb) {
        group_item<GROUP_BIG_BLOCK, GROUP_BIG_REG>(v, ge.items[blockIdx.x - nb], ge.slots, ge.mass, ge.ref,
                                                    w0 + blockIdx.y);
        return;
    }
    const int w = w0 + blockIdx.y, tid = threadIdx.x, b = blockIdx.x;
    const int n = it.nA + it.nB;
    const int ncell = it.nc[0] * it.nc[1] * it.nc[2], n2 = 2 * ncell;
    const bool table = n2 <= CELL_WAVE_TAB;  // uniform
    const int i = b * CELL_BLOCK + tid, lane = tid & 63, wv = tid >> 6;
    if (table) {
        // each wave clears its own row (its LDS operations complete in order)
        for (int k = lane; k < n2; k += 64) wcnt[wv][k] = 0;
    } else {
        for (int c = tid; c < n2; c += CELL_BLOCK) lcnt[c] = 0;
        __syncthreads();
    }
    const bool valid = i < n;
    int c = -1;
    const size_t o = (size_t)w * cb.sortStride + it.sortBase + i;
    if (valid) {
        const bool a = i < it.nA;
        const int rp = v.rowPos[(a ? it.rowA + i : it.rowB + (i - it.nA))];
        const float4 q = v.posq[cb.pos[(size_t)w * cb.posStride + (a ? it.aBegin + i : it.bBegin + (i - it.nA))]];
        c = cell_of(q.x, q.y, q.z, v.box + (size_t)w * 12, it.nc) + (a ? 0 : ncell);
        cb.cellOf[(size_t)w * cb.cellOfStride + it.sortBase + i] = c;
        // w: the atom's Jacobian row position (the table load does not wait for the gather)
        cb.unsorted[o] = make_float4(q.x, q.y, q.z, __int_as_float(rp));
    }
    // rank among the wave's lower lanes of the same cell, and the group size
    // (each lane compares its cell with every lane's; the lowest lane of a
    // group is its leader)
    int inWave = 0, group = 0;
#pragma unroll 16
    for (int j = 0; j < 64; j++) {
        const int cj = __builtin_amdgcn_readlane(c, j);
        inWave += (cj == c) & (j < lane);
        group += cj == c;
    }
    const bool leader = inWave == 0;
    int rank = 0;
    int* tab = cb.blockCnt + (size_t)w * cb.blkStride + it.blkBase;
    if (table) {
        // every wave's group sizes, then each atom adds the counts of the
        // earlier waves of its cell
        if (valid && leader) wcnt[wv][c] = (unsigned short)group;
        __syncthreads();
        if (valid) {
            int base = 0;
            for (int k = 0; k < wv; k++) base += wcnt[k][c];
            rank = base + inWave;
            cb.rank[o] = rank;
        }
        for (int cc = tid; cc < n2; cc += CELL_BLOCK) {
            int t = 0;
#pragma unroll
            for (int k = 0; k < WAVES; k++) t += wcnt[k][cc];
            tab[(size_t)cc * nb + b] = t;
        }
        return;
    }
    // many cells: the waves take their cells' counters in wave order
    for (int k = 0; k < WAVES; k++) {
        __syncthreads();
        if (wv == k && valid) {
            const int base = lcnt[c];
            __builtin_amdgcn_wave_barrier();
            if (leader) lcnt[c] = base + group;
            rank = base + inWave;
        }
    }
    if (valid) cb.rank[o] = rank;
    __syncthreads();
    for (int cc = tid; cc < n2; cc += CELL_BLOCK) tab[(size_t)cc * nb + b] = lcnt[cc];
}

// Placement with the scan folded in: every placing workgroup scans the whole
// (cell, workgroup) count table itself (cell-major, workgroup-minor; a few
// thousand integers, all loads issued before the sums) and keeps the start of
// its own atoms of every cell in LDS; workgroup 0 also writes the cell starts.
// This replaces a one-workgroup scan launch between rank and place.
constexpr int CELL_SCAN_PER = 16;  // table entries per thread held in registers by the scan

__global__ __launch_bounds__(CELL_BLOCK) void k_coord_place(CoordBuffers cb, CoordItem it, int item, int w0)
{
    __shared__ int sOff[CELL_LDS];
    __shared__ int sTab[CELL_BLOCK * CELL_SCAN_PER];
    __shared__ int waveTot[CELL_BLOCK / 64];
    const int nb = gridDim.x;
    const int w = w0 + blockIdx.y, b = blockIdx.x, tid = threadIdx.x;
    const int n = it.nA + it.nB;
    const int ncell = it.nc[0] * it.nc[1] * it.nc[2], n2 = 2 * ncell, len = n2 * nb;
    const int* tab = cb.blockCnt + (size_t)w * cb.blkStride + it.blkBase;
    int* st = cb.cellStart + (size_t)w * (cb.cellStride + cb.nItems) + it.cellBase + item;
    // this atom's rank and cell, loaded before the scan (their round trip overlaps it)
    const int i = b * CELL_BLOCK + tid;
    const size_t o = (size_t)w * cb.sortStride + it.sortBase + min(i, n - 1);
    const int cAtom = cb.cellOf[(size_t)w * cb.cellOfStride + it.sortBase + min(i, n - 1)];
    const int rAtom = cb.rank[o];
    const float4 qAtom = cb.unsorted[o];
    // exclusive scan of the table, in rounds of CELL_BLOCK * CELL_SCAN_PER entries
    int carry = 0;
    for (int r0 = 0; r0 < len; r0 += CELL_BLOCK * CELL_SCAN_PER) {
        const int e0 = r0 + tid * CELL_SCAN_PER;
        // the round's entries staged through LDS with coalesced loads (each
        // thread then takes CELL_SCAN_PER consecutive ones: read directly,
        // every load of a wave would touch 64 different lines)
        {
            int y[CELL_SCAN_PER];
#pragma unroll
            for (int u = 0; u < CELL_SCAN_PER; u++) y[u] = tab[min(r0 + tid + u * CELL_BLOCK, len - 1)];
            __syncthreads();  // the previous round's entries consumed
#pragma unroll
            for (int u = 0; u < CELL_SCAN_PER; u++) sTab[tid + u * CELL_BLOCK] = y[u];
            __syncthreads();
        }
        int x[CELL_SCAN_PER];
#pragma unroll
        for (int k = 0; k < CELL_SCAN_PER; k += 4) {
            const int4 t4 = *reinterpret_cast<const int4*>(&sTab[tid * CELL_SCAN_PER + k]);
            x[k] = t4.x;
            x[k + 1] = t4.y;
            x[k + 2] = t4.z;
            x[k + 3] = t4.w;
        }
        int tsum = 0;
#pragma unroll
        for (int k = 0; k < CELL_SCAN_PER; k++)
            if (e0 + k < len) tsum += x[k];
        const int lane = tid & 63, wid = tid >> 6;
        int incl = tsum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d, 64);
            if (lane >= d) incl += t;
        }
        __syncthreads();  // waveTot of the previous round consumed
        if (lane == 63) waveTot[wid] = incl;
        __syncthreads();
        int run = carry + incl - tsum;
        for (int q = 0; q < wid; q++) run += waveTot[q];
        int roundTot = 0;
        for (int q = 0; q < CELL_BLOCK / 64; q++) roundTot += waveTot[q];
        int c = e0 / nb, bb = e0 - c * nb;  // (cell, workgroup) of entry e0, advanced entry by entry
#pragma unroll
        for (int k = 0; k < CELL_SCAN_PER; k++) {
            const int e = e0 + k;
            if (e < len) {
                if (bb == b) sOff[c] = run;
                if (b == 0 && bb == 0) st[c] = run;  // start of cell c
                run += x[k];
            }
            if (++bb == nb) {
                bb = 0;
                c++;
            }
        }
        carry += roundTot;
    }
    if (b == 0 && tid == 0) st[n2] = carry;
    __syncthreads();
    if (i < n) cb.sorted[(size_t)w * cb.sortStride + it.sortBase + sOff[cAtom] + rAtom] = qAtom;
}

struct SwitchD {
    double r0, rn, rm, ir0;
    double preRes, preDfunc, preSecDev;
    int intPow, in, im;
};

__device__ inline SwitchD make_switch(const CoordItem& it)
{
    SwitchD s;
    s.r0 = it.r0;
    s.rn = it.rn;
    s.rm = it.rm;
    s.ir0 = it.ir0;
    s.preRes = it.preRes;
    s.preDfunc = it.preDfunc;
    s.preSecDev = it.preSecDev;
    s.intPow = it.intPow;
    s.in = (int)s.rn;
    s.im = (int)s.rm;
    return s;
}

// f and df/dr of the rational switch (1-x^n)/(1-x^m), the reference's branch
// order (CommonGluedKernels.cpp:979-1007).  FP64: in FP32 the numerator and
// denominator cancel near x = 1 (1 - x^12 loses ~4 digits at |x-1| = 1e-4),
// which breaks the force tolerance.
__device__ inline void switch_eval(const SwitchD& s, double r, double& f, double& dfdr)
{
    const double x = r * s.ir0;
    const double dx1 = x - 1.0;
    if (fabs(dx1) < 1e-4) {
        f = s.preRes + dx1 * (s.preDfunc + 0.5 * dx1 * s.preSecDev);
        dfdr = (s.preDfunc + dx1 * s.preSecDev) * s.ir0;
    } else if (x < 1e-10) {
        f = 1.0;
        dfdr = 0.0;
    } else {
        double xn, xm;
        if (s.intPow) {
            xn = ipowd(x, s.in);
            xm = ipowd(x, s.im);
        } else {
            xn = pow(x, s.rn);
            xm = pow(x, s.rm);
        }
        const double den = 1.0 - xm, ix = 1.0 / x;
        f = (1.0 - xn) / den;
        const double dfx = (s.rm * xm * ix - s.rn * xn * ix + (s.rn - s.rm) * xn * xm * ix) / (den * den);
        dfdr = dfx * s.ir0;
    }
}

// n = 6, m = 12: (1 - x^6) / (1 - x^12) = 1 / (1 + x^6) for every x (no
// cancellation, no Taylor branch), so from x^2 = r^2 / r0^2 alone:
//   f = 1 / (1 + x^6),  (df/dr) / r = -6 x^4 f^2 / r0^2.
// The reference's formula with its Taylor expansion for |x - 1| < 1e-4
// (CommonGluedKernels.cpp:979-1007) agrees with it to ~1e-12 per pair.
__device__ inline void switch_612(double r2, double ir0sq, double& f, double& s)
{
    const double x2 = r2 * ir0sq;
    const double x4 = x2 * x2;
    f = 1.0 / (1.0 + x4 * x2);
    s = -6.0 * ir0sq * x4 * (f * f);
}

constexpr int COORD_BLOCK = 256;  // 4 waves
constexpr int COORD_PF = 8;       // partner loads in flight per lane
constexpr double COORD_FIX = 1099511627776.0;  // 2^40: gradient terms (|term| < ~1e3, row sums < 2^23)
constexpr double COORD_UNFIX = 1.0 / 1099511627776.0;

__device__ inline long long to_fix(double x) { return __double2ll_rn(x * COORD_FIX); }

// CN terms f in [0, 1] as 128-bit fixed point in units of 2^-64 (hi * 2^64 +
// lo): exact for f >= 2^-11, below that the error is < 2^-64.  Sums of these
// are exact integer sums, independent of order.
struct U128 {
    unsigned long long lo;
    long long hi;
};

__device__ inline void add128(U128& a, unsigned long long lo, long long hi)
{
    const unsigned long long t = a.lo + lo;
    a.hi += hi + (t < lo ? 1 : 0);
    a.lo = t;
}

__device__ inline void add_cn_term(U128& a, double f)
{
    const double ip = floor(f);
    const double frac = f - ip;  // exact
    add128(a, (unsigned long long)(frac * 18446744073709551616.0), (long long)ip);
}

// Pair passes without a cell list (no cutoff, or a box the cell grid does
// not fit): the partner group is scanned in chunks.  SIDE_A: one home atom
// per wave; SIDE_B: two home atoms per wave (lanes 0-31, 32-63).
template <bool SIDE_A>
__global__ __launch_bounds__(COORD_BLOCK) void k_coord_pairs(StepView v, CoordBuffers cb,
                                                            const CoordItem* __restrict__ items, int item, int w0)
{
    constexpr int HPW = SIDE_A ? 1 : 2;   // homes per wave
    constexpr int LPH = 64 / HPW;         // lanes per home
    __shared__ float4 qPos[COORD_BLOCK / 64][128];
    __shared__ int qHome[COORD_BLOCK / 64][128];
    const int w = w0 + blockIdx.y;
    const CoordItem it = items[item];
    const int nHome = SIDE_A ? it.nA : it.nB;
    const int nPart = SIDE_A ? it.nB : it.nA;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int h = lane / LPH, l = lane % LPH;
    if (!SIDE_A && blockIdx.x == 0 && threadIdx.x == 0) {
        // pass A has finished (launch boundary): CN and the pair count to their
        // outputs, accumulators cleared for the next step
        const size_t k = (size_t)w * cb.nItems + item;
        // CN = (H 2^40 + L) 2^-64 = H 2^-24 + L 2^-64
        v.cvVal[(size_t)w * v.V + it.value] =
            (double)cb.cnAcc[k] * 5.9604644775390625e-08 + (double)cb.cnAccLo[k] * 5.421010862427522e-20;
        cb.npairs[k] = (long long)cb.pairAcc[k];
        cb.cnAcc[k] = 0;
        cb.cnAccLo[k] = 0;
        cb.pairAcc[k] = 0;
    }
    const int home0 = (blockIdx.x * (COORD_BLOCK / 64) + wv) * HPW;
    if (home0 >= nHome) return;  // whole wave idle (no workgroup barrier below)
    const int home = home0 + h;
    const bool homeOk = home < nHome;
    const WalkerBox wb = load_box(v, w);
    const double* bx = wb.b;
    const SwitchD sw = make_switch(it);

    // home positions of this wave (both halves), this lane's partner range
    const float4 ph =
        load_pos(v, w, cb.slots[SIDE_A ? it.aBegin + min(home, nHome - 1) : it.bBegin + min(home, nHome - 1)]);
    const int homeIdx = home;
    float4 hp[HPW];
#pragma unroll
    for (int k = 0; k < HPW; k++) {
        const int src = k * LPH;
        hp[k] = make_float4(__shfl(ph.x, src, 64), __shfl(ph.y, src, 64), __shfl(ph.z, src, 64), 0.f);
    }
    int pBegin = 0, pCount = 0;
    if (homeOk) {
        const int chunk = (nPart + LPH - 1) / LPH;
        pBegin = l * chunk;
        pCount = max(0, min(nPart, pBegin + chunk) - pBegin);
    }
    int maxCount = pCount;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) maxCount = max(maxCount, __shfl_xor(maxCount, o, 64));

    long long gx[HPW], gy[HPW], gz[HPW];
#pragma unroll
    for (int k = 0; k < HPW; k++) gx[k] = gy[k] = gz[k] = 0;
    U128 cn{0ull, 0ll};
    unsigned long long np = 0;
    float4* Q = qPos[wv];
    int* QH = qHome[wv];
    unsigned head = 0, tail = 0;

    // FP64 evaluation of queue entry e (exact predicate, reference switching)
    auto evaluate = [&](int e) {
        const float4 pp = Q[e & 127];
        const int hh = QH[e & 127];
        const float4 pH = hh == 0 ? hp[0] : hp[HPW - 1];
        const float4 pa = SIDE_A ? pH : pp;  // dr = r_A - r_B in both passes
        const float4 pb = SIDE_A ? pp : pH;
        double dx = (double)pa.x - (double)pb.x, dy = (double)pa.y - (double)pb.y, dz = (double)pa.z - (double)pb.z;
        if (v.pbc) pbc_exact(dx, dy, dz, bx);
        const double r2 = dot3_exact(dx, dy, dz);
        if (r2 < 1e-20) return;
        if (it.rc2 > 0.0 && !(r2 < it.rc2)) return;
        const double r = sqrt(r2);
        if (r * sw.ir0 > 1000.0) return;
        double f, dfdr;
        switch_eval(sw, r, f, dfdr);
        const double s = (SIDE_A ? 1.0 : -1.0) * dfdr / r;
        const long long qx = to_fix(s * dx), qy = to_fix(s * dy), qz = to_fix(s * dz);
        if (hh == 0) {
            gx[0] += qx;
            gy[0] += qy;
            gz[0] += qz;
        } else {
            gx[HPW - 1] += qx;
            gy[HPW - 1] += qy;
            gz[HPW - 1] += qz;
        }
        if (SIDE_A) {
            add_cn_term(cn, f);
            np++;
        }
    };

    // partners in batches of COORD_PF: every load of a batch in flight before the tests
    for (int k0 = 0; k0 < maxCount; k0 += COORD_PF) {
        float4 pf[COORD_PF];
#pragma unroll
        for (int u = 0; u < COORD_PF; u++) {  // clamped to a valid partner: no divergent loads
            const int j = pCount > 0 ? pBegin + min(k0 + u, pCount - 1) : 0;
            pf[u] = load_pos(v, w, cb.slots[SIDE_A ? it.bBegin + j : it.aBegin + j]);
        }
#pragma unroll
        for (int u = 0; u < COORD_PF; u++) {
            const int k = k0 + u;
            if (k >= maxCount) break;
            const float4 pp = pf[u];
            const bool cand = k < pCount;  // every partner: the exact predicate decides
            const unsigned long long m = __ballot(cand);
            if (cand) {
                const unsigned slot = tail + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
                Q[slot & 127] = pp;
                QH[slot & 127] = h;
            }
            tail += (unsigned)__popcll(m);
            if (tail - head >= 64) {
                __builtin_amdgcn_wave_barrier();
                evaluate((int)(head + lane));
                head += 64;
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < (int)(tail - head)) evaluate((int)(head + lane));

    // integer (exact) wave sums; lane 0 of each home's half writes its row
#pragma unroll
    for (int k = 0; k < HPW; k++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            gx[k] += __shfl_xor(gx[k], o, 64);
            gy[k] += __shfl_xor(gy[k], o, 64);
            gz[k] += __shfl_xor(gz[k], o, 64);
        }
    }
    if (l == 0 && homeOk) {
        const int k = h;
        const long long sx = k == 0 ? gx[0] : gx[HPW - 1], sy = k == 0 ? gy[0] : gy[HPW - 1],
                        sz = k == 0 ? gz[0] : gz[HPW - 1];
        store_row(v, w, (SIDE_A ? it.rowA : it.rowB) + homeIdx, (double)sx * COORD_UNFIX, (double)sy * COORD_UNFIX,
                  (double)sz * COORD_UNFIX);
    }
    if (SIDE_A) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long lo = __shfl_xor(cn.lo, o, 64);
            const long long hi = __shfl_xor(cn.hi, o, 64);
            add128(cn, lo, hi);
            np += __shfl_xor(np, o, 64);
        }
        if (lane == 0) {
            // the wave's 128-bit sum V as (V >> 40) and (V mod 2^40): the low
            // accumulator absorbs 2^24 waves without overflow, so two 64-bit
            // atomics add V exactly
            const size_t k = (size_t)w * cb.nItems + item;
            const unsigned long long L = cn.lo & ((1ull << 40) - 1ull);
            const long long H = (long long)((unsigned long long)cn.hi << 24) + (long long)(cn.lo >> 40);
            atomicAdd(&cb.cnAccLo[k], L);
            atomicAdd((unsigned long long*)&cb.cnAcc[k], (unsigned long long)H);
            atomicAdd(&cb.pairAcc[k], np);
        }
    }
}

// Cell-blocked pair pass (cell list): one workgroup per home
// cell of either group -- workgroups [0, ncell) take the group-A atoms of a
// cell and form CN and the A rows, workgroups [ncell, 2 ncell) the group-B
// atoms and the B rows -- so each pass visits every in-cutoff pair once, and
// both passes are one launch.  A workgroup reads the cell starts of its home
// cell and its 27 neighbour cells of the other group, then stages the homes
// and the neighbours' atoms in LDS with coalesced loads (one round trip;
// more than COORD_PCAP partners or COORD_HCAP homes are taken in chunks).
// A task is one group-A home and a contiguous range of its partner blocks
// (each wave takes an equal share of the cell's (home, block) units), or
// two group-B homes, one per half-wave; lanes pre-check their partners in FP32 (1e-5 relative margin),
// candidates are queued in LDS and evaluated 64 at a time on the exact FP64
// predicate with FP64 switching.  Each lane sums its row terms of a task in
// FP64 in its (fixed) queue order; a DPP wave sum (fixed order) gives the
// task's row, converted once to 2^40 fixed point, and the tasks' integers
// add exactly in the home's LDS accumulator, so rows do not depend on the
// order of the waves.  Each lane's CN terms are summed in FP64 likewise and
// converted once to 128-bit fixed point (units of 2^-64); CN and the pair
// count are summed per workgroup and added with 64-bit atomics; the last
// group-A workgroup of a walker to finish (ticket after a fence) converts
// CN, stores it and the pair count, and clears the accumulators.
// The workgroup size CB (256, 512 or 1,024 threads) is a launch-shape choice
// of the autotuner; every size gives the same pairs and the same exact
// integer sums, the FP64 per-lane partial sums differ in rounding only.
constexpr int COORD_PCAP = 2560;  // partners staged in LDS per chunk
constexpr int COORD_HCAP = 128;   // homes staged in LDS per chunk

// Staged atoms are held twice: the positions as given (the exact FP64
// evaluation), and for the FP32 pre-check their image nearest the home
// cell's centre, relative to that centre (SoA).  A pair within the cutoff
// is within 1.5 cell edges of the centre (the home is within half an edge),
// and a cell edge exceeds the cutoff with at least 3 cells per axis, so the
// nearest images of both atoms are the pair's minimum image: the pre-check
// needs no per-pair image arithmetic.  Two workgroups of 512 threads fit in
// a CU's LDS.
template <int CB>
struct CoordCellShared {
    static constexpr int WAVES = CB / 64;
    float4 part[COORD_PCAP];
    float pw[3][COORD_PCAP];
    float4 home[COORD_HCAP];
    float hw[3][COORD_HCAP];
    long long g[COORD_HCAP][3];
    double cen[3];
    short q[WAVES][128];
    int cellB[27], pre[28];
    int homeB, homeN;
    unsigned long long lo[WAVES], np[WAVES];
    long long hi[WAVES];
};

// One home cell.  SIDE_A: homes are group-A atoms (a task is one home and
// a contiguous range of its partner blocks); otherwise group-B
// atoms, two homes per task, one per half-wave (a B atom has few partners,
// so a task per home would end in a mostly empty evaluation round).
template <bool SIDE_A, int CB>
__device__ void coord_cell(const StepView& v, const CoordBuffers& cb, const CoordItem& it, int item, int w, int hc,
                           CoordCellShared<CB>& sh)
{
    constexpr int COORD_CB = CB, COORD_CB_WAVES = CB / 64;
    constexpr int HPT = SIDE_A ? 1 : 2;  // homes per task
    constexpr int LPH = 64 / HPT;        // lanes per home
    const int nx = it.nc[0], ny = it.nc[1], nz = it.nc[2], ncell = nx * ny * nz;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int* st = cb.cellStart + (size_t)w * (cb.cellStride + cb.nItems) + it.cellBase + item;
    const float4* sorted = cb.sorted + (size_t)w * cb.sortStride + it.sortBase;
    int cnt = 0;  // lanes 0..26 of wave 0: atoms in neighbour cell `lane`
    if (tid < 27) {
        int cx = hc % nx, cy = (hc / nx) % ny, cz = hc / (nx * ny);
        cx = (cx + tid % 3 - 1 + nx) % nx;
        cy = (cy + (tid / 3) % 3 - 1 + ny) % ny;
        cz = (cz + tid / 9 - 1 + nz) % nz;
        const int pc = (cz * ny + cy) * nx + cx + (SIDE_A ? ncell : 0);
        const int b = st[pc], e = st[pc + 1];
        sh.cellB[tid] = b;
        cnt = e - b;
    } else if (tid == 32) {
        const int c = hc + (SIDE_A ? 0 : ncell);
        const int b = st[c], e = st[c + 1];
        sh.homeB = b;
        sh.homeN = e - b;
        sh.pre[0] = 0;
    }
    const WalkerBox wb = load_box(v, w);
    const double* bx = wb.b;
    // the home cell's centre (in LDS: it is needed only while staging) and
    // the nearest image relative to it
    if (tid == 64) {
        const int hcx = hc % nx, hcy = (hc / nx) % ny, hcz = hc / (nx * ny);
        sh.cen[0] = ((double)hcx + 0.5) * bx[0] / (double)nx;
        sh.cen[1] = ((double)hcy + 0.5) * bx[4] / (double)ny;
        sh.cen[2] = ((double)hcz + 0.5) * bx[8] / (double)nz;
    }
    auto rel = [&](float x, int d) -> float {
        double t = (double)x - sh.cen[d];
        if (v.pbc) t -= bx[4 * d] * rint(t * bx[9 + d]);
        return (float)t;
    };
    if (wv == 0) {  // inclusive scan of the 27 counts: pre[c + 1] = atoms in cells 0..c
        int incl = cnt;
#pragma unroll
        for (int d = 1; d < 32; d <<= 1) {
            const int t = __shfl_up(incl, d, 64);
            if (lane >= d) incl += t;
        }
        if (lane < 27) sh.pre[lane + 1] = incl;
    }
    const SwitchD sw = make_switch(it);
    const float rc2f = (float)(it.rc2 * (1.0 + 1e-5));
    __syncthreads();
    const int P = sh.pre[27], H = sh.homeN, homeB = sh.homeB;
    double cnLane = 0.0;  // this lane's CN terms, in its fixed evaluation order
    unsigned long long np = 0;
    short* Q = sh.q[wv];
    static_assert(COORD_HCAP <= 256, "one staged home per thread");
    for (int h0 = 0; h0 < H; h0 += COORD_HCAP) {
        const int hn = min(COORD_HCAP, H - h0);
        for (int p0 = 0; p0 < max(P, 1); p0 += COORD_PCAP) {
            const int pn = min(COORD_PCAP, P - p0);  // 0: no partners (the homes' rows are zero)
            // first chunk: this thread's home too, loaded with the partners
            // and written to LDS behind the same barrier (one round trip)
            // (both defined on every path: nothing is carried across the tasks)
            float4 hq = make_float4(0.f, 0.f, 0.f, 0.f);
            if (p0 == 0) hq = sorted[homeB + h0 + min(tid, hn - 1)];
            // every source index first (neighbour cell of flat index q:
            // pre[c] <= q < pre[c + 1]), then all loads, then the stores
            constexpr int PER = (COORD_PCAP + COORD_CB - 1) / COORD_CB;
            float4 pv[PER];
#pragma unroll
            for (int u = 0; u < PER; u++) pv[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (pn > 0) {
                int src[PER];
#pragma unroll
                for (int u = 0; u < PER; u++) {
                    const int q = p0 + min(tid + u * COORD_CB, pn - 1);
                    int lo = 0, hi = 27;
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (sh.pre[mid] <= q) lo = mid;
                        else hi = mid;
                    }
                    src[u] = sh.cellB[lo] + (q - sh.pre[lo]);
                }
#pragma unroll
                for (int u = 0; u < PER; u++) pv[u] = sorted[src[u]];
            }
            __syncthreads();  // the previous chunk's tasks and row stores are done with LDS
            if (p0 == 0 && tid < hn) {
                sh.home[tid] = hq;
                sh.hw[0][tid] = rel(hq.x, 0);
                sh.hw[1][tid] = rel(hq.y, 1);
                sh.hw[2][tid] = rel(hq.z, 2);
                sh.g[tid][0] = sh.g[tid][1] = sh.g[tid][2] = 0;
            }
            if (pn > 0) {
#pragma unroll
                for (int u = 0; u < PER; u++) {
                    const int i = tid + u * COORD_CB;
                    if (i < pn) {
                        sh.part[i] = pv[u];
                        sh.pw[0][i] = rel(pv[u].x, 0);
                        sh.pw[1][i] = rel(pv[u].y, 1);
                        sh.pw[2][i] = rel(pv[u].z, 2);
                    }
                }
            }
            __syncthreads();
            // tasks: HPT homes and a range of their partner blocks.  Group
            // A: the (home, block) units of the chunk in home-major order,
            // split into one contiguous, equal range per wave, so that the
            // waves finish together and each ends at most a few tasks (one
            // per home it touches) with a partly filled evaluation round.
            // Group B: two homes per task, all of their partner blocks.
            const int nBlk = (pn + LPH - 1) / LPH;
            const int nGroups = (hn + HPT - 1) / HPT;
            auto run_task = [&](int grp, int pbBeg, int pbEnd) {
                const int myHome = grp * HPT + (HPT == 2 ? (lane >> 5) : 0);
                const bool myOk = myHome < hn;
                const int hcl = min(myHome, hn - 1);
                const float hx = sh.hw[0][hcl], hy = sh.hw[1][hcl], hz = sh.hw[2][hcl];
                const float4 hp0 = sh.home[grp * HPT];
                const float4 hp1 = sh.home[min(grp * HPT + HPT - 1, hn - 1)];
                double acc[HPT][3];  // this lane's row terms of the task per home, in queue order
#pragma unroll
                for (int k = 0; k < HPT; k++) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
                unsigned head = 0, tail = 0;
                // FP64 evaluation of queue entry e, dr = r_A - r_B (exact predicate, reference switching)
                auto evaluate = [&](unsigned e) {
                    const int ent = Q[e & 127];
                    const int half = HPT == 2 ? (ent >> 14) : 0;
                    const float4 pp = sh.part[ent & 0x3FFF];
                    const float4 hq = half ? hp1 : hp0;
                    const float4 pa = SIDE_A ? hq : pp;
                    const float4 pb = SIDE_A ? pp : hq;
                    double dx = (double)pa.x - (double)pb.x, dy = (double)pa.y - (double)pb.y,
                           dz = (double)pa.z - (double)pb.z;
                    if (v.pbc) pbc_exact(dx, dy, dz, bx);
                    const double r2 = dot3_exact(dx, dy, dz);
                    if (r2 < 1e-20) return;
                    if (it.rc2 > 0.0 && !(r2 < it.rc2)) return;
                    double f, sc;
                    if (it.rat612) {
                        if (r2 * it.ir0sq > 1e6) return;  // x > 1000
                        switch_612(r2, it.ir0sq, f, sc);
                    } else {
                        const double r = sqrt(r2);
                        if (r * sw.ir0 > 1000.0) return;
                        double dfdr;
                        switch_eval(sw, r, f, dfdr);
                        sc = dfdr / r;
                    }
                    if (!SIDE_A) sc = -sc;
#pragma unroll
                    for (int k = 0; k < HPT; k++) {
                        const double m = k == half ? sc : 0.0;
                        acc[k][0] += m * dx;
                        acc[k][1] += m * dy;
                        acc[k][2] += m * dz;
                    }
                    if (SIDE_A) {
                        cnLane += f;
                        np++;
                    }
                };
                // queue this lane's partner pi when it passes the pre-check;
                // evaluate a full round as soon as 64 entries wait
                auto push = [&](int pi, bool cand) {
                    const unsigned long long m = __ballot(cand);
                    if (cand)
                        Q[(tail + (unsigned)__popcll(m & ((1ull << lane) - 1ull))) & 127] =
                            (short)(pi | (HPT == 2 ? (lane >> 5) << 14 : 0));
                    tail += (unsigned)__popcll(m);
                    if (tail - head >= 64) {
                        __builtin_amdgcn_wave_barrier();
                        evaluate(head + lane);
                        head += 64;
                        __builtin_amdgcn_wave_barrier();
                    }
                };
                auto precheck = [&](int pi, float px, float py, float pz) {
                    const float fx = hx - px, fy = hy - py, fz = hz - pz;
                    return myOk && pi < pn && fx * fx + fy * fy + fz * fz < rc2f;
                };
                for (int pbk = pbBeg; pbk < pbEnd; pbk++) {
                    const int pi = pbk * LPH + (lane % LPH);
                    const int pcl = min(pi, pn - 1);
                    push(pi, precheck(pi, sh.pw[0][pcl], sh.pw[1][pcl], sh.pw[2][pcl]));
                }
                __builtin_amdgcn_wave_barrier();
                if (lane < (int)(tail - head)) evaluate(head + lane);
                // the task's row of each home: a deterministic FP64 wave
                // reduce-scatter of the lanes' sums (lane j holds value j),
                // converted once to 2^40 fixed point;
                // the tasks' integers then add exactly in LDS
                double tot[3 * HPT];
#pragma unroll
                for (int k = 0; k < HPT; k++) {
                    tot[3 * k] = acc[k][0];
                    tot[3 * k + 1] = acc[k][1];
                    tot[3 * k + 2] = acc[k][2];
                }
                const double val = wave_reduce_scatter_small<3 * HPT>(tot);  // lane j: total of value j
                if (lane < 3 * HPT) {
                    const int k = lane / 3, c = lane - 3 * k;
                    const int h = grp * HPT + k;
                    if (h < hn) atomicAdd((unsigned long long*)&sh.g[h][c], (unsigned long long)to_fix(val));
                }
            };
            if (SIDE_A) {
                const int units = hn * nBlk;
                const int u1 = (int)((long long)units * (wv + 1) / COORD_CB_WAVES);
                for (int u = (int)((long long)units * wv / COORD_CB_WAVES); u < u1;) {
                    const int grp = u / nBlk, b = u - grp * nBlk, e = min(nBlk, b + (u1 - u));
                    run_task(grp, b, e);
                    u += e - b;
                }
            } else {
                for (int t = wv; t < nGroups; t += COORD_CB_WAVES) run_task(t, 0, nBlk);
            }
        }
        __syncthreads();
        for (int i = tid; i < hn; i += COORD_CB)  // w of a staged atom: its Jacobian row position
            store_row_at(v, (size_t)w * v.R + __float_as_int(sh.home[i].w), (double)sh.g[i][0] * COORD_UNFIX,
                         (double)sh.g[i][1] * COORD_UNFIX, (double)sh.g[i][2] * COORD_UNFIX);
    }
    if (!SIDE_A) return;
    U128 cn{0ull, 0ll};
    add_cn_term(cn, cnLane);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long lo = __shfl_xor(cn.lo, o, 64);
        const long long hi = __shfl_xor(cn.hi, o, 64);
        add128(cn, lo, hi);
        np += __shfl_xor(np, o, 64);
    }
    if (lane == 0) {
        sh.lo[wv] = cn.lo;
        sh.hi[wv] = cn.hi;
        sh.np[wv] = np;
    }
    __syncthreads();
    if (tid == 0) {
        U128 t{0ull, 0ll};
        unsigned long long tn = 0;
        for (int q = 0; q < COORD_CB_WAVES; q++) {
            add128(t, sh.lo[q], sh.hi[q]);
            tn += sh.np[q];
        }
        // the workgroup's 128-bit CN sum V as (V >> 40) and (V mod 2^40): the
        // low accumulator absorbs 2^24 workgroups without overflow, so two
        // 64-bit atomics add V exactly
        const size_t k = (size_t)w * cb.nItems + item;
        atomicAdd(&cb.cnAccLo[k], t.lo & ((1ull << 40) - 1ull));
        atomicAdd((unsigned long long*)&cb.cnAcc[k],
                  (unsigned long long)((long long)((unsigned long long)t.hi << 24) + (long long)(t.lo >> 40)));
        atomicAdd(&cb.pairAcc[k], tn);
        __threadfence();
        if (atomicAdd(&cb.ticket[k], 1u) == (unsigned)ncell - 1u) {
            // every group-A workgroup of this walker has added its sums
            __threadfence();
            const long long Hc = (long long)atomicAdd((unsigned long long*)&cb.cnAcc[k], 0ull);
            const unsigned long long Lc = atomicAdd(&cb.cnAccLo[k], 0ull);
            const unsigned long long Pc = atomicAdd(&cb.pairAcc[k], 0ull);
            // CN = (H 2^40 + L) 2^-64 = H 2^-24 + L 2^-64
            v.cvVal[(size_t)w * v.V + it.value] = (double)Hc * 5.9604644775390625e-08 + (double)Lc * 5.421010862427522e-20;
            cb.npairs[k] = (long long)Pc;
            atomicExch((unsigned long long*)&cb.cnAcc[k], 0ull);
            atomicExch(&cb.cnAccLo[k], 0ull);
            atomicExch(&cb.pairAcc[k], 0ull);
            atomicExch(&cb.ticket[k], 0u);
        }
    }
}

template <int CB>
__global__ __launch_bounds__(CB) __attribute__((amdgpu_waves_per_eu(CB == 256 ? 2 : 4))) void k_coord_cells(StepView v, CoordBuffers cb, CoordItem it, int item, int w0)
{
    __shared__ CoordCellShared<CB> sh;
    const int w = w0 + blockIdx.y;
    const int ncell = it.nc[0] * it.nc[1] * it.nc[2];
    if ((int)blockIdx.x < ncell) coord_cell<true, CB>(v, cb, it, item, w, blockIdx.x, sh);
    else coord_cell<false, CB>(v, cb, it, item, w, blockIdx.x - ncell, sh);
}

bool launch_coordination(const StepView& v, const CoordBuffers& cb, const CoordItem* dItems, const CoordItem& hIt,
                         int item, int w0, int nw, hipStream_t s, const GroupExtra* extra)
{
    const int n = hIt.nA + hIt.nB;
    bool carried = false;
    if (hIt.useCells) {
        const int ncell2 = 2 * hIt.nc[0] * hIt.nc[1] * hIt.nc[2];
        if (coord_sort_fits(ncell2, n)) {
            const int nb = coord_blocks(n);
            GroupExtra ge{};
            if (extra && extra->n > 0) {
                ge = *extra;
                carried = true;
            }
            hipLaunchKernelGGL(k_coord_rank, dim3(nb + ge.n, nw), dim3(CELL_BLOCK), 0, s, v, cb, hIt, item, w0, ge);
            hipLaunchKernelGGL(k_coord_place, dim3(nb, nw), dim3(CELL_BLOCK), 0, s, cb, hIt, item, w0);
        } else {
            hipLaunchKernelGGL(k_coord_zero, dim3((ncell2 + 255) / 256, nw), dim3(256), 0, s, cb, hIt.cellBase,
                               ncell2, w0);
            hipLaunchKernelGGL(k_coord_bin, dim3((n + 255) / 256, nw), dim3(256), 0, s, v, cb, dItems, item, w0);
            hipLaunchKernelGGL(k_coord_scan, dim3(1, nw), dim3(1024), 0, s, cb, dItems, item, w0);
            hipLaunchKernelGGL(k_coord_fill, dim3((n + 255) / 256, nw), dim3(256), 0, s, v, cb, dItems, item, w0);
        }
    }
    const int wavesA = hIt.nA, wavesB = (hIt.nB + 1) / 2;
    const dim3 gA((wavesA + 3) / 4, nw), gB((wavesB + 3) / 4, nw);
    if (hIt.useCells) {
        const int ncell = hIt.nc[0] * hIt.nc[1] * hIt.nc[2];
        const dim3 g(2 * ncell, nw);
        if (hIt.cellBlock == 256)
            hipLaunchKernelGGL(k_coord_cells<256>, g, dim3(256), 0, s, v, cb, hIt, item, w0);
        else if (hIt.cellBlock == 1024)
            hipLaunchKernelGGL(k_coord_cells<1024>, g, dim3(1024), 0, s, v, cb, hIt, item, w0);
        else
            hipLaunchKernelGGL(k_coord_cells<512>, g, dim3(512), 0, s, v, cb, hIt, item, w0);
    } else {
        hipLaunchKernelGGL((k_coord_pairs<true>), gA, dim3(COORD_BLOCK), 0, s, v, cb, dItems, item, w0);
        hipLaunchKernelGGL((k_coord_pairs<false>), gB, dim3(COORD_BLOCK), 0, s, v, cb, dItems, item, w0);
    }
    return carried;
}

}  // namespace glued

namespace glued {

// ---------------------------------------------------------------------------
// Path CV (s, z), MSD metric, frames 1-indexed, no PBC:
// CommonGluedKernels.cpp:667-794, in two phases so that a path of many frames
// and atoms spreads over the whole chip:
//   MSD blocks  msd_wf = (|x'_w|^2 - 2 x'_w . r'_f + |r'_f|^2) / M on
//               coordinates centred on the path's mean frame c: x' = x - c
//               and r' = r - c are formed in FP64 from the FP32 values, where
//               they are exact; |r'_f|^2 comes from the host.  The cross term
//               is a contraction over K = 3M on the matrix cores: one block
//               takes a tile of PATH_TW walkers x PATH_TF frames and one K
//               slice, stages positions, frames and centre in LDS per chunk
//               of PATH_KC values (FP32, as stored), and accumulates with
//               v_mfma_f64_16x16x4f64 (exact products of the FP32-valued
//               operands, FP64 sums); the four waves' partials are added in
//               wave order and the slice's partial goes to a [W][slices][N]
//               buffer (with |x'|^2 per slice).  Every output element has its
//               own fixed accumulation order, so a walker's partials do not
//               depend on the other walkers of its tile (the one-walker launch
//               and the batched one give the same bits).  FP32 matrix cores
//               were measured first: their rounded sums left s and z ~5e-8
//               off, which a narrow hill width turns into 1e-4 gradient
//               errors.
//   row blocks  (walker, PATH_TILE atoms): every block rebuilds msd_f (slice
//               partials added in slice order), w_f, S and s with the same
//               fixed-order reduction (so all blocks agree bit for bit); lane
//               = atom, wave = frame phase; the per-phase partial rows are
//               summed in LDS in phase order.  x - r is formed in FP64 (exact
//               for FP32 inputs); sums are FP64.  Tile 0 also stores msd_f and
//               writes s, z and d(s, z)/d(lambda):
//   ds/dlambda = -(1/S) sum_f (f+1-s) msd_f w_f,  dz/dlambda = (<msd>_w - z)/lambda.
// With S <= 1e-300 (every frame's weight underflows) s = z = 0 as in the
// reference, and the rows and lambda derivatives are 0 (the reference divides
// by S there).  No frame cap at 128 (:3901); up to PATH_MAX_FRAMES.
// ---------------------------------------------------------------------------
constexpr int PATH_MSD_THREADS = 256;  // threads of an MSD block (k_path_msd, k_cv_fused)
constexpr int PATH_ROWS_UNROLL = 8;    // frames per lane whose loads are issued together (row blocks)
constexpr int PATH_LDS_ROW = PATH_KC + 4;  // padded LDS row of an operand tile (floats)

typedef double path_v4d __attribute__((ext_vector_type(4)));

// One MSD work unit: walker tile [wt0, wt0 + nwt) x frame tile x K slice.
__device__ void path_msd_block(const StepView& v, const PathItem& it, const int* __restrict__ slots,
                               const float* __restrict__ frames, int unit, int wt0, int nwt)
{
    // operand tiles (walker positions, frames: FP32 as stored) and the
    // chunk's centre; the tiles are reused for the FP64 reduction
    __shared__ __align__(16) float sOp[2 * PATH_TW * PATH_LDS_ROW];
    __shared__ float sC[PATH_KC];
    float* sX = sOp;
    float* sR = sOp + PATH_TW * PATH_LDS_ROW;
    double* sRed = reinterpret_cast<double*>(sOp);  // [4][256] wave partials, then [16][32] |x'|^2
    static_assert(sizeof(double) * (4 * 256 + PATH_TW * 32) <= sizeof(float) * 2 * PATH_TW * PATH_LDS_ROW,
                  "the reduction buffers must fit the operand tiles");
    const int M = it.M, N = it.N, K = 3 * M;
    const int ft = unit % it.frameTiles, sl = unit / it.frameTiles;
    const int f0 = ft * PATH_TF;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const float* C = v.pathCentre + it.centreOff;  // centre [K]
    const float* R = frames + it.frameOff;         // frames [N][K]
    const int* tab = slots + it.begin;
    const int nChunks = (K + PATH_KC - 1) / PATH_KC;
    const int cBegin = sl * it.chunksPerSlice, cEnd = min(nChunks, cBegin + it.chunksPerSlice);
    path_v4d d = {0.0, 0.0, 0.0, 0.0};
    double x2[2] = {0.0, 0.0};  // |x'|^2 of this thread's (walker row, atom) entries
    // staging roles: thread t holds atom (t & 31) of walker rows (t >> 5) and (t >> 5) + 8
    const int sa = t & 31, sr0 = t >> 5;
    // MFMA roles: lane l supplies A[i = l % 16][k = l / 16] (walker row i) and
    // B[k = l / 16][j = l % 16] (frame j) and ends with D[l / 16 + 4 e][l % 16]
    // (the FP64 instruction's output rows are interleaved by 4)
    const int mi = lane & 15, mk = lane >> 4;
    const bool rowOk = mi < nwt, colOk = f0 + mi < N;
    for (int ch = cBegin; ch < cEnd; ch++) {
        const int k0 = ch * PATH_KC, kc = min(PATH_KC, K - k0);
        const int j0 = k0 / 3, na = kc / 3;
        __syncthreads();
        // walkers: table entries, then positions; frames: 16 x kc values,
        // coalesced along K; the centre (clamped: every load issued, none
        // waits on a branch)
        int gi[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int r = min(sr0 + 8 * u, nwt - 1);
            gi[u] = tab[(size_t)(wt0 + r) * v.pStride + j0 + min(sa, na - 1)];
        }
        float4 q[2];
#pragma unroll
        for (int u = 0; u < 2; u++) q[u] = v.posq[gi[u]];
        float rv[PATH_TF * PATH_KC / PATH_MSD_THREADS];
#pragma unroll
        for (int u = 0; u < PATH_TF * PATH_KC / PATH_MSD_THREADS; u++) {
            const int e = t + u * PATH_MSD_THREADS;
            const int fr = e / PATH_KC, k = e % PATH_KC;
            rv[u] = R[(size_t)min(f0 + fr, N - 1) * K + k0 + min(k, kc - 1)];
        }
        const float cv = C[k0 + min(t, kc - 1)];
        const float* cj = C + 3 * (j0 + min(sa, na - 1));
        const double c0 = cj[0], c1 = cj[1], c2 = cj[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int r = sr0 + 8 * u;
            float* dst = sX + r * PATH_LDS_ROW + 3 * sa;
            dst[0] = q[u].x;
            dst[1] = q[u].y;
            dst[2] = q[u].z;
            if (r < nwt && sa < na) {
                const double xa = (double)q[u].x - c0, ya = (double)q[u].y - c1, za = (double)q[u].z - c2;
                x2[u] += xa * xa + ya * ya + za * za;
            }
        }
#pragma unroll
        for (int u = 0; u < PATH_TF * PATH_KC / PATH_MSD_THREADS; u++) {
            const int e = t + u * PATH_MSD_THREADS;
            sR[(e / PATH_KC) * PATH_LDS_ROW + e % PATH_KC] = rv[u];
        }
        if (t < PATH_KC) sC[t] = cv;
        __syncthreads();
        // wave `wave` takes K steps [wave * STEPS, wave * STEPS + STEPS) of
        // the chunk; entries of invalid rows, frames or K are zero operands
        constexpr int STEPS = PATH_KC / 4 / 4;
#pragma unroll
        for (int st = 0; st < STEPS; st++) {
            const int k = (wave * STEPS + st) * 4 + mk;
            const double ck = sC[k];
            const bool kOk = k < kc;
            const double a = (rowOk && kOk) ? (double)sX[mi * PATH_LDS_ROW + k] - ck : 0.0;
            const double b = (colOk && kOk) ? (double)sR[mi * PATH_LDS_ROW + k] - ck : 0.0;
            d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d, 0, 0, 0);
        }
    }
    __syncthreads();  // operand tiles free: reuse as the reduction buffer
#pragma unroll
    for (int e = 0; e < 4; e++) sRed[wave * 256 + ((lane >> 4) + 4 * e) * 16 + (lane & 15)] = d[e];
#pragma unroll
    for (int u = 0; u < 2; u++) sRed[4 * 256 + (sr0 + 8 * u) * 32 + sa] = x2[u];
    __syncthreads();
    double* part = v.pathPart + it.partOff;
    {
        const int e = t, wi = e / 16, fj = e % 16;
        const double cr = sRed[e] + sRed[256 + e] + sRed[512 + e] + sRed[768 + e];
        if (wi < nwt && f0 + fj < N) part[(size_t)(wt0 + wi) * v.partStride + sl * N + f0 + fj] = cr;
    }
    if (ft == 0 && t < nwt) {
        double a = 0.0;
        for (int k = 0; k < 32; k++) a += sRed[4 * 256 + t * 32 + k];
        part[(size_t)(wt0 + t) * v.partStride + it.slices * N + sl] = a;
    }
}

// Per-kind MSD launch: grid (units of every path, walker tiles of the launch).
__global__ __launch_bounds__(PATH_MSD_THREADS) void k_path_msd(StepView v, const PathItem* __restrict__ items, int nItems,
                                                  const int* __restrict__ slots, const float* __restrict__ frames,
                                                  int w0, int nw)
{
    int p = 0;
    while (p + 1 < nItems && (int)blockIdx.x >= items[p + 1].msdBlk0) p++;
    const int wt0 = w0 + blockIdx.y * PATH_TW;
    path_msd_block(v, items[p], slots, frames, blockIdx.x - items[p].msdBlk0, wt0, min(PATH_TW, w0 + nw - wt0));
}

// Contact-map values from the chunk partials an earlier launch of the step
// wrote: wave k adds maps k, k + waves, ... with the additions of the chunks'
// own ticketed form (wave_sum_partials in chunk order), so either form gives
// the same bits.
__device__ void cm_finish(const StepView& v, const GroupExtra& ge, int w)
{
    const int nwv = blockDim.x >> 6;
    for (int q = threadIdx.x >> 6; q < ge.nCm; q += nwv) {
        const GroupCVItem& it = ge.cm[q];
        const int nChunks = ((it.end - it.begin) / 2 + CM_CHUNK - 1) / CM_CHUNK;
        const double tot = wave_sum_partials(ge.cmPart + (size_t)w * ge.nCmBlocks + it.chunk0, nChunks);
        if ((threadIdx.x & 63) == 0) v.cvVal[(size_t)w * v.V + it.value] = tot;
    }
}

// CARRY: the launch also runs the plan's wide group CVs as ge.n extra
// workgroups per walker row (512 threads each; they share no data with the
// rows).  ge.cmBlock >= 0: that workgroup of every walker row (the one after
// the row blocks) adds the contact-map chunk partials instead.  8 phases: at
// most 128 VGPRs, so two workgroups fit a CU and a launch of up to 512
// workgroups is one residency round (the carrying form came to 129 with the
// contact-map branch, and the 288 workgroups of the 16 x 100k-atom plan then
// took two rounds: 15.0 -> 19.6 us).  At most 128 VGPRs: two 8-phase workgroups
// per CU, so a launch of up to 512 workgroups is one residency round (at 129
// the carrying launch of the 16 x 100k-atom plan, 288 workgroups, took a
// second round: 15.0 -> 19.6 us).
template <int PH, bool CARRY>
__global__ __launch_bounds__(64 * PH) __attribute__((amdgpu_waves_per_eu(PH == 8 ? 4 : 1))) void k_path_rows(StepView v, const PathItem* __restrict__ items,
                                                                int nItems, const int* __restrict__ slots,
                                                                const float* __restrict__ frames, int w0,
                                                                GroupExtra ge)
{
    __shared__ double shW[PATH_MAX_FRAMES];
    __shared__ double shRed[6][64 * PH];
    __shared__ double sh[4 * 16];
    static_assert(!CARRY || 64 * PH == GROUP_BIG_BLOCK_PATH, "the carried group CVs run on the launch's block size");
    // blocks past the rows: the contact-map sums (cmBlock), then the group
    // CVs; one branch region for both (two separate branches cost a VGPR
    // and a spill)
    if ((int)blockIdx.x >= (int)gridDim.x - ge.n - (ge.cmBlock >= 0 ? 1 : 0)) {
        if ((int)blockIdx.x == ge.cmBlock) {
            cm_finish(v, ge, w0 + blockIdx.y);
        } else if constexpr (CARRY) {
            const int nb = gridDim.x - ge.n;
            group_item<GROUP_BIG_BLOCK_PATH, GROUP_BIG_REG>(v, ge.items[blockIdx.x - nb], ge.slots, ge.mass, ge.ref,
                                                            w0 + blockIdx.y);
        }
        return;
    }
    int p = 0;
    while (p + 1 < nItems && (int)blockIdx.x >= items[p + 1].rowBlk0) p++;
    const PathItem& it = items[p];
    const int tile = blockIdx.x - it.rowBlk0;
    const int w = w0 + blockIdx.y;
    const int M = it.M, N = it.N;
    const double lam = it.lambda;
    double* msd = v.pathMsd + (size_t)w * v.msdStride + it.msdOff;
    const double* part = v.pathPart + (size_t)w * v.partStride + it.partOff;
    const double* r2 = v.pathR2 + it.msdOff;
    const int lane = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const int j = tile * PATH_TILE + lane;
    // this lane's position and row indices, loaded before the weight table
    // so that their round trips overlap the MSD loads (clamped: no branch)
    const int jc = min(j, M - 1);
    const float4 q = gather_pos(v, slots, v.pStride, w, it.begin + jc);
    const size_t rowS = row_index(v, w, it.row0 + 2 * jc), rowZ = row_index(v, w, it.row0 + 2 * jc + 1);
    // the slice partials: all loads issued together (indices clamped into
    // the slice count), then added in slice order
    const int nsl = it.slices;
    double x2 = 0.0;
    {
        double pv[PATH_MAX_SLICES];
#pragma unroll
        for (int sl = 0; sl < PATH_MAX_SLICES; sl++) pv[sl] = part[nsl * N + min(sl, nsl - 1)];
#pragma unroll
        for (int sl = 0; sl < PATH_MAX_SLICES; sl++)
            if (sl < nsl) x2 += pv[sl];
    }
    double t[2] = {0.0, 0.0};
    for (int f = threadIdx.x; f < N; f += blockDim.x) {
        double pv[PATH_MAX_SLICES];
#pragma unroll
        for (int sl = 0; sl < PATH_MAX_SLICES; sl++) pv[sl] = part[min(sl, nsl - 1) * N + f];
        const double r2f = r2[f];
        double cr = 0.0;
#pragma unroll
        for (int sl = 0; sl < PATH_MAX_SLICES; sl++)
            if (sl < nsl) cr += pv[sl];
        const double m = (x2 - 2.0 * cr + r2f) / M;
        if (tile == 0) msd[f] = m;  // read back below by this thread, and by the frame adjoint
        const double wv = exp(-lam * m);
        shW[f] = wv;
        t[0] += wv;
        t[1] += (f + 1) * wv;
    }
    block_sum<2, PH>(t, sh);  // ends with a barrier: shW is complete
    const double S = t[0];
    const bool ok = S > 1e-300;
    const double s = ok ? t[1] / S : 0.0;
    const double z = ok ? -log(S) / lam : 0.0;
    if (tile == 0) {
        double a[2] = {0.0, 0.0};
        for (int f = threadIdx.x; f < N; f += blockDim.x) {
            a[0] += ((f + 1) - s) * msd[f] * shW[f];
            a[1] += msd[f] * shW[f];
        }
        block_sum<2, PH>(a, sh);
        if (threadIdx.x == 0) {
            v.cvVal[(size_t)w * v.V + it.valueS] = s;
            v.cvVal[(size_t)w * v.V + it.valueS + 1] = z;
            double* aux = v.pathAux + ((size_t)w * v.nPath + it.aux) * 2;
            aux[0] = ok ? -a[0] / S : 0.0;
            aux[1] = ok ? (a[1] / S - z) / lam : 0.0;
        }
    }
    double js0 = 0, js1 = 0, js2 = 0, jz0 = 0, jz1 = 0, jz2 = 0;
    if (j < M) {
        // frames f = ph, ph + PH, ... in increasing order; the loads
        // of PATH_ROWS_UNROLL of them are issued before any is used
        const float* F = frames + it.frameOff + (size_t)j * 3;
        for (int fb = ph; fb < N; fb += PH * PATH_ROWS_UNROLL) {
            float r[PATH_ROWS_UNROLL][3];
#pragma unroll
            for (int u = 0; u < PATH_ROWS_UNROLL; u++) {  // clamped frame: unconditional loads
                const float* rp = F + (size_t)min(fb + u * PH, N - 1) * M * 3;
                r[u][0] = rp[0];
                r[u][1] = rp[1];
                r[u][2] = rp[2];
            }
#pragma unroll
            for (int u = 0; u < PATH_ROWS_UNROLL; u++) {
                const int f = fb + u * PH;
                if (f < N) {
                    const double wv = shW[f];
                    const double dx = (double)q.x - r[u][0], dy = (double)q.y - r[u][1], dz = (double)q.z - r[u][2];
                    const double a = ((f + 1) - s) * wv;
                    js0 += a * dx;
                    js1 += a * dy;
                    js2 += a * dz;
                    jz0 += wv * dx;
                    jz1 += wv * dy;
                    jz2 += wv * dz;
                }
            }
        }
    }
    shRed[0][threadIdx.x] = js0;
    shRed[1][threadIdx.x] = js1;
    shRed[2][threadIdx.x] = js2;
    shRed[3][threadIdx.x] = jz0;
    shRed[4][threadIdx.x] = jz1;
    shRed[5][threadIdx.x] = jz2;
    __syncthreads();
    if (ph == 0 && j < M) {
        double r6[6];
#pragma unroll
        for (int k = 0; k < 6; k++) {
            double acc = 0.0;
            for (int q = 0; q < PH; q++) acc += shRed[k][q * 64 + lane];
            r6[k] = acc;
        }
        const double scS = ok ? -2.0 * lam / (M * S) : 0.0, scZ = ok ? 2.0 / (M * S) : 0.0;
        store_row_at(v, rowS, scS * r6[0], scS * r6[1], scS * r6[2]);
        store_row_at(v, rowZ, scZ * r6[3], scZ * r6[4], scZ * r6[5]);
    }
}

// ---------------------------------------------------------------------------
// Adjoint of the path CV with respect to its reference frames (an extension:
// the reference has no frame derivatives).  With V(s, z) and the MSD metric,
//   dV/dr_fj = (2/M) (w_f/S) (x_j - r_fj) [lambda Vs (f+1-s) - Vz],
// Vs, Vz = dU/ds, dU/dz of the last evaluation.  Over the walkers, with
// upstream weights g_b,
//   dL/dr_fj = sum_b c_bf (x_bj - r_fj),
//   c_bf = g_b (2/M) (w_bf/S_b) [lambda Vs_b (f+1-s_b) - Vz_b].
// Stage 1 forms c (one workgroup per walker, w_f, S and s exactly as the
// forward's row blocks form them); stage 2 runs one thread per (frame, atom)
// and adds the walkers in index order (deterministic), FP64.  Where every
// weight underflowed (S = 0) the walker contributes nothing, as its rows.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64 * PATH_PHASES) void k_path_frames_coef(StepView v, const PathItem* __restrict__ items,
                                                                       int pi, const double* __restrict__ gV,
                                                                       double* __restrict__ coef)
{
    __shared__ double sh[4 * 16];
    const PathItem& it = items[pi];
    const int w = blockIdx.x;
    const int N = it.N;
    const double lam = it.lambda;
    const double* msd = v.pathMsd + (size_t)w * v.msdStride + it.msdOff;
    double t[2] = {0.0, 0.0};
    for (int f = threadIdx.x; f < N; f += blockDim.x) {
        const double wv = exp(-lam * msd[f]);
        t[0] += wv;
        t[1] += (f + 1) * wv;
    }
    block_sum<2, PATH_PHASES>(t, sh);
    const double S = t[0];
    const bool ok = S > 1e-300;
    const double s = ok ? t[1] / S : 0.0;
    const double Vs = v.dUdCV[(size_t)w * v.V + it.valueS], Vz = v.dUdCV[(size_t)w * v.V + it.valueS + 1];
    const double pre = ok ? gV[w] * 2.0 / ((double)it.M * S) : 0.0;
    for (int f = threadIdx.x; f < N; f += blockDim.x) {
        const double wv = exp(-lam * msd[f]);
        coef[(size_t)w * N + f] = pre * wv * (lam * Vs * ((f + 1) - s) - Vz);
    }
}

__global__ __launch_bounds__(256) void k_path_frames_grad(StepView v, const PathItem* __restrict__ items, int pi,
                                                          const int* __restrict__ slots,
                                                          const float* __restrict__ frames,
                                                          const double* __restrict__ coef, double* __restrict__ gF,
                                                          int W)
{
    const PathItem& it = items[pi];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= it.N * it.M) return;
    const int f = t / it.M, j = t - f * it.M;
    const float* r = frames + it.frameOff + (size_t)t * 3;  // frame f, atom j
    const double r0 = r[0], r1 = r[1], r2 = r[2];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int b = 0; b < W; b++) {
        const float4 q = gather_pos(v, slots, v.pStride, b, it.begin + j);
        const double c = coef[(size_t)b * it.N + f];
        a0 += c * ((double)q.x - r0);
        a1 += c * ((double)q.y - r1);
        a2 += c * ((double)q.z - r2);
    }
    gF[(size_t)t * 3] = a0;
    gF[(size_t)t * 3 + 1] = a1;
    gF[(size_t)t * 3 + 2] = a2;
}

void launch_path_frames_adjoint(const StepView& v, const PathItem* items, int pi, int N, int M, const int* slots,
                                const float* frames, const double* gV, double* coef, double* gF, int W, hipStream_t s)
{
    hipLaunchKernelGGL(k_path_frames_coef, dim3(W), dim3(64 * PATH_PHASES), 0, s, v, items, pi, gV, coef);
    const int n = N * M;
    hipLaunchKernelGGL(k_path_frames_grad, dim3((n + 255) / 256), dim3(256), 0, s, v, items, pi, slots, frames, coef,
                       gF, W);
}

// Row blocks of 8 frame phases, or of 4 when there are more blocks than fit
// on the chip at once (each lane then covers twice the frames, and twice as
// many blocks are resident).  Up to 256 frames both forms form w_f, S and s
// with the same additions (a thread per frame, the waves past the frames add
// zeros), so the adjoint's coefficients match either.
// phases: 4 or 8 waves per row block (the autotuner's choice; 0 = 4 when the
// blocks exceed one residency round and the paths have at most 256 frames)
// Returns true when the launch carried the wide group CVs of `extra` (their
// workgroups are 512 threads: with 8 phases).  The contact-map sums of
// `extra` (extra->nCm > 0) run with any phase count.
bool launch_path_rows(const StepView& v, const PathItem* items, int n, int nRowBlocks, int maxN, const int* slots,
                      const float* frames, int w0, int nw, int phases, hipStream_t s, const GroupExtra* extra)
{
    if (n <= 0) return false;
    if (phases == 0) phases = (long long)nRowBlocks * nw > 512 && maxN <= 256 ? 4 : PATH_PHASES;
    static_assert(64 * PATH_PHASES == GROUP_BIG_BLOCK_PATH, "8 phases: the carrying launch");
    GroupExtra cmOnly{};
    cmOnly.cmBlock = -1;
    if (extra) {
        cmOnly.cm = extra->cm;
        cmOnly.nCm = extra->nCm;
        cmOnly.nCmBlocks = extra->nCmBlocks;
        cmOnly.cmPart = extra->cmPart;
    }
    const int cmb = cmOnly.nCm > 0 ? 1 : 0;
    if (cmb) cmOnly.cmBlock = nRowBlocks;
    if (phases == 4) {
        hipLaunchKernelGGL((k_path_rows<4, false>), dim3(nRowBlocks + cmb, nw), dim3(64 * 4), 0, s, v, items, n,
                           slots, frames, w0, cmOnly);
        return false;
    }
    if (extra && extra->n > 0) {
        GroupExtra ge = *extra;
        ge.cmBlock = cmOnly.cmBlock;
        hipLaunchKernelGGL((k_path_rows<PATH_PHASES, true>), dim3(nRowBlocks + cmb + extra->n, nw),
                           dim3(64 * PATH_PHASES), 0, s, v, items, n, slots, frames, w0, ge);
        return true;
    }
    hipLaunchKernelGGL((k_path_rows<PATH_PHASES, false>), dim3(nRowBlocks + cmb, nw), dim3(64 * PATH_PHASES), 0, s,
                       v, items, n, slots, frames, w0, cmOnly);
    return false;
}

void launch_cv_path(const StepView& v, const PathItem* items, int n, int nMsdBlocks, int nRowBlocks, int maxN,
                    const int* slots, const float* frames, int w0, int nw, int phases, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_path_msd, dim3(nMsdBlocks, (nw + PATH_TW - 1) / PATH_TW), dim3(PATH_MSD_THREADS), 0, s, v,
                       items, n, slots, frames, w0, nw);
    launch_path_rows(v, items, n, nRowBlocks, maxN, slots, frames, w0, nw, phases, s, nullptr);
}

// ---------------------------------------------------------------------------
// Contact maps of more than CM_CHUNK pairs, in chunks of CM_CHUNK pairs (one
// 256-thread workgroup each, every gather of a thread issued before its
// arithmetic), so a long map spreads over many CUs.  Each chunk writes its
// rows and its partial sum; the last chunk of the map to finish adds the
// partials in chunk order and writes the value, so the value does not depend
// on which chunk finished last.  In the batched step of a plan with a path CV
// the chunks stop at their partials and the path-row launch that follows adds
// them (cm_finish): the ticket's agent-scope fence, which writes back the
// L2 with every row stored so far, costs a chunk about 4 us.
// ---------------------------------------------------------------------------
constexpr int CM_THREADS = 256;
constexpr int CM_PER_THREAD = CM_CHUNK / CM_THREADS;
static_assert(CM_CHUNK % CM_THREADS == 0, "whole pairs per thread");

__device__ void cm_chunk_block(const StepView& v, const CVFusedPlan& p, int lb, int w)
{
    __shared__ double sh[16];
    int q = 0;
    while (q + 1 < p.nCm && lb >= p.cm[q + 1].chunk0) q++;
    const GroupCVItem& it = p.cm[q];
    const int c = lb - it.chunk0;
    const int np = (it.end - it.begin) / 2;
    const int nChunks = (np + CM_CHUNK - 1) / CM_CHUNK;
    const int p0 = c * CM_CHUNK, n = min(CM_CHUNK, np - p0);
    const int tid = threadIdx.x;
    const WalkerBox wb = load_box(v, w);
    const float* prm = p.gref + it.refOff;
    float4 pa[CM_PER_THREAD], pb[CM_PER_THREAD], pp[CM_PER_THREAD];
    size_t ra[CM_PER_THREAD], rb[CM_PER_THREAD];
#pragma unroll
    for (int r = 0; r < CM_PER_THREAD; r++) {  // clamped index: no divergent loads
        const int k = p0 + min(tid + r * CM_THREADS, n - 1);
        pa[r] = gather_pos(v, p.gslots, v.gStride, w, it.begin + 2 * k);
        pb[r] = gather_pos(v, p.gslots, v.gStride, w, it.begin + 2 * k + 1);
        pp[r] = make_float4(prm[4 * k], prm[4 * k + 1], prm[4 * k + 2], prm[4 * k + 3]);
        ra[r] = row_index(v, w, it.row0 + 2 * k);
        rb[r] = row_index(v, w, it.row0 + 2 * k + 1);
    }
    double s[1] = {0.0};
#pragma unroll
    for (int r = 0; r < CM_PER_THREAD; r++) {
        if (tid + r * CM_THREADS < n) {
            double dx = (double)pb[r].x - pa[r].x, dy = (double)pb[r].y - pa[r].y, dz = (double)pb[r].z - pa[r].z;
            if (v.pbc) pbc_exact(dx, dy, dz, wb.b);
            double coeff;
            s[0] += contact_pair(dx, dy, dz, pp[r], coeff);
            store_row_at(v, ra[r], -coeff * dx, -coeff * dy, -coeff * dz);
            store_row_at(v, rb[r], coeff * dx, coeff * dy, coeff * dz);
        }
    }
    block_sum<1, CM_THREADS / 64>(s, sh);
    double* part = p.cmPart + (size_t)w * p.nCmBlocks + it.chunk0;
    unsigned int* cnt = p.cmCount + (size_t)w * p.nCm + it.chunkItem;
    if (tid == 0) part[c] = s[0];
    if (p.cmDefer) return;  // the path-row launch adds the partials
    if (!last_block(cnt, (unsigned int)nChunks)) return;
    if (tid >= 64) return;
    const double tot = wave_sum_partials(part, nChunks);
    if (tid == 0) {
        v.cvVal[(size_t)w * v.V + it.value] = tot;
        *cnt = 0u;
    }
}

__global__ __launch_bounds__(CM_THREADS) void k_cv_cm_chunks(StepView v, CVFusedPlan p, int w0)
{
    cm_chunk_block(v, p, blockIdx.x, w0 + blockIdx.y);
}

void launch_cv_cm_chunks(const StepView& v, const CVFusedPlan& p, int w0, int nw, hipStream_t s)
{
    if (p.nCmBlocks <= 0) return;
    hipLaunchKernelGGL(k_cv_cm_chunks, dim3(p.nCmBlocks, nw), dim3(CM_THREADS), 0, s, v, p, w0);
}

// ---------------------------------------------------------------------------
// Batched CV stage: every group CV, path CV, contact-map chunk and few-atom
// CV of every walker in one launch (blocks [0, nGroup) group items, then path
// MSD blocks, then contact-map chunks, then blocks of 256 few-atom items).
// Same device code as the per-kind kernels.
// ---------------------------------------------------------------------------
static_assert(GROUP_BLOCK == PATH_MSD_THREADS, "the MSD blocks of the fused CV stage run on GROUP_BLOCK threads");
static_assert(GROUP_BLOCK == CM_THREADS, "the contact-map chunks of the fused CV stage run on GROUP_BLOCK threads");

__global__ __launch_bounds__(GROUP_BLOCK) void k_cv_fused(StepView v, CVFusedPlan p)
{
    const int w = blockIdx.y;
    const int b = blockIdx.x;
    if (b < p.nGroup) {
        group_item<GROUP_BLOCK, GROUP_REG>(v, p.group[b], p.gslots, p.gmass, p.gref, w);
    } else if (b < p.nGroup + p.nPathMsdBlocks) {
        // path q's (walker tile, unit) items spread over its msdBlkF blocks of
        // every walker row: item = local block * W + walker row
        const int lb = b - p.nGroup;
        int q = 0;
        while (q + 1 < p.nPath && lb >= p.path[q + 1].msdBlkF0) q++;
        const PathItem& it = p.path[q];
        const int W = gridDim.y;
        const int item = (lb - it.msdBlkF0) * W + w;
        const int tile = item / it.units;
        if (tile * PATH_TW < W)
            path_msd_block(v, it, p.pslots, p.frames, item % it.units, tile * PATH_TW, min(PATH_TW, W - tile * PATH_TW));
    } else if (b < p.nGroup + p.nPathMsdBlocks + p.nCmBlocks) {
        cm_chunk_block(v, p, b - p.nGroup - p.nPathMsdBlocks, w);
    } else {
        const int i = (b - p.nGroup - p.nPathMsdBlocks - p.nCmBlocks) * GROUP_BLOCK + threadIdx.x;
        if (i < p.nSmall) small_item(v, p.small[i], w);
    }
}

void launch_cv_fused(const StepView& v, const CVFusedPlan& p, int nw, hipStream_t s)
{
    const int nb = p.nGroup + p.nPathMsdBlocks + p.nCmBlocks + (p.nSmall + GROUP_BLOCK - 1) / GROUP_BLOCK;
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_cv_fused, dim3(nb, nw), dim3(GROUP_BLOCK), 0, s, v, p);
}

}  // namespace glued
