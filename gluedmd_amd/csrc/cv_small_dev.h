// Few-atom CVs (distance, angle, dihedral, position) and the geometry they
// share with the other CV kernels: the exact minimum image, the slot gather
// and the Jacobian-row store.  A header so that every kernel translation
// unit that evaluates few-atom CVs (the CV kernels and the one-workgroup
// step for small plans) has the same device code.
#pragma once
#include "glued_internal.h"
#include <hip/hip_runtime.h>
#include <cmath>

namespace glued {

// ---------------------------------------------------------------------------
// Minimum image (triclinic, Z then Y then X), CommonGluedKernels.cpp:74-91.
// bx = {a, b, c, 1/ax, 1/by, 1/cz}.  No FMA contraction: the oracle is built
// with -ffp-contract=off and must see the same rounding.
// ---------------------------------------------------------------------------
__device__ inline void pbc_exact(double& x, double& y, double& z, const double* __restrict__ bx)
{
#pragma clang fp contract(off)
    double s3 = floor(z * bx[11] + 0.5);
    x = x - s3 * bx[6];
    y = y - s3 * bx[7];
    z = z - s3 * bx[8];
    double s2 = floor(y * bx[10] + 0.5);
    x = x - s2 * bx[3];
    y = y - s2 * bx[4];
    double s1 = floor(x * bx[9] + 0.5);
    x = x - s1 * bx[0];
}

// The walker's box (the 3x3 matrix and the reciprocal diagonal) read once
// into scalar registers.  Every caller's walker index is uniform over the
// workgroup.  Minimum-image code that reads the box from memory inside a loop
// would issue those loads after earlier row stores, and on this target such a
// load waits for the stores to complete.
struct WalkerBox {
    double b[12];
};

__device__ inline WalkerBox load_box(const StepView& v, int w)
{
    WalkerBox r;
    const double* bx = v.box + (size_t)w * 12;
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const double x = bx[k];
        r.b[k] = __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(x)),
                                  __builtin_amdgcn_readfirstlane(__double2loint(x)));
    }
    return r;
}

__device__ inline double dot3_exact(double x, double y, double z)
{
#pragma clang fp contract(off)
    return x * x + y * y + z * z;
}

__device__ inline float4 load_pos(const StepView& v, int w, int slot)
{
    const WalkerInfo wi = v.walkers[w];
    int g = v.gidx[(size_t)w * v.S + slot];
    return v.posq[wi.posOff + g];
}

// Index of Jacobian row `row` of walker w in the row arrays.  CV code reads
// these with its operands, before any row is stored: the vector memory
// counter also counts stores, so a table load issued after a row store first
// waits for that store to complete.
__device__ inline size_t row_index(const StepView& v, int w, int row)
{
    return (size_t)w * v.R + v.rowPos[row];
}

__device__ inline void store_row_at(const StepView& v, size_t i, double gx, double gy, double gz)
{
    v.jrow[i] = make_float4((float)gx, (float)gy, (float)gz, 0.f);
}

__device__ inline void store_row(const StepView& v, int w, int row, double gx, double gy, double gz)
{
    store_row_at(v, row_index(v, w, row), gx, gy, gz);
}

// ---------------------------------------------------------------------------
// Few-atom CVs, one thread per (item, walker), FP64.
// distance :177-229, angle :238-325, dihedral :335-444, position :1084-1114
// ---------------------------------------------------------------------------
__device__ inline void small_item(const StepView& v, const SmallCVItem& it, int w)
{
    if (it.type == CV_ENERGY) {  // value supplied by the caller; Jacobian via the dense scatter
        v.cvVal[(size_t)w * v.V + it.value] = v.energyU[w];
        return;
    }
    double p[4][3];
    const int na = it.type == CV_DISTANCE ? 2 : it.type == CV_ANGLE ? 3 : it.type == CV_DIHEDRAL ? 4 : 1;
    // every gather issued before the first use (independent loads in flight together)
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const float4 q = load_pos(v, w, it.slot[a < na ? a : 0]);
        p[a][0] = q.x;
        p[a][1] = q.y;
        p[a][2] = q.z;
    }
    size_t ri[4];
#pragma unroll
    for (int a = 0; a < 4; a++) ri[a] = row_index(v, w, it.row0 + (a < na ? a : 0));
    const WalkerBox wb = load_box(v, w);
    const double* bx = wb.b;
    double val = 0.0;
    double* cvv = v.cvVal + (size_t)w * v.V;
    if (it.type == CV_DISTANCE) {
        double dx = p[1][0] - p[0][0], dy = p[1][1] - p[0][1], dz = p[1][2] - p[0][2];
        if (v.pbc) pbc_exact(dx, dy, dz, bx);
        double d = sqrt(dx * dx + dy * dy + dz * dz);
        double inv = d > 1e-10 ? 1.0 / d : 0.0;
        val = d;
        store_row_at(v, ri[0], -dx * inv, -dy * inv, -dz * inv);
        store_row_at(v, ri[1], dx * inv, dy * inv, dz * inv);
    } else if (it.type == CV_ANGLE) {
        double r1[3], r2[3];
        for (int k = 0; k < 3; k++) {
            r1[k] = p[0][k] - p[1][k];
            r2[k] = p[2][k] - p[1][k];
        }
        if (v.pbc) {
            pbc_exact(r1[0], r1[1], r1[2], bx);
            pbc_exact(r2[0], r2[1], r2[2], bx);
        }
        double l1 = sqrt(r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2]);
        double l2 = sqrt(r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]);
        double i1 = l1 > 1e-10 ? 1.0 / l1 : 0.0, i2 = l2 > 1e-10 ? 1.0 / l2 : 0.0;
        double n1[3], n2[3];
        for (int k = 0; k < 3; k++) {
            n1[k] = r1[k] * i1;
            n2[k] = r2[k] * i2;
        }
        double c = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
        c = fmax(-1.0, fmin(1.0, c));
        val = acos(c);
        double s = sqrt(1.0 - c * c);
        double is = s > 1e-6 ? 1.0 / s : 0.0;
        double ga[3], gc[3];
        for (int k = 0; k < 3; k++) {
            ga[k] = -(n2[k] - c * n1[k]) * i1 * is;
            gc[k] = -(n1[k] - c * n2[k]) * i2 * is;
        }
        store_row_at(v, ri[0], ga[0], ga[1], ga[2]);
        store_row_at(v, ri[1], -(ga[0] + gc[0]), -(ga[1] + gc[1]), -(ga[2] + gc[2]));
        store_row_at(v, ri[2], gc[0], gc[1], gc[2]);
    } else if (it.type == CV_DIHEDRAL) {
        double b1[3], b2[3], b3[3];
        for (int k = 0; k < 3; k++) {
            b1[k] = p[1][k] - p[0][k];
            b2[k] = p[2][k] - p[1][k];
            b3[k] = p[3][k] - p[2][k];
        }
        if (v.pbc) {
            pbc_exact(b1[0], b1[1], b1[2], bx);
            pbc_exact(b2[0], b2[1], b2[2], bx);
            pbc_exact(b3[0], b3[1], b3[2], bx);
        }
        double t[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2], b1[0] * b2[1] - b1[1] * b2[0]};
        double u[3] = {b2[1] * b3[2] - b2[2] * b3[1], b2[2] * b3[0] - b2[0] * b3[2], b2[0] * b3[1] - b2[1] * b3[0]};
        double d = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
        double t2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
        double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        val = atan2(d * (b1[0] * u[0] + b1[1] * u[1] + b1[2] * u[2]), t[0] * u[0] + t[1] * u[1] + t[2] * u[2]);
        double iT = t2 > 1e-10 ? 1.0 / t2 : 0.0, iU = u2 > 1e-10 ? 1.0 / u2 : 0.0, iD = d > 1e-10 ? 1.0 / d : 0.0;
        double c1 = (b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2]) * iD;
        double c2 = (b3[0] * b2[0] + b3[1] * b2[1] + b3[2] * b2[2]) * iD;
        double aT = d * iT, aU = d * iU, bT = (d + c1) * iT, bU = c2 * iU;
        double ja[3], jb[3], jd[3];
        for (int k = 0; k < 3; k++) {
            ja[k] = -aT * t[k];
            jd[k] = aU * u[k];
            jb[k] = bT * t[k] + bU * u[k];
        }
        store_row_at(v, ri[0], ja[0], ja[1], ja[2]);
        store_row_at(v, ri[1], jb[0], jb[1], jb[2]);
        store_row_at(v, ri[2], -(ja[0] + jb[0] + jd[0]), -(ja[1] + jb[1] + jd[1]), -(ja[2] + jb[2] + jd[2]));
        store_row_at(v, ri[3], jd[0], jd[1], jd[2]);
    } else if (it.type == CV_POSITION) {
        val = p[0][it.comp];
        store_row_at(v, ri[0], it.comp == 0, it.comp == 1, it.comp == 2);
    }
    cvv[it.value] = val;
}

}  // namespace glued
