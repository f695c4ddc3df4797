/*
 * glued_oracle.c -- FP64 CPU restatement of GLUED's enhanced-sampling step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
 * implementation in gluedmd_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or as the
 * timed CPU baseline).  The product path never calls it.
 *
 * Every function restates the formula of the reference GPU kernel it cites
 * (platforms/common/src/CommonGluedKernels.cpp in MarvinTaterra/GluedMD), in
 * FP64, from FP32 inputs widened to double.  Where the reference stores a
 * parameter as float (coordination r0/n/m, path lambda and frames, RMSD
 * reference, DRMSD reference distances, contact-map parameters) the caller
 * passes the float-rounded value, exactly as the reference kernel sees it.
 *
 * Functions the reference does not have (coordination cutoff, Kabsch RMSD,
 * hill-list metadynamics, path lambda derivatives) are defined from the
 * mathematics; their pins are listed in the test files and in the header
 * comment of each function.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fPIC -shared).
 * -ffp-contract=off keeps a*b+c as two rounded operations so the FP64
 * predicates (coordination cutoff, hill truncation, OPES merge) match the GPU's
 * explicitly rounded __dmul_rn/__dadd_rn sequence bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Geometry: triclinic minimum image, Z then Y then X.                       */
/* CommonGluedKernels.cpp:74-91 (deltaPBC).                                  */
/* box = {ax,ay,az, bx,by,bz, cx,cy,cz} (OpenMM reduced form).               */
/* ------------------------------------------------------------------------ */
static void delta_pbc(double d[3], const double* box)
{
    const double ix = 1.0 / box[0], iy = 1.0 / box[4], iz = 1.0 / box[8];
    double s3 = floor(d[2] * iz + 0.5);
    d[0] -= s3 * box[6];
    d[1] -= s3 * box[7];
    d[2] -= s3 * box[8];
    double s2 = floor(d[1] * iy + 0.5);
    d[0] -= s2 * box[3];
    d[1] -= s2 * box[4];
    double s1 = floor(d[0] * ix + 0.5);
    d[0] -= s1 * box[0];
}

static void vsub(double o[3], const double* b, const double* a)
{
    o[0] = b[0] - a[0];
    o[1] = b[1] - a[1];
    o[2] = b[2] - a[2];
}

static double vdot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void vcross(double o[3], const double* a, const double* b)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* Exported for the tests of the minimum-image rule itself. */
EXPORT void orc_delta_pbc(double* d, const double* box) { delta_pbc(d, box); }

/* ------------------------------------------------------------------------ */
/* Distance CV.  CommonGluedKernels.cpp:177-229.                              */
/* g[0..2] = d/dA, g[3..5] = d/dB.                                           */
/* ------------------------------------------------------------------------ */
EXPORT double orc_cv_distance(const double* pos, int a, int b, const double* box, int pbc, double* g)
{
    double dr[3];
    vsub(dr, pos + 3 * b, pos + 3 * a);
    if (pbc) delta_pbc(dr, box);
    double d = sqrt(vdot(dr, dr));
    double inv = d > 1e-10 ? 1.0 / d : 0.0;
    for (int k = 0; k < 3; k++) {
        g[k] = -dr[k] * inv;
        g[3 + k] = dr[k] * inv;
    }
    return d;
}

/* ------------------------------------------------------------------------ */
/* Angle CV (vertex at the middle atom).  CommonGluedKernels.cpp:238-325.     */
/* ------------------------------------------------------------------------ */
EXPORT double orc_cv_angle(const double* pos, const int* at, const double* box, int pbc, double* g)
{
    double r1[3], r2[3];
    vsub(r1, pos + 3 * at[0], pos + 3 * at[1]); /* A - B */
    vsub(r2, pos + 3 * at[2], pos + 3 * at[1]); /* C - B */
    if (pbc) {
        delta_pbc(r1, box);
        delta_pbc(r2, box);
    }
    double l1 = sqrt(vdot(r1, r1)), l2 = sqrt(vdot(r2, r2));
    double i1 = l1 > 1e-10 ? 1.0 / l1 : 0.0, i2 = l2 > 1e-10 ? 1.0 / l2 : 0.0;
    double n1[3], n2[3];
    for (int k = 0; k < 3; k++) {
        n1[k] = r1[k] * i1;
        n2[k] = r2[k] * i2;
    }
    double c = vdot(n1, n2);
    if (c > 1.0) c = 1.0;
    if (c < -1.0) c = -1.0;
    double th = acos(c);
    double s = sqrt(1.0 - c * c);
    double is = s > 1e-6 ? 1.0 / s : 0.0;
    for (int k = 0; k < 3; k++) {
        g[k] = -(n2[k] - c * n1[k]) * i1 * is;
        g[6 + k] = -(n1[k] - c * n2[k]) * i2 * is;
        g[3 + k] = -(g[k] + g[6 + k]);
    }
    return th;
}

/* ------------------------------------------------------------------------ */
/* Dihedral CV, Blondel-Karplus.  CommonGluedKernels.cpp:335-444.            */
/* g = 4 rows of 3 (atoms A,B,C,D).                                          */
/* ------------------------------------------------------------------------ */
EXPORT double orc_cv_dihedral(const double* pos, const int* at, const double* box, int pbc, double* g)
{
    double b1[3], b2[3], b3[3], t[3], u[3];
    vsub(b1, pos + 3 * at[1], pos + 3 * at[0]);
    vsub(b2, pos + 3 * at[2], pos + 3 * at[1]);
    vsub(b3, pos + 3 * at[3], pos + 3 * at[2]);
    if (pbc) {
        delta_pbc(b1, box);
        delta_pbc(b2, box);
        delta_pbc(b3, box);
    }
    vcross(t, b1, b2);
    vcross(u, b2, b3);
    double d = sqrt(vdot(b2, b2));
    double t2 = vdot(t, t), u2 = vdot(u, u);
    double phi = atan2(d * vdot(b1, u), vdot(t, u));
    double invT2 = t2 > 1e-10 ? 1.0 / t2 : 0.0;
    double invU2 = u2 > 1e-10 ? 1.0 / u2 : 0.0;
    double invD = d > 1e-10 ? 1.0 / d : 0.0;
    double c1 = vdot(b1, b2) * invD, c2 = vdot(b3, b2) * invD;
    double aT = d * invT2, aU = d * invU2, bTB = (d + c1) * invT2, bUB = c2 * invU2;
    for (int k = 0; k < 3; k++) {
        double ja = -aT * t[k];
        double jd = aU * u[k];
        double jb = bTB * t[k] + bUB * u[k];
        g[k] = ja;
        g[3 + k] = jb;
        g[6 + k] = -(ja + jb + jd);
        g[9 + k] = jd;
    }
    return phi;
}

/* ------------------------------------------------------------------------ */
/* COM distance.  CommonGluedKernels.cpp:456-539.  Group 1 = atoms[0..n1),   */
/* group 2 = atoms[n1..n).  masses per listed atom.  COM without PBC,        */
/* difference with PBC.                                                      */
/* ------------------------------------------------------------------------ */
EXPORT double orc_cv_com_distance(const double* pos, const int* atoms, int n1, int n, const double* m,
                                  const double* box, int pbc, double* g)
{
    double c1[3] = {0, 0, 0}, c2[3] = {0, 0, 0}, M1 = 0, M2 = 0;
    for (int j = 0; j < n; j++) {
        const double* p = pos + 3 * atoms[j];
        double* c = j < n1 ? c1 : c2;
        for (int k = 0; k < 3; k++) c[k] += m[j] * p[k];
        if (j < n1) M1 += m[j]; else M2 += m[j];
    }
    double dr[3];
    for (int k = 0; k < 3; k++) dr[k] = c2[k] / M2 - c1[k] / M1;
    if (pbc) delta_pbc(dr, box);
    double d = sqrt(vdot(dr, dr));
    double inv = d > 1e-10 ? 1.0 / d : 0.0;
    for (int j = 0; j < n; j++) {
        double w = j < n1 ? -m[j] / M1 : m[j] / M2;
        for (int k = 0; k < 3; k++) g[3 * j + k] = dr[k] * inv * w;
    }
    return d;
}

/* ------------------------------------------------------------------------ */
/* Radius of gyration.  CommonGluedKernels.cpp:556-639.                       */
/* COM without PBC; displacement from COM with PBC.                          */
/* ------------------------------------------------------------------------ */
EXPORT double orc_cv_gyration(const double* pos, const int* atoms, int n, const double* m, const double* box,
                              int pbc, double* g)
{
    double c[3] = {0, 0, 0}, M = 0;
    for (int j = 0; j < n; j++) {
        const double* p = pos + 3 * atoms[j];
        for (int k = 0; k < 3; k++) c[k] += m[j] * p[k];
        M += m[j];
    }
    for (int k = 0; k < 3; k++) c[k] /= M;
    double rg2 = 0.0;
    for (int j = 0; j < n; j++) {
        double dr[3];
        vsub(dr, pos + 3 * atoms[j], c);
        if (pbc) delta_pbc(dr, box);
        rg2 += m[j] * vdot(dr, dr);
        for (int k = 0; k < 3; k++) g[3 * j + k] = dr[k];
    }
    rg2 /= M;
    double rg = sqrt(rg2);
    double inv = rg > 1e-10 ? 1.0 / rg : 0.0;
    for (int j = 0; j < n; j++)
        for (int k = 0; k < 3; k++) g[3 * j + k] *= (m[j] / M) * inv;
    return rg;
}

/* ------------------------------------------------------------------------ */
/* Coordination.  CommonGluedKernels.cpp:899-1031.                           */
/* f(x) = (1-x^n)/(1-x^m), x = r/r0; Taylor expansion for |x-1| < 1e-4;      */
/* r^2 < 1e-20 and x > 1000 skipped.  dr = r_A - r_B (reference order).      */
/* Extension (no counterpart in the reference): cutoff > 0 keeps only pairs  */
/* with r^2 < cutoff^2, evaluated in FP64 from the FP32 inputs with no FMA    */
/* contraction, so the neighbour set is bit-exact with the GPU cell list.    */
/* *npairs = number of pairs that passed every predicate.                    */
/* ------------------------------------------------------------------------ */
static void rational_switch(double x, double r0, double rn, double rm, double* f, double* dfdr)
{
    double preRes = rn / rm;
    double preDfunc = 0.5 * rn * (rn - rm) / rm;
    double preSecDev = rn * (rm * rm - 3.0 * rm * (rn - 1.0) + rn * (2.0 * rn - 3.0)) / (6.0 * rm);
    double dx1 = x - 1.0;
    if (fabs(dx1) < 1e-4) {
        *f = preRes + dx1 * (preDfunc + 0.5 * dx1 * preSecDev);
        *dfdr = (preDfunc + dx1 * preSecDev) / r0;
    } else if (x < 1e-10) {
        *f = 1.0;
        *dfdr = 0.0;
    } else {
        double xn = pow(x, rn), xm = pow(x, rm);
        double den = 1.0 - xm;
        *f = (1.0 - xn) / den;
        double dfx = (rm * xm / x - rn * xn / x + (rn - rm) * xn * xm / x) / (den * den);
        *dfdr = dfx / r0;
    }
}

EXPORT double orc_switch(double x, double r0, double rn, double rm, double* dfdr)
{
    double f;
    rational_switch(x, r0, rn, rm, &f, dfdr);
    return f;
}

EXPORT double orc_cv_coordination(const double* pos, const int* A, int nA, const int* B, int nB, double r0,
                                  double rn, double rm, double cutoff, const double* box, int pbc, double* gA,
                                  double* gB, long long* npairs)
{
    double cn = 0.0;
    long long cnt = 0;
    double rc2 = cutoff > 0.0 ? cutoff * cutoff : 0.0;
    memset(gA, 0, sizeof(double) * 3 * (size_t)nA);
    memset(gB, 0, sizeof(double) * 3 * (size_t)nB);
    for (int k = 0; k < nA; k++) {
        const double* pk = pos + 3 * (size_t)A[k];
        for (int l = 0; l < nB; l++) {
            const double* pl = pos + 3 * (size_t)B[l];
            double dr[3];
            vsub(dr, pk, pl);
            if (pbc) delta_pbc(dr, box);
            double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            if (r2 < 1e-20) continue;
            if (cutoff > 0.0 && !(r2 < rc2)) continue;
            double r = sqrt(r2);
            double x = r / r0;
            if (x > 1000.0) continue;
            double f, dfdr;
            rational_switch(x, r0, rn, rm, &f, &dfdr);
            cn += f;
            cnt++;
            double s = dfdr / r;
            for (int c = 0; c < 3; c++) {
                gA[3 * k + c] += s * dr[c];
                gB[3 * l + c] -= s * dr[c];
            }
        }
    }
    if (npairs) *npairs = cnt;
    return cn;
}

/* ------------------------------------------------------------------------ */
/* Symmetric 4x4 eigen-solve by cyclic Jacobi rotations (FP64).              */
/* Used by the Kabsch RMSD below; a different method from the GPU's          */
/* characteristic-polynomial Newton solve, so a slip in either shows up.     */
/* ------------------------------------------------------------------------ */
static void jacobi4(double a[4][4], double evals[4], double evecs[4][4])
{
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) evecs[i][j] = (i == j);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int i = 0; i < 4; i++)
            for (int j = i + 1; j < 4; j++) off += a[i][j] * a[i][j];
        if (off < 1e-300) break;
        for (int p = 0; p < 4; p++)
            for (int q = p + 1; q < 4; q++) {
                if (fabs(a[p][q]) < 1e-300) continue;
                double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 4; k++) {
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 4; k++) {
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 4; k++) {
                    double vkp = evecs[k][p], vkq = evecs[k][q];
                    evecs[k][p] = c * vkp - s * vkq;
                    evecs[k][q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < 4; i++) evals[i] = a[i][i];
}

/* ------------------------------------------------------------------------ */
/* RMSD.  fit == 0: no-fit RMSD of CommonGluedKernels.cpp:810-881            */
/*   rmsd = sqrt(sum |r-ref|^2 / n), g = (r-ref)/(n rmsd).                   */
/* fit == 1 (no counterpart in the reference): RMSD after optimal            */
/*   translation + rotation (Kabsch), by Horn's quaternion: the largest       */
/*   eigenvalue L of the 4x4 key matrix built from S = sum y x^T (centred)   */
/*   gives msd = (Gx + Gy - 2L)/n; the gradient at the optimum is            */
/*   (x_c - R y_c)/(n rmsd) (the rotation and centring terms vanish there).  */
/* rot (may be NULL) receives R row-major, with R y_c ~ x_c.                 */
/* ------------------------------------------------------------------------ */
EXPORT double orc_cv_rmsd(const double* pos, const int* atoms, int n, const double* ref, int fit, double* g,
                          double* rot)
{
    if (!fit) {
        double s = 0.0;
        for (int j = 0; j < n; j++) {
            double d[3];
            vsub(d, pos + 3 * (size_t)atoms[j], ref + 3 * j);
            s += vdot(d, d);
            for (int k = 0; k < 3; k++) g[3 * j + k] = d[k];
        }
        double r = sqrt(s / n);
        double sc = (r > 1e-10 ? 1.0 / r : 0.0) / n;
        for (int j = 0; j < 3 * n; j++) g[j] *= sc;
        return r;
    }
    double xc[3] = {0, 0, 0}, yc[3] = {0, 0, 0};
    for (int j = 0; j < n; j++)
        for (int k = 0; k < 3; k++) {
            xc[k] += pos[3 * (size_t)atoms[j] + k];
            yc[k] += ref[3 * j + k];
        }
    for (int k = 0; k < 3; k++) {
        xc[k] /= n;
        yc[k] /= n;
    }
    double S[3][3] = {{0}}, Gx = 0, Gy = 0;
    for (int j = 0; j < n; j++) {
        double x[3], y[3];
        vsub(x, pos + 3 * (size_t)atoms[j], xc);
        vsub(y, ref + 3 * j, yc);
        Gx += vdot(x, x);
        Gy += vdot(y, y);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) S[a][b] += y[a] * x[b];
    }
    double K[4][4];
    double Sxx = S[0][0], Sxy = S[0][1], Sxz = S[0][2], Syx = S[1][0], Syy = S[1][1], Syz = S[1][2], Szx = S[2][0],
           Szy = S[2][1], Szz = S[2][2];
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
    double ev[4], V[4][4];
    jacobi4(K, ev, V);
    int im = 0;
    for (int i = 1; i < 4; i++)
        if (ev[i] > ev[im]) im = i;
    double q0 = V[0][im], qx = V[1][im], qy = V[2][im], qz = V[3][im];
    double qn = sqrt(q0 * q0 + qx * qx + qy * qy + qz * qz);
    q0 /= qn; qx /= qn; qy /= qn; qz /= qn;
    double R[3][3];
    R[0][0] = q0 * q0 + qx * qx - qy * qy - qz * qz;
    R[0][1] = 2 * (qx * qy - q0 * qz);
    R[0][2] = 2 * (qx * qz + q0 * qy);
    R[1][0] = 2 * (qy * qx + q0 * qz);
    R[1][1] = q0 * q0 - qx * qx + qy * qy - qz * qz;
    R[1][2] = 2 * (qy * qz - q0 * qx);
    R[2][0] = 2 * (qz * qx - q0 * qy);
    R[2][1] = 2 * (qz * qy + q0 * qx);
    R[2][2] = q0 * q0 - qx * qx - qy * qy + qz * qz;
    double msd = (Gx + Gy - 2.0 * ev[im]) / n;
    if (msd < 0) msd = 0;
    double r = sqrt(msd);
    double sc = (r > 1e-10 ? 1.0 / r : 0.0) / n;
    for (int j = 0; j < n; j++) {
        double x[3], y[3];
        vsub(x, pos + 3 * (size_t)atoms[j], xc);
        vsub(y, ref + 3 * j, yc);
        for (int a = 0; a < 3; a++) {
            double ry = R[a][0] * y[0] + R[a][1] * y[1] + R[a][2] * y[2];
            g[3 * j + a] = (x[a] - ry) * sc;
        }
    }
    if (rot)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) rot[3 * a + b] = R[a][b];
    return r;
}

/* ------------------------------------------------------------------------ */
/* Path CV (s, z), MSD metric, frames 1-indexed.                             */
/* CommonGluedKernels.cpp:667-794: w_f = exp(-lambda msd_f), S = sum w,      */
/* s = sum (f+1) w / S, z = -ln S / lambda,                                  */
/* ds/dx_j = -2 lambda/(M S) sum (f+1-s) w (x_j - r_fj),                     */
/* dz/dx_j =  2/(M S) sum w (x_j - r_fj).  No PBC (the reference ignores it). */
/* Extension for the adjoint: ds/dlambda = -(1/S) sum (f+1-s) msd_f w_f and  */
/* dz/dlambda = (<msd>_w - z)/lambda.                                        */
/* ------------------------------------------------------------------------ */
EXPORT void orc_cv_path(const double* pos, const int* atoms, int M, const double* frames, int N, double lambda,
                        double* sz, double* gs, double* gz, double* dlam)
{
    double* w = (double*)malloc(sizeof(double) * N);
    double* msd = (double*)malloc(sizeof(double) * N);
    double S = 0.0, sNum = 0.0;
    for (int f = 0; f < N; f++) {
        double acc = 0.0;
        for (int j = 0; j < M; j++) {
            double d[3];
            vsub(d, pos + 3 * (size_t)atoms[j], frames + 3 * ((size_t)f * M + j));
            acc += vdot(d, d);
        }
        msd[f] = acc / M;
        w[f] = exp(-lambda * msd[f]);
        S += w[f];
        sNum += (f + 1) * w[f];
    }
    double s = S > 1e-300 ? sNum / S : 0.0;
    double z = S > 1e-300 ? -log(S) / lambda : 0.0;
    sz[0] = s;
    sz[1] = z;
    /* every weight underflowed: s = z = 0 as in the reference; rows and the  */
    /* lambda derivatives are 0 (the reference divides by S = 0 there)        */
    const int ok = S > 1e-300;
    double scS = ok ? -2.0 * lambda / (M * S) : 0.0, scZ = ok ? 2.0 / (M * S) : 0.0;
    for (int j = 0; j < M; j++) {
        double js[3] = {0, 0, 0}, jz[3] = {0, 0, 0};
        for (int f = 0; f < N; f++) {
            double d[3];
            vsub(d, pos + 3 * (size_t)atoms[j], frames + 3 * ((size_t)f * M + j));
            for (int k = 0; k < 3; k++) {
                js[k] += ((f + 1) - s) * w[f] * d[k];
                jz[k] += w[f] * d[k];
            }
        }
        for (int k = 0; k < 3; k++) {
            gs[3 * j + k] = scS * js[k];
            gz[3 * j + k] = scZ * jz[k];
        }
    }
    if (dlam) {
        double a = 0.0, b = 0.0;
        for (int f = 0; f < N; f++) {
            a += ((f + 1) - s) * msd[f] * w[f];
            b += msd[f] * w[f];
        }
        dlam[0] = ok ? -a / S : 0.0;
        dlam[1] = ok ? (b / S - z) / lambda : 0.0;
    }
    free(w);
    free(msd);
}

/* ------------------------------------------------------------------------ */
/* Contact map.  CommonGluedKernels.cpp:1236-1323.  prm = 4 per pair         */
/* [r0, nn, mm, w]; integer powers; |den| < 1e-8 gives s = nn/mm with zero   */
/* derivative.  dr = r_B - r_A.                                              */
/* ------------------------------------------------------------------------ */
static double ipow(double x, int n)
{
    double r = 1.0;
    for (int i = 0; i < n; i++) r *= x;
    return r;
}

EXPORT double orc_cv_contactmap(const double* pos, const int* pairs, int np, const double* prm, const double* box,
                                int pbc, double* g)
{
    double cv = 0.0;
    for (int p = 0; p < np; p++) {
        double r0 = prm[4 * p], w = prm[4 * p + 3];
        int nn = (int)prm[4 * p + 1], mm = (int)prm[4 * p + 2];
        double dr[3];
        vsub(dr, pos + 3 * (size_t)pairs[2 * p + 1], pos + 3 * (size_t)pairs[2 * p]);
        if (pbc) delta_pbc(dr, box);
        double d = sqrt(vdot(dr, dr));
        double invd = d > 1e-10 ? 1.0 / d : 0.0;
        double x = d / r0;
        double xn = ipow(x, nn), xm = ipow(x, mm);
        double num = 1.0 - xn, den = 1.0 - xm, s, dsdd;
        if (fabs(den) < 1e-8) {
            s = (double)nn / (double)mm;
            dsdd = 0.0;
        } else {
            s = num / den;
            double dxn = nn > 0 ? nn * ipow(x, nn - 1) : 0.0;
            double dxm = mm > 0 ? mm * ipow(x, mm - 1) : 0.0;
            dsdd = ((-dxn * den + dxm * num) / (den * den)) / r0;
        }
        cv += w * s;
        double c = w * dsdd * invd;
        for (int k = 0; k < 3; k++) {
            g[6 * p + k] = -c * dr[k];
            g[6 * p + 3 + k] = c * dr[k];
        }
    }
    return cv;
}

/* ------------------------------------------------------------------------ */
/* Distance RMSD.  CommonGluedKernels.cpp:1130-1224.                         */
/* ------------------------------------------------------------------------ */
EXPORT double orc_cv_drmsd(const double* pos, const int* pairs, int np, const double* d0, const double* box, int pbc,
                           double* g)
{
    double ss = 0.0;
    for (int p = 0; p < np; p++) {
        double dr[3];
        vsub(dr, pos + 3 * (size_t)pairs[2 * p + 1], pos + 3 * (size_t)pairs[2 * p]);
        if (pbc) delta_pbc(dr, box);
        double d = sqrt(vdot(dr, dr));
        double dev = d - d0[p];
        ss += dev * dev;
        double invd = d > 1e-10 ? 1.0 / d : 0.0;
        for (int k = 0; k < 3; k++) {
            g[6 * p + k] = dev * invd * dr[k]; /* scaled below */
        }
    }
    double v = sqrt(ss / np);
    double inv = v > 1e-10 ? 1.0 / v : 0.0;
    for (int p = 0; p < np; p++)
        for (int k = 0; k < 3; k++) {
            double c = g[6 * p + k] * inv / np;
            g[6 * p + k] = -c;
            g[6 * p + 3 + k] = c;
        }
    return v;
}

/* ------------------------------------------------------------------------ */
/* Chain-rule scatter into the 2^32 fixed-point field-major force buffer.    */
/* CommonGluedKernels.cpp:2430-2456: per row F = -(dU/dCV * (double)g) in    */
/* double, truncated toward zero by the (mm_long) cast, added per component. */
/* ------------------------------------------------------------------------ */
EXPORT void orc_scatter_fixed(int nrows, const int* atom, const float* gx, const float* gy, const float* gz,
                              const int* cvIdx, const double* dUdCV, long long* force, int padded)
{
    const double scale = 4294967296.0;
    for (int j = 0; j < nrows; j++) {
        double u = dUdCV[cvIdx[j]];
        long long fx = (long long)(-(u * (double)gx[j]) * scale);
        long long fy = (long long)(-(u * (double)gy[j]) * scale);
        long long fz = (long long)(-(u * (double)gz[j]) * scale);
        force[atom[j]] += fx;
        force[atom[j] + padded] += fy;
        force[atom[j] + 2 * (size_t)padded] += fz;
    }
}

/* ------------------------------------------------------------------------ */
/* Grid metadynamics evaluation.  CommonGluedKernels.cpp:3574-3636.          */
/* Multilinear interpolation over 2^D corners; periodic wrap, non-periodic   */
/* clamp to [0, N-1] with lo <= N-2.  Returns V; dV[d] = dV/ds_d.            */
/* ------------------------------------------------------------------------ */
EXPORT double orc_metad_grid_eval(const double* cv, int D, const double* grid, const double* origin,
                                  const double* invSpacing, const int* npts, const int* strides, const int* periodic,
                                  double* dV)
{
    int lo[3], hi[3];
    double al[3];
    for (int d = 0; d < D; d++) {
        double fr = (cv[d] - origin[d]) * invSpacing[d];
        int N = npts[d];
        if (N <= 1) {
            lo[d] = hi[d] = 0;
            al[d] = 0.0;
        } else if (periodic[d]) {
            fr = fr - floor(fr / (double)N) * (double)N;
            if (fr < 0.0) fr += N;
            lo[d] = (int)fr;
            if (lo[d] >= N) lo[d] = N - 1;
            al[d] = fr - lo[d];
            hi[d] = (lo[d] + 1) % N;
        } else {
            if (fr < 0.0) fr = 0.0;
            if (fr > (double)(N - 1)) fr = (double)(N - 1);
            lo[d] = (int)fr;
            if (lo[d] >= N - 1) lo[d] = N - 2;
            al[d] = fr - lo[d];
            hi[d] = lo[d] + 1;
        }
    }
    double V = 0.0, g[3] = {0, 0, 0};
    for (int c = 0; c < (1 << D); c++) {
        int idx = 0;
        for (int d = 0; d < D; d++) idx += (((c >> d) & 1) ? hi[d] : lo[d]) * strides[d];
        double val = grid[idx], w = 1.0;
        for (int d = 0; d < D; d++) w *= ((c >> d) & 1) ? al[d] : (1.0 - al[d]);
        V += w * val;
        for (int d = 0; d < D; d++) {
            double wo = 1.0;
            for (int e = 0; e < D; e++)
                if (e != d) wo *= ((c >> e) & 1) ? al[e] : (1.0 - al[e]);
            g[d] += (((c >> d) & 1) ? 1.0 : -1.0) * wo * val;
        }
    }
    for (int d = 0; d < D; d++) dV[d] = g[d] * invSpacing[d];
    return V;
}

/* ------------------------------------------------------------------------ */
/* Grid deposit.  CommonGluedKernels.cpp:3645-3679: node position            */
/* origin + i/invSpacing, periodic min-image with period N/invSpacing via    */
/* round(), exponent = sum dc^2/(2 sigma^2), added when exponent <= 8.       */
/* ------------------------------------------------------------------------ */
EXPORT void orc_metad_grid_deposit(double* grid, int D, int total, const double* origin, const double* invSpacing,
                                   const int* npts, const int* strides, const int* periodic, const double* center,
                                   const double* sigma, double height)
{
    for (int idx = 0; idx < total; idx++) {
        double e = 0.0;
        for (int d = 0; d < D; d++) {
            int bi = (idx / strides[d]) % npts[d];
            double p = origin[d] + (double)bi / invSpacing[d];
            double dc = p - center[d];
            if (periodic[d]) {
                double per = (double)npts[d] / invSpacing[d];
                dc -= round(dc / per) * per;
            }
            e += dc * dc / (2.0 * sigma[d] * sigma[d]);
        }
        if (e <= 8.0) grid[idx] += height * exp(-e);
    }
}

/* ------------------------------------------------------------------------ */
/* Hill-list metadynamics (no counterpart in the reference, which is grid    */
/* only).  Defined as the grid deposit rule of CommonGluedKernels.cpp:       */
/* 3664-3677 evaluated at the CV point itself instead of at grid nodes:      */
/*   E_k = sum_d dc_d^2 * (1/(2 sigma_d^2)), dc periodic min-image when      */
/*   period_d > 0 (dc -= rint(dc * (1/P)) P, ties to even);                   */
/*   V = sum_k h_k e^-E_k [E_k<=8];                                          */
/*   dV/ds_d = sum_k h_k exp(-E_k) [E_k<=8] (-dc_d / sigma_d^2).             */
/* centers[k*D+d], heights[k].  Pinned by equality with grid MetaD at grid   */
/* nodes (tests) and brute force.                                            */
/* ------------------------------------------------------------------------ */
EXPORT double orc_hills_eval(const double* cv, int D, int nh, const double* centers, const double* heights,
                             const double* sigma, const double* period, double* dV)
{
    double V = 0.0, g[8] = {0}, i2s2[8], is2[8], ip[8];
    for (int d = 0; d < D; d++) {
        i2s2[d] = 1.0 / (2.0 * sigma[d] * sigma[d]);
        is2[d] = 1.0 / (sigma[d] * sigma[d]);
        ip[d] = period[d] > 0.0 ? 1.0 / period[d] : 0.0;
    }
    for (int k = 0; k < nh; k++) {
        double e = 0.0, dc[8];
        for (int d = 0; d < D; d++) {
            double x = cv[d] - centers[(size_t)k * D + d];
            if (period[d] > 0.0) x -= rint(x * ip[d]) * period[d];
            dc[d] = x;
            e += x * x * i2s2[d];
        }
        if (!(e <= 8.0)) continue;
        double t = heights[k] * exp(-e);
        V += t;
        for (int d = 0; d < D; d++) g[d] -= t * dc[d] * is2[d];
    }
    for (int d = 0; d < D; d++) dV[d] = g[d];
    return V;
}

/* Per-hill exponent, exported so the tests can see how far a hill sits from */
/* the E = 8 truncation edge (a hill whose predicate differs between an FP32  */
/* and an FP64 CV value is reported, not hidden).                            */
EXPORT double orc_hill_min_margin(const double* cv, int D, int nh, const double* centers, const double* sigma,
                                  const double* period)
{
    double m = 1e300, i2s2[8], ip[8];
    for (int d = 0; d < D; d++) {
        i2s2[d] = 1.0 / (2.0 * sigma[d] * sigma[d]);
        ip[d] = period[d] > 0.0 ? 1.0 / period[d] : 0.0;
    }
    for (int k = 0; k < nh; k++) {
        double e = 0.0;
        for (int d = 0; d < D; d++) {
            double x = cv[d] - centers[(size_t)k * D + d];
            if (period[d] > 0.0) x -= rint(x * ip[d]) * period[d];
            e += x * x * i2s2[d];
        }
        double a = fabs(e - 8.0);
        if (a < m) m = a;
    }
    return m;
}

/* ------------------------------------------------------------------------ */
/* OPES evaluation.  CommonGluedKernels.cpp:2642-2723.                       */
/* ------------------------------------------------------------------------ */
EXPORT double orc_opes_eval(const double* s, int D, int nk, const double* centers, const double* sigmas,
                            const double* logw, double sum_uprob, double invGF, double kT, double cutoff2, double eps,
                            double barrier, double* dV)
{
    for (int d = 0; d < D; d++) dV[d] = 0.0;
    if (nk == 0) return -barrier;
    double kfloor = exp(-0.5 * cutoff2), prob = 0.0, acc[16] = {0};
    for (int k = 0; k < nk; k++) {
        double n2 = 0.0, dc[16], sg[16];
        for (int d = 0; d < D; d++) {
            dc[d] = s[d] - centers[(size_t)k * D + d];
            sg[d] = sigmas[(size_t)k * D + d];
            n2 += dc[d] * dc[d] / (sg[d] * sg[d]);
        }
        if (!(n2 < cutoff2)) continue;
        double G = exp(-0.5 * n2), kern = G - kfloor;
        if (kern <= 0.0) continue;
        double h = exp(logw[k]);
        if (!(h > 0.0 && h < 1e308)) continue;
        prob += h * kern;
        for (int d = 0; d < D; d++) acc[d] += h * G * (-dc[d] / (sg[d] * sg[d]));
    }
    double pz = sum_uprob > 0.0 ? prob * (double)nk / sum_uprob : 0.0;
    if (!(pz >= 0.0)) pz = 0.0;
    double arg = pz + eps;
    double V = invGF * kT * log(arg > 1e-30 ? arg : 1e-30);
    if (prob > 0.0) {
        double f = invGF * kT * (pz / (pz + eps)) / prob;
        for (int d = 0; d < D; d++) dV[d] = f * acc[d];
    }
    return V;
}

/* Variant constants.  CommonBiasKernels.cpp:553-588.                        */
/* out = {invGF, cutoff2, eps, barrier}; sigma0 broadened by sqrt(gamma) for */
/* EXPLORE when not adaptive (applied by the caller).                        */
EXPORT void orc_opes_constants(int variant, double gamma, double kT, double* out)
{
    double invGF, cutoff2, eps, barrier;
    if (variant == 1) {
        invGF = 1.0;
        cutoff2 = 2.0 * gamma;
        eps = 0.0;
        barrier = gamma * kT;
    } else if (variant == 2) {
        invGF = gamma - 1.0;
        cutoff2 = 2.0 * gamma;
        eps = exp(-gamma / (gamma - 1.0));
        barrier = gamma * kT;
    } else {
        invGF = (gamma - 1.0) / gamma;
        cutoff2 = 2.0 * gamma / invGF;
        eps = exp(-1.0 / (invGF * (1.0 - invGF)));
        barrier = kT / (1.0 - invGF);
    }
    out[0] = invGF;
    out[1] = cutoff2;
    out[2] = eps;
    out[3] = barrier;
}

/* Adaptive-sigma Welford accumulator.  CommonGluedKernels.cpp:3009-3027.     */
EXPORT void orc_opes_welford(const double* cv, int D, double* mean, double* M2, int* nSamples)
{
    nSamples[0]++;
    int n = nSamples[0];
    for (int d = 0; d < D; d++) {
        double delta = cv[d] - mean[d];
        mean[d] += delta / n;
        M2[d] += delta * (cv[d] - mean[d]);
    }
}

/* ------------------------------------------------------------------------ */
/* OPES deposit.  CommonGluedKernels.cpp:2752-3003 (opesGatherDeposit):       */
/* weight accumulators and neff, Silverman / adaptive sigma, height          */
/* correction, strict-min Mahalanobis compression with the existing kernel's */
/* sigma (first index wins ties), GMM merge or append, sum_uprob updated     */
/* (merge variance f1 s1^2 + f2 s2^2 + f1 f2 (c1-c2)^2: equal to the         */
/* reference's f1(s1^2+c1^2)+f2(s2^2+c2^2)-cm^2, whose cancellation loses    */
/* ~11 digits when |c| ~ 3e4 and sigma ~ 0.05)                               */
/* O(N) on append and recomputed O(N^2) on merge -- the reference's own       */
/* algorithm, kept here as the oracle.                                       */
/* Returns the merged index, -1 for an append, -2 for no deposit, -3 when    */
/* the table is full.                                                        */
/* st_d = {sum_uprob, sumWeights, sum_w, sum_w2}; st_i = {numKernels,        */
/* numAllocated, stepCount, nSamples}.                                       */
/* ------------------------------------------------------------------------ */
EXPORT int orc_opes_deposit(const double* cvnew, int D, double V, double invKT, int variant, int adaptiveStride,
                            double gamma, double sigmaMin, const double* sigma0, double cutoff2, int maxKernels,
                            double* centers, double* sigmas, double* logw, double* st_d, int* st_i, double* mean,
                            double* M2)
{
    if (adaptiveStride > 0 && st_i[3] < adaptiveStride) return -2;
    int nBefore = st_i[0];
    if (variant != 1 && adaptiveStride == 0) st_i[3]++;
    int n = st_i[3];
    double c[16], sg[16];
    for (int d = 0; d < D; d++) c[d] = cvnew[d];
    double lw_raw, h_raw;
    if (variant == 2) {
        lw_raw = 0.0;
        h_raw = 1.0;
    } else {
        lw_raw = V * invKT;
        h_raw = exp(lw_raw);
    }
    st_d[2] += h_raw;
    st_d[3] += h_raw * h_raw;
    st_i[2]++;
    double sw = st_d[2], sw2 = st_d[3];
    double neff = (1.0 + sw) * (1.0 + sw) / (1.0 + sw2);
    double size = variant == 2 ? (double)st_i[2] : neff;
    for (int d = 0; d < D; d++) {
        double s;
        if (variant == 1) {
            s = sigma0[d];
        } else if (adaptiveStride > 0) {
            double factor = variant == 2 ? 1.0 : gamma;
            double var = n > 1 ? M2[d] / factor / (double)(n - 1) : 0.0;
            s = var > 0.0 ? sqrt(var) : 1.0;
            if (s < sigmaMin) s = sigmaMin;
        } else {
            s = sigma0[d] * pow(size * (D + 2) / 4.0, -1.0 / (D + 4.0));
            if (s < sigmaMin) s = sigmaMin;
        }
        sg[d] = s;
    }
    double lw_new = lw_raw;
    if (variant != 1 && adaptiveStride == 0)
        for (int d = 0; d < D; d++) lw_new += log(sigma0[d] / sg[d]);
    double h_new = exp(lw_new);

    int mk = -1;
    double mn = 1.0;
    for (int kp = 0; kp < nBefore; kp++) {
        double n2 = 0.0;
        int over = 0;
        for (int d = 0; d < D; d++) {
            double x = (c[d] - centers[(size_t)kp * D + d]) / sigmas[(size_t)kp * D + d];
            n2 += x * x;
            if (n2 >= mn) {
                over = 1;
                break;
            }
        }
        if (!over) {
            mn = n2;
            mk = kp;
        }
    }
    int nAfter, ret;
    if (mk >= 0) {
        double lo = logw[mk], ho = exp(lo), ht = ho + h_new;
        double lmax = lo > lw_new ? lo : lw_new, lmin = lo < lw_new ? lo : lw_new;
        double lm = lmax + log1p(exp(lmin - lmax));
        double f1 = ho / ht, f2 = h_new / ht;
        for (int d = 0; d < D; d++) {
            double c1 = centers[(size_t)mk * D + d], s1 = sigmas[(size_t)mk * D + d];
            double cm = f1 * c1 + f2 * c[d];
            /* f1 (s1^2 + c1^2) + f2 (s2^2 + c2^2) - cm^2 of the reference, in
               the algebraically identical form without its cancellation */
            double dcm = c1 - c[d];
            double ss = f1 * s1 * s1 + f2 * sg[d] * sg[d] + f1 * f2 * dcm * dcm;
            if (ss < 0.0) ss = 0.0;
            double sm = sqrt(ss);
            if (sm < sigmaMin) sm = sigmaMin;
            centers[(size_t)mk * D + d] = cm;
            sigmas[(size_t)mk * D + d] = sm;
        }
        logw[mk] = lm;
        nAfter = nBefore;
        ret = mk;
    } else {
        int k = st_i[1]++;
        if (k < maxKernels) {
            for (int d = 0; d < D; d++) {
                centers[(size_t)k * D + d] = c[d];
                sigmas[(size_t)k * D + d] = sg[d];
            }
            logw[k] = lw_new;
            st_i[0]++;
            nAfter = nBefore + 1;
            ret = -1;
        } else {
            st_i[1]--;
            nAfter = nBefore;
            ret = -3;
        }
    }
    st_d[1] += h_raw;
    double epsd = exp(-0.5 * cutoff2);
    if (mk < 0 && nAfter > nBefore) {
        double delta = h_new * (1.0 - epsd);
        for (int j = 0; j < nBefore; j++) {
            double n2 = 0.0;
            int skip = 0;
            for (int d = 0; d < D; d++) {
                double dc = centers[(size_t)j * D + d] - c[d];
                n2 += dc * dc / (sg[d] * sg[d]);
                if (n2 >= cutoff2) {
                    skip = 1;
                    break;
                }
            }
            if (!skip) delta += h_new * (exp(-0.5 * n2) - epsd);
        }
        for (int kp = 0; kp < nBefore; kp++) {
            double n2 = 0.0;
            int skip = 0;
            for (int d = 0; d < D; d++) {
                double dc = c[d] - centers[(size_t)kp * D + d], s = sigmas[(size_t)kp * D + d];
                n2 += dc * dc / (s * s);
                if (n2 >= cutoff2) {
                    skip = 1;
                    break;
                }
            }
            if (!skip) delta += exp(logw[kp]) * (exp(-0.5 * n2) - epsd);
        }
        st_d[0] += delta;
    } else {
        double su = 0.0;
        for (int j = 0; j < nAfter; j++)
            for (int kp = 0; kp < nAfter; kp++) {
                double n2 = 0.0;
                int skip = 0;
                for (int d = 0; d < D; d++) {
                    double dc = centers[(size_t)j * D + d] - centers[(size_t)kp * D + d],
                           s = sigmas[(size_t)kp * D + d];
                    n2 += dc * dc / (s * s);
                    if (n2 >= cutoff2) {
                        skip = 1;
                        break;
                    }
                }
                if (skip) continue;
                su += exp(logw[kp]) * (exp(-0.5 * n2) - epsd);
            }
        st_d[0] = su;
    }
    return ret;
}

/* Full O(N^2) sum_uprob of a kernel table (used to build restart blobs in    */
/* the tests and to check the incremental GPU update).                       */
EXPORT double orc_opes_sum_uprob(int nk, int D, const double* centers, const double* sigmas, const double* logw,
                                 double cutoff2)
{
    double epsd = exp(-0.5 * cutoff2), su = 0.0;
    for (int j = 0; j < nk; j++)
        for (int kp = 0; kp < nk; kp++) {
            double n2 = 0.0;
            int skip = 0;
            for (int d = 0; d < D; d++) {
                double dc = centers[(size_t)j * D + d] - centers[(size_t)kp * D + d], s = sigmas[(size_t)kp * D + d];
                n2 += dc * dc / (s * s);
                if (n2 >= cutoff2) {
                    skip = 1;
                    break;
                }
            }
            if (skip) continue;
            su += exp(logw[kp]) * (exp(-0.5 * n2) - epsd);
        }
    return su;
}

/* ------------------------------------------------------------------------ */
/* OPES expanded ensemble.  CommonGluedKernels.cpp:3050-3125.                 */
/* ------------------------------------------------------------------------ */
EXPORT double orc_opes_expanded_eval(const double* ecv, int n, const double* logW, double invKT, double logZ,
                                     double* dV)
{
    double mx = -1e300;
    for (int l = 0; l < n; l++) {
        double lw = logW[l] - ecv[l] * invKT;
        if (lw > mx) mx = lw;
    }
    double se = 0.0;
    for (int l = 0; l < n; l++) se += exp(logW[l] - ecv[l] * invKT - mx);
    double logQ = mx + log(se);
    for (int l = 0; l < n; l++) dV[l] = exp(logW[l] - ecv[l] * invKT - logQ);
    return -(1.0 / invKT) * (logQ - logZ);
}

EXPORT void orc_opes_expanded_update_logz(const double* ecv, int n, const double* logW, double invKT, double* logZ,
                                          int* nUpd)
{
    double mx = -1e300;
    for (int l = 0; l < n; l++) {
        double lw = logW[l] - ecv[l] * invKT;
        if (lw > mx) mx = lw;
    }
    double se = 0.0;
    for (int l = 0; l < n; l++) se += exp(logW[l] - ecv[l] * invKT - mx);
    double logQ = mx + log(se);
    int k = nUpd[0];
    if (k == 0) {
        logZ[0] = logQ;
    } else {
        double A = logZ[0] + log((double)k), B = logQ, m = A > B ? A : B;
        logZ[0] = m + log(exp(A - m) + exp(B - m)) - log((double)(k + 1));
    }
    nUpd[0] = k + 1;
}

/* ------------------------------------------------------------------------ */
/* OPES multithermal.  CommonGluedKernels.cpp:3145-3203.                      */
/* dbeta[l] = beta_l - beta_0.                                               */
/* ------------------------------------------------------------------------ */
EXPORT double orc_multithermal_eval(double U, const double* dbeta, const double* dF, int n, double kT0, double* dVdU)
{
    double mx = -1e300;
    for (int l = 0; l < n; l++) {
        double d = -dbeta[l] * U + dF[l] / kT0;
        if (d > mx) mx = d;
    }
    double s = 0.0, ds = 0.0;
    for (int l = 0; l < n; l++) {
        double e = exp(-dbeta[l] * U + dF[l] / kT0 - mx);
        s += e;
        ds += e * dbeta[l];
    }
    *dVdU = kT0 * (ds / s);
    return -kT0 * (mx + log(s / (double)n));
}

EXPORT void orc_multithermal_update(double U, double V, const double* dbeta, double* dF, double* rct,
                                    double* counter, int n, double kT0)
{
    double r = rct[0];
    counter[0] += 1.0;
    double cnt = counter[0], logc = log(cnt - 1.0);
    double arg = (V - r) / kT0 - logc;
    double inc = arg > 0.0 ? kT0 * (arg + log1p(exp(-arg))) : kT0 * log1p(exp(arg));
    for (int l = 0; l < n; l++) {
        double dl = -dbeta[l] * U + (V - r + dF[l]) / kT0 - logc;
        double dec = dl > 0.0 ? kT0 * (dl + log1p(exp(-dl))) : kT0 * log1p(exp(dl));
        dF[l] += inc - dec;
    }
    rct[0] = r + inc + kT0 * log1p(-1.0 / cnt);
}

/* ------------------------------------------------------------------------ */
/* Static restraints (no state, no deposits).  Parameters arrive already    */
/* rounded to FP32, as the reference stores them (CommonBiasKernels.cpp:    */
/* 118-126, 676-678, 700-707).                                               */
/* ------------------------------------------------------------------------ */

/* Harmonic, CommonGluedKernels.cpp:2506-2540: V = sum k/2 (s - s0)^2; with  */
/* the periodic flag the difference is folded to the nearest 2 pi image.     */
/* p = [k_0, s0_0, k_1, s0_1, ...].                                          */
EXPORT double orc_harmonic_eval(const double* s, int D, const double* p, int periodic, double* g)
{
    const double TWO_PI = 6.283185307179586;
    double V = 0.0;
    for (int d = 0; d < D; d++) {
        double ds = s[d] - p[2 * d + 1];
        if (periodic) ds -= TWO_PI * round(ds / TWO_PI);
        V += 0.5 * p[2 * d] * ds * ds;
        g[d] = p[2 * d] * ds;
    }
    return V;
}

/* Linear, CommonGluedKernels.cpp:2542-2566: V = sum k s, dV/ds = k.         */
EXPORT double orc_linear_eval(const double* s, int D, const double* k, double* g)
{
    double V = 0.0;
    for (int d = 0; d < D; d++) {
        V += k[d] * s[d];
        g[d] = k[d];
    }
    return V;
}

/* One-sided wall, CommonGluedKernels.cpp:2568-2616.  lower = 0: delta =    */
/* s - at, else at - s; only delta > 0 counts: V = kappa delta^n e^(eps      */
/* delta), dV/ds = +-kappa (n delta^(n-1) + eps delta^n) e^(eps delta), with */
/* n = 1 and n = 2 in closed form and pow() otherwise.                       */
/* p = [at, kappa, eps, n] per CV.                                           */
EXPORT double orc_wall_eval(const double* s, int D, const double* p, int lower, double* g)
{
    double V = 0.0;
    for (int d = 0; d < D; d++) {
        const double at = p[4 * d], kappa = p[4 * d + 1], eps = p[4 * d + 2], n = p[4 * d + 3];
        const double delta = lower ? at - s[d] : s[d] - at;
        g[d] = 0.0;
        if (delta <= 0.0) continue;
        double powered, dPow;
        if (n == 1.0) {
            powered = delta;
            dPow = 1.0;
        } else if (n == 2.0) {
            powered = delta * delta;
            dPow = 2.0 * delta;
        } else {
            powered = pow(delta, n);
            dPow = n * pow(delta, n - 1.0);
        }
        const double e = exp(eps * delta);
        V += kappa * powered * e;
        g[d] = (lower ? -1.0 : 1.0) * kappa * (dPow + eps * powered) * e;
    }
    return V;
}

/* ------------------------------------------------------------------------ */
/* PBMetaD, CommonGluedKernels.cpp:3697-3795.  One 1-D grid per CV;         */
/* c = {origin, 1/spacing, points, periodic, sigma, offset} per sub-grid,   */
/* grid = the sub-grids back to back.                                        */
/* ------------------------------------------------------------------------ */
EXPORT double orc_pbmetad_eval(const double* s, int N, const double* c, const double* grid, double kT, double* Vi,
                               double* wi, double* g)
{
    double gi[64];
    for (int i = 0; i < N; i++) {
        const double* ci = c + 6 * i;
        const int np = (int)ci[2];
        double frac = (s[i] - ci[0]) * ci[1];
        int lo, hi;
        if (ci[3] != 0.0) {
            frac = frac - floor(frac / (double)np) * (double)np;
            if (frac < 0.0) frac += (double)np;
            lo = (int)frac;
            if (lo >= np) lo = np - 1;
            hi = (lo + 1) % np;
        } else {
            if (frac < 0.0) frac = 0.0;
            if (frac > (double)(np - 1)) frac = (double)(np - 1);
            lo = (int)frac;
            if (lo >= np - 1) lo = np - 2;
            hi = lo + 1;
        }
        const double alpha = frac - lo;
        const double* gr = grid + (int)ci[5];
        Vi[i] = gr[lo] * (1.0 - alpha) + gr[hi] * alpha;
        gi[i] = (gr[hi] - gr[lo]) * ci[1];
    }
    double bmin = Vi[0];
    for (int i = 1; i < N; i++)
        if (Vi[i] < bmin) bmin = Vi[i];
    double ene = 0.0;
    for (int i = 0; i < N; i++) ene += exp((bmin - Vi[i]) / kT);
    for (int i = 0; i < N; i++) {
        wi[i] = exp((bmin - Vi[i]) / kT) / ene;
        g[i] = wi[i] * gi[i];
    }
    return -kT * (log(ene) - log((double)N)) + bmin;
}

/* One sub-grid deposit: a Gaussian of the given height at `center`.        */
EXPORT void orc_pbmetad_deposit(double* grid, const double* ci, double center, double height)
{
    const int np = (int)ci[2];
    const double per = (double)np / ci[1], sigma = ci[4];
    for (int idx = 0; idx < np; idx++) {
        const double pos = ci[0] + (double)idx / ci[1];
        double dc = pos - center;
        if (ci[3] != 0.0) dc -= round(dc / per) * per;
        const double e = dc * dc / (2.0 * sigma * sigma);
        if (e <= 8.0) grid[idx] += height * exp(-e);
    }
}
