/*
 * CPU oracle for the short-range multipole operators of PhotoNs
 * (src/operator.c): the M2L translation with the erfc-modified kernel of the
 * long/short split, the L2P evaluation of a local expansion and the P2M
 * moments, with QUADRUPOLE and OCTUPOLE as inc/typedef.h:12-13 define them
 * (NMULTI = 20 coefficients).
 *
 * TEST INFRASTRUCTURE ONLY: called by tests/ and by bench.py's accuracy
 * check, never by the product library.
 *
 * Two forms:
 *   *_ref32  the reference's FP32 arithmetic, operation for operation, so
 *            that the results equal the reference's own compiled functions
 *            bit for bit (tests/golden/m2l_ref.npz, made from them);
 *   *_f64    an independent FP64 form: the radial derivatives of
 *            phi(r) = erfc(r / 2 rs) / r derived here, and the M2L / L2P /
 *            P2M contractions written as sums over multi-indices.  Pinned by
 *            finite differences of phi and by the Taylor truncation order
 *            against direct sums (tests/test_oracle_multipole.py).
 *
 * Coefficient layout (inc/operator.h): 0; X Y Z; XX XY XZ YY YZ ZZ;
 * XXX XXY XXZ XYY XYZ XZZ YYY YYZ YZZ ZZZ.  M[a] holds sum m (-d)^a / a!
 * (src/operator.c:14-100); L[a] holds the a-th derivative of the potential.
 */
#include <math.h>
#include <stdlib.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#include "oracle.h"

enum { X = 1, Y, Z, XX, XY, XZ, YY, YZ, ZZ, XXX, XXY, XXZ, XYY, XYZ, XZZ, YYY, YYZ, YZZ, ZZZ };

/* exponents (a, b, c) of each coefficient index */
static const int kExp[ORACLE_NMULTI][3] = {
    {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2},
    {3, 0, 0}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {1, 1, 1}, {1, 0, 2}, {0, 3, 0}, {0, 2, 1}, {0, 1, 2}, {0, 0, 3}};

static int index_of(int a, int b, int c)
{
    for (int k = 0; k < ORACLE_NMULTI; ++k)
        if (kExp[k][0] == a && kExp[k][1] == b && kExp[k][2] == c)
            return k;
    return -1;
}

/* ------------------------------------------------------------------ FP32 */

/* src/operator.c:331-474 under LONGSHORT, QUADRUPOLE, OCTUPOLE: the same
 * operations in the same order and types (float cpuType, double literals
 * and libm calls promoting where the reference's expressions do). */
void oracle_m2l_ref32(float dx, float dy, float dz, const float *M, float rs, float *L)
{
    const float r2 = dx * dx + dy * dy + dz * dz;
    /* abs() of the reference is the integer one: every r2 < 1 returns here */
    if (abs((int)r2) <= 1.0e-15)
        return;
    const float dr = (float)sqrt((double)r2);
    const float ir = (float)(1.0 / dr);
    const float ir2 = ir * ir;
    const float ir3 = ir * ir * ir;
    const float ir4 = ir3 * ir;
    const float ir5 = ir * ir * ir3;
    const float ir6 = ir2 * ir4;
    const float ir7 = ir5 * ir2;
    const float irs = (float)(1.0 / rs);
    const float irs2 = irs * irs;
    const float irs3 = irs2 * irs;
    const float irs5 = irs3 * irs2;
    const float drs = (float)(0.5 * dr / rs);
    const float coeff = (float)(1.0 / sqrt(M_PI));
    const float fexp = (float)(exp((double)(-drs * drs)) * coeff);
    const float ferc = (float)erfc((double)drs);
    const float f0 = ir * ferc;
    const float f1 = -ir3 * (ferc + dr * fexp * irs);
    const float f2 = (float)(3.0 * ir5 * ferc + (3.0 * irs * ir4 + 0.5 * ir2 * irs3) * fexp);
    const float f3 = (float)(-3.0 * 5.0 * ir7 * ferc - (15.0 * ir6 * irs + 2.5 * ir4 * irs3 + 0.25 * ir2 * irs5) * fexp);

    float D[ORACLE_NMULTI], F[ORACLE_NMULTI];
    D[0] = f0;
    D[X] = f1 * dx;
    D[Y] = f1 * dy;
    D[Z] = f1 * dz;
    F[0] = M[0] * D[0] + M[X] * D[X] + M[Y] * D[Y] + M[Z] * D[Z];
    F[X] = M[0] * D[X];
    F[Y] = M[0] * D[Y];
    F[Z] = M[0] * D[Z];

    D[XX] = f2 * dx * dx + f1;
    D[XY] = f2 * dx * dy;
    D[XZ] = f2 * dx * dz;
    D[YY] = f2 * dy * dy + f1;
    D[YZ] = f2 * dy * dz;
    D[ZZ] = f2 * dz * dz + f1;
    F[0] += M[XX] * D[XX] + M[XY] * D[XY] + M[XZ] * D[XZ] + M[YY] * D[YY] + M[YZ] * D[YZ] + M[ZZ] * D[ZZ];
    F[X] += M[X] * D[XX] + M[Y] * D[XY] + M[Z] * D[XZ];
    F[Y] += M[X] * D[XY] + M[Y] * D[YY] + M[Z] * D[YZ];
    F[Z] += M[X] * D[XZ] + M[Y] * D[YZ] + M[Z] * D[ZZ];
    for (int k = XX; k <= ZZ; ++k)
        F[k] = M[0] * D[k];

    D[XXX] = f3 * dx * dx * dx + f2 * dx * 3;
    D[XYY] = f3 * dx * dy * dy + f2 * dx;
    D[XZZ] = f3 * dx * dz * dz + f2 * dx;
    D[XXY] = f3 * dx * dx * dy + f2 * dy;
    D[YYY] = f3 * dy * dy * dy + f2 * dy * 3;
    D[YZZ] = f3 * dy * dz * dz + f2 * dy;
    D[XXZ] = f3 * dx * dx * dz + f2 * dz;
    D[YYZ] = f3 * dy * dy * dz + f2 * dz;
    D[ZZZ] = f3 * dz * dz * dz + f2 * dz * 3;
    D[XYZ] = f3 * dx * dy * dz;
    F[0] += M[XXX] * D[XXX] + M[XXY] * D[XXY] + M[XXZ] * D[XXZ] + M[XYY] * D[XYY] + M[XYZ] * D[XYZ] +
            M[YYY] * D[YYY] + M[YYZ] * D[YYZ] + M[YZZ] * D[YZZ] + M[XZZ] * D[XZZ] + M[ZZZ] * D[ZZZ];
    F[X] += M[XX] * D[XXX] + M[XY] * D[XXY] + M[XZ] * D[XXZ] + M[YY] * D[XYY] + M[YZ] * D[XYZ] + M[ZZ] * D[XZZ];
    F[Y] += M[YY] * D[YYY] + M[YZ] * D[YYZ] + M[XY] * D[XYY] + M[ZZ] * D[YZZ] + M[XZ] * D[XYZ] + M[XX] * D[XXY];
    F[Z] += M[ZZ] * D[ZZZ] + M[XZ] * D[XZZ] + M[YZ] * D[YZZ] + M[XX] * D[XXZ] + M[XY] * D[XYZ] + M[YY] * D[YYZ];
    F[XX] += M[X] * D[XXX] + M[Y] * D[XXY] + M[Z] * D[XXZ];
    F[XY] += M[X] * D[XXY] + M[Y] * D[XYY] + M[Z] * D[XYZ];
    F[XZ] += M[X] * D[XXZ] + M[Y] * D[XYZ] + M[Z] * D[XZZ];
    F[YY] += M[X] * D[XYY] + M[Y] * D[YYY] + M[Z] * D[YYZ];
    F[YZ] += M[X] * D[XYZ] + M[Y] * D[YYZ] + M[Z] * D[YZZ];
    F[ZZ] += M[X] * D[XZZ] + M[Y] * D[YZZ] + M[Z] * D[ZZZ];
    for (int k = XXX; k <= ZZZ; ++k)
        F[k] = M[0] * D[k];

    for (int k = 0; k < ORACLE_NMULTI; ++k)
        L[k] += F[k];
}

/* src/operator.c:203-262 for one leaf: acc[d][n] += the force of the local
 * expansion L about center at each particle n of the leaf (FP32). */
void oracle_l2p_ref32(int npart, const int *posi, double int2pos, const float *center, const float *L,
                      float *acc_x, float *acc_y, float *acc_z)
{
    for (int n = 0; n < npart; ++n) {
        const float dx = (float)(int2pos * posi[3 * n]) - center[0];
        const float dy = (float)(int2pos * posi[3 * n + 1]) - center[1];
        const float dz = (float)(int2pos * posi[3 * n + 2]) - center[2];
        float fx = L[X], fy = L[Y], fz = L[Z];
        fx += (L[XX] * dx + L[XY] * dy + L[XZ] * dz);
        fy += (L[XY] * dx + L[YY] * dy + L[YZ] * dz);
        fz += (L[XZ] * dx + L[YZ] * dy + L[ZZ] * dz);
        fx += L[XXX] * dx * dx / 2 + L[XXY] * dx * dy + L[XXZ] * dx * dz + L[XYZ] * dy * dz + L[XYY] * dy * dy / 2 +
              L[XZZ] * dz * dz / 2;
        fy += L[YYY] * dy * dy / 2 + L[YYZ] * dy * dz + L[XYY] * dx * dy + L[XYZ] * dx * dz + L[YZZ] * dz * dz / 2 +
              L[XXY] * dx * dx / 2;
        fz += L[ZZZ] * dz * dz / 2 + L[XZZ] * dx * dz + L[YZZ] * dy * dz + L[XYZ] * dx * dy + L[XXZ] * dx * dx / 2 +
              L[YYZ] * dy * dy / 2;
        acc_x[n] += fx;
        acc_y[n] += fy;
        acc_z[n] += fz;
    }
}

/* src/operator.c:14-98: the multipole of npart particles (integer grid
 * positions) about center, mass per particle `mass` (FP32). */
void oracle_p2m_ref32(int npart, const int *posi, double int2pos, const float *center, float mass, float *M)
{
    for (int n = 0; n < ORACLE_NMULTI; ++n)
        M[n] = 0.0f;
    for (int n = 0; n < npart; ++n) {
        const float d0 = (float)(int2pos * posi[3 * n]) - center[0];
        const float d1 = (float)(int2pos * posi[3 * n + 1]) - center[1];
        const float d2 = (float)(int2pos * posi[3 * n + 2]) - center[2];
        const float m0 = mass, m1 = -m0, m2 = m0, m3 = -m0;
        M[0] += m0;
        M[X] += m1 * d0;
        M[Y] += m1 * d1;
        M[Z] += m1 * d2;
        M[XX] += m2 * d0 * d0 / 2;
        M[XY] += m2 * d0 * d1;
        M[XZ] += m2 * d0 * d2;
        M[YY] += m2 * d1 * d1 / 2;
        M[YZ] += m2 * d1 * d2;
        M[ZZ] += m2 * d2 * d2 / 2;
        M[XXX] += m3 * d0 * d0 * d0 / 6;
        M[XXY] += m3 * d1 * d0 * d0 / 2;
        M[XXZ] += m3 * d2 * d0 * d0 / 2;
        M[XYY] += m3 * d1 * d1 * d0 / 2;
        M[XYZ] += m3 * d2 * d1 * d0;
        M[XZZ] += m3 * d2 * d2 * d0 / 2;
        M[YYY] += m3 * d1 * d1 * d1 / 6;
        M[YYZ] += m3 * d1 * d2 * d1 / 2;
        M[YZZ] += m3 * d2 * d2 * d1 / 2;
        M[ZZZ] += m3 * d2 * d2 * d2 / 6;
    }
}

/* src/operator.c:100-165: tM += M moved to a centre (dx, dy, dz) = new
 * centre - old centre (FP32). */
void oracle_m2m_ref32(float dx, float dy, float dz, const float *M, float *tM)
{
    float t[ORACLE_NMULTI];
    t[0] = M[0];
    t[X] = M[X] + M[0] * dx;
    t[Y] = M[Y] + M[0] * dy;
    t[Z] = M[Z] + M[0] * dz;
    t[XX] = M[XX] + M[X] * dx + M[0] * dx * dx / 2;
    t[XY] = M[XY] + M[X] * dy + M[Y] * dx + M[0] * dx * dy;
    t[XZ] = M[XZ] + M[X] * dz + M[Z] * dx + M[0] * dx * dz;
    t[YY] = M[YY] + M[Y] * dy + M[0] * dy * dy / 2;
    t[YZ] = M[YZ] + M[Y] * dz + M[Z] * dy + M[0] * dy * dz;
    t[ZZ] = M[ZZ] + M[Z] * dz + M[0] * dz * dz / 2;
    t[XXX] = M[XXX] + M[XX] * dx + M[X] * dx * dx / 2 + M[0] * dx * dx * dx / 6;
    t[YYY] = M[YYY] + M[YY] * dy + M[Y] * dy * dy / 2 + M[0] * dy * dy * dy / 6;
    t[ZZZ] = M[ZZZ] + M[ZZ] * dz + M[Z] * dz * dz / 2 + M[0] * dz * dz * dz / 6;
    t[XXY] = M[XXY] + M[XX] * dy + M[XY] * dx + M[X] * dx * dy + M[Y] * dx * dx / 2 + M[0] * dx * dx * dy / 2;
    t[XXZ] = M[XXZ] + M[XX] * dz + M[XZ] * dx + M[X] * dx * dz + M[Z] * dx * dx / 2 + M[0] * dx * dx * dz / 2;
    t[XYY] = M[XYY] + M[YY] * dx + M[XY] * dy + M[Y] * dx * dy + M[X] * dy * dy / 2 + M[0] * dx * dy * dy / 2;
    t[XZZ] = M[XZZ] + M[ZZ] * dx + M[XZ] * dz + M[Z] * dx * dz + M[X] * dz * dz / 2 + M[0] * dx * dz * dz / 2;
    t[YZZ] = M[YZZ] + M[ZZ] * dy + M[YZ] * dz + M[Z] * dy * dz + M[Y] * dz * dz / 2 + M[0] * dy * dz * dz / 2;
    t[YYZ] = M[YYZ] + M[YY] * dz + M[YZ] * dy + M[Y] * dy * dz + M[Z] * dy * dy / 2 + M[0] * dy * dy * dz / 2;
    t[XYZ] = M[XYZ] + M[XY] * dz + M[YZ] * dx + M[XZ] * dy + M[X] * dy * dz + M[Y] * dx * dz + M[Z] * dy * dx +
             M[0] * dx * dy * dz;
    for (int n = 0; n < ORACLE_NMULTI; ++n)
        tM[n] += t[n];
}

/* src/operator.c:477-580: toL += L moved to a centre (dx, dy, dz) = new
 * centre - old centre (FP32; the octupole terms of the first-order
 * coefficients are summed in FP64 there, as the reference's 0.5 literals do). */
void oracle_l2l_ref32(float dx, float dy, float dz, const float *L, float *toL)
{
    toL[0] += L[0] + L[X] * dx + L[Y] * dy + L[Z] * dz;
    toL[X] += L[X];
    toL[Y] += L[Y];
    toL[Z] += L[Z];
    toL[0] += L[XX] * dx * dx / 2 + L[XY] * dx * dy + L[XZ] * dx * dz + L[YZ] * dy * dz + L[YY] * dy * dy / 2 +
              L[ZZ] * dz * dz / 2;
    toL[X] += L[XX] * dx + L[XY] * dy + L[XZ] * dz;
    toL[Y] += L[XY] * dx + L[YY] * dy + L[YZ] * dz;
    toL[Z] += L[XZ] * dx + L[YZ] * dy + L[ZZ] * dz;
    for (int n = XX; n <= ZZ; ++n)
        toL[n] += L[n];
    toL[0] += L[XXX] * dx * dx * dx / 6 + L[XXY] * dx * dx * dy / 2 + L[XXZ] * dx * dx * dz / 2 +
              L[XYY] * dx * dy * dy / 2 + L[XYZ] * dx * dy * dz + L[XZZ] * dx * dz * dz / 2 +
              L[YYY] * dy * dy * dy / 6 + L[YZZ] * dy * dz * dz / 2 + L[YYZ] * dy * dy * dz / 2 +
              L[ZZZ] * dz * dz * dz / 6;
    toL[X] += L[XXX] * dx * dx * 0.5 + L[XXY] * dx * dy + L[XXZ] * dx * dz + L[XYY] * dy * dy * 0.5 +
              L[XYZ] * dy * dz + L[XZZ] * dz * dz * 0.5;
    toL[Y] += L[XXY] * dx * dx * 0.5 + L[XYY] * dx * dy + L[XYZ] * dx * dz + L[YYY] * dy * dy * 0.5 +
              L[YYZ] * dy * dz + L[YZZ] * dz * dz * 0.5;
    toL[Z] += L[XXZ] * dx * dx * 0.5 + L[XYZ] * dx * dy + L[XZZ] * dx * dz + L[YYZ] * dy * dy * 0.5 +
              L[YZZ] * dy * dz + L[ZZZ] * dz * dz * 0.5;
    toL[XX] += L[XXX] * dx + L[XXY] * dy + L[XXZ] * dz;
    toL[XY] += L[XXY] * dx + L[XYY] * dy + L[XYZ] * dz;
    toL[XZ] += L[XXZ] * dx + L[XYZ] * dy + L[XZZ] * dz;
    toL[YY] += L[XYY] * dx + L[YYY] * dy + L[YYZ] * dz;
    toL[YZ] += L[XYZ] * dx + L[YYZ] * dy + L[YZZ] * dz;
    toL[ZZ] += L[XZZ] * dx + L[YZZ] * dy + L[ZZZ] * dz;
    toL[XXX] += L[XXX];
    toL[XZZ] += L[XZZ];
    toL[XYY] += L[XYY];
    toL[XXY] += L[XXY];
    toL[XXZ] += L[XXZ];
    toL[YYY] += L[YYY];
    toL[XYZ] += L[XYZ];
    toL[YYZ] += L[YYZ];
    toL[YZZ] += L[YZZ];
    toL[ZZZ] += L[ZZZ];
}

/* ------------------------------------------------------------------ FP64 */

/* f[k] = ((1/r) d/dr)^k phi for phi(r) = erfc(r / 2 rs) / r, k = 0..3, with
 * q = exp(-r^2 / 4 rs^2) / (rs sqrt(pi)) = -d/dr erfc(r / 2 rs):
 *   f0 = E/r
 *   f1 = -E/r^3 - q/r^2
 *   f2 = 3E/r^5 + 3q/r^4 + q/(2 rs^2 r^2)
 *   f3 = -15E/r^7 - 15q/r^6 - 5q/(2 rs^2 r^4) - q/(4 rs^4 r^2) */
void oracle_radial_f64(double r, double rs, double *f)
{
    const double E = erfc(r / (2.0 * rs));
    const double q = exp(-r * r / (4.0 * rs * rs)) / (rs * sqrt(M_PI));
    const double s2 = rs * rs;
    f[0] = E / r;
    f[1] = -E / pow(r, 3) - q / (r * r);
    f[2] = 3.0 * E / pow(r, 5) + 3.0 * q / pow(r, 4) + q / (2.0 * s2 * r * r);
    f[3] = -15.0 * E / pow(r, 7) - 15.0 * q / pow(r, 6) - 5.0 * q / (2.0 * s2 * pow(r, 4)) -
           q / (4.0 * s2 * s2 * r * r);
}

/* derivative of phi with exponents e (order <= 3) at x, from the radial f */
static double deriv(const int e[3], const double x[3], const double f[4])
{
    int idx[3], n = 0;
    for (int d = 0; d < 3; ++d)
        for (int k = 0; k < e[d]; ++k)
            idx[n++] = d;
    switch (n) {
    case 0:
        return f[0];
    case 1:
        return f[1] * x[idx[0]];
    case 2:
        return f[2] * x[idx[0]] * x[idx[1]] + (idx[0] == idx[1] ? f[1] : 0.0);
    default: {
        const int i = idx[0], j = idx[1], k = idx[2];
        return f[3] * x[i] * x[j] * x[k] +
               f[2] * ((i == j ? x[k] : 0.0) + (i == k ? x[j] : 0.0) + (j == k ? x[i] : 0.0));
    }
    }
}

/* L[b] += sum over a with |a| + |b| <= 3 of M[a] * D(a + b) at (dx, dy, dz) =
 * target centre - source centre. */
void oracle_m2l_f64(double dx, double dy, double dz, const double *M, double rs, double *L)
{
    const double x[3] = {dx, dy, dz};
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    double f[4];
    oracle_radial_f64(r, rs, f);
    for (int b = 0; b < ORACLE_NMULTI; ++b) {
        const int nb = kExp[b][0] + kExp[b][1] + kExp[b][2];
        double s = 0.0;
        for (int a = 0; a < ORACLE_NMULTI; ++a) {
            if (kExp[a][0] + kExp[a][1] + kExp[a][2] + nb > 3)
                continue;
            const int e[3] = {kExp[a][0] + kExp[b][0], kExp[a][1] + kExp[b][1], kExp[a][2] + kExp[b][2]};
            s += M[a] * deriv(e, x, f);
        }
        L[b] += s;
    }
}

/* F_i = sum over b with |b| <= 2 of L[b + e_i] d^b / b!  (the gradient of the
 * local expansion at offset d from its centre). */
void oracle_l2p_f64(double dx, double dy, double dz, const double *L, double *F)
{
    const double d[3] = {dx, dy, dz};
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int b = 0; b < ORACLE_NMULTI; ++b) {
            const int *e = kExp[b];
            if (e[0] + e[1] + e[2] > 2)
                continue;
            const int k = index_of(e[0] + (i == 0), e[1] + (i == 1), e[2] + (i == 2));
            double t = L[k];
            for (int c = 0; c < 3; ++c)
                for (int p = 0; p < e[c]; ++p)
                    t *= d[c] / (double)(p + 1);
            s += t;
        }
        F[i] = s;
    }
}

/* M[a] = sum over particles of mass (-d)^a / a!, d = position - center. */
void oracle_p2m_f64(int n, const double *pos, double mass, const double *center, double *M)
{
    for (int a = 0; a < ORACLE_NMULTI; ++a)
        M[a] = 0.0;
    for (int p = 0; p < n; ++p) {
        const double d[3] = {pos[3 * p] - center[0], pos[3 * p + 1] - center[1], pos[3 * p + 2] - center[2]};
        for (int a = 0; a < ORACLE_NMULTI; ++a) {
            double t = mass;
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < kExp[a][c]; ++k)
                    t *= -d[c] / (double)(k + 1);
            M[a] += t;
        }
    }
}

/* The M2L task loop (task_compute_m2l, src/fmm.c:1188-1217) in FP64: for each
 * task, L[t] += M2L(center[t] - center[s], M[s]); separations with r^2 < 1
 * are skipped as in the reference. */
void oracle_m2l_tasks_f64(int ncell, const double *center, const double *M, long long ntask, const int *task_t,
                          const int *task_s, double rs, double *L)
{
    (void)ncell;
    for (long long k = 0; k < ntask; ++k) {
        const int t = task_t[k], s = task_s[k];
        const double dx = center[3 * t] - center[3 * s];
        const double dy = center[3 * t + 1] - center[3 * s + 1];
        const double dz = center[3 * t + 2] - center[3 * s + 2];
        if (dx * dx + dy * dy + dz * dz < 1.0)
            continue;
        oracle_m2l_f64(dx, dy, dz, M + (size_t)ORACLE_NMULTI * s, rs, L + (size_t)ORACLE_NMULTI * t);
    }
}

/* The reference's task loop (src/fmm.c:1188-1217) in its FP32 arithmetic:
 * L[t] += m2l(center[t] - center[s], M[s]) task by task. */
void oracle_m2l_tasks_ref32(int ncell, const float *center, const float *M, long long ntask, const int *task_t,
                            const int *task_s, float rs, float *L)
{
    (void)ncell;
    for (long long k = 0; k < ntask; ++k) {
        const int t = task_t[k], s = task_s[k];
        oracle_m2l_ref32(center[3 * t] - center[3 * s], center[3 * t + 1] - center[3 * s + 1],
                         center[3 * t + 2] - center[3 * s + 2], M + (size_t)ORACLE_NMULTI * s, rs,
                         L + (size_t)ORACLE_NMULTI * t);
    }
}

/* tM[a] += sum over b <= a of M[b] s^(a-b) / (a-b)!: the moments about a
 * centre moved by s = new - old (since M[a] = sum m (-d)^a / a!). */
void oracle_m2m_f64(double sx, double sy, double sz, const double *M, double *tM)
{
    const double sv[3] = {sx, sy, sz};
    for (int a = 0; a < ORACLE_NMULTI; ++a) {
        double acc = 0.0;
        for (int b = 0; b < ORACLE_NMULTI; ++b) {
            int e[3], ok = 1;
            for (int c = 0; c < 3; ++c) {
                e[c] = kExp[a][c] - kExp[b][c];
                ok &= e[c] >= 0;
            }
            if (!ok)
                continue;
            double t = M[b];
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < e[c]; ++k)
                    t *= sv[c] / (double)(k + 1);
            acc += t;
        }
        tM[a] += acc;
    }
}

/* toL[b] += sum over a >= b, |a| <= 3, of L[a] s^(a-b) / (a-b)!: the Taylor
 * coefficients about a centre moved by s = new - old. */
void oracle_l2l_f64(double sx, double sy, double sz, const double *L, double *toL)
{
    const double sv[3] = {sx, sy, sz};
    for (int b = 0; b < ORACLE_NMULTI; ++b) {
        double acc = 0.0;
        for (int a = 0; a < ORACLE_NMULTI; ++a) {
            int e[3], ok = 1;
            for (int c = 0; c < 3; ++c) {
                e[c] = kExp[a][c] - kExp[b][c];
                ok &= e[c] >= 0;
            }
            if (!ok)
                continue;
            double t = L[a];
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < e[c]; ++k)
                    t *= sv[c] / (double)(k + 1);
            acc += t;
        }
        toL[b] += acc;
    }
}
