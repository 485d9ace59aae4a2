/* vvt_oracle.c -- CPU oracle of the 3D horizontal-convection tank step
 * (test infrastructure; see vvt_oracle.h for what it composes).
 *
 * Every routine follows the routine of the reference it names, term by term
 * and in the same operation order, with the metric factors of the annulus
 * (r, 1/r, 1/r^2) and the centrifugal buoyancy removed (the Cartesian form,
 * as velvor2dhorconvtank.f takes the annulus plane's code to r = 1), and the
 * hybrid-upwind advection derivatives of rot3dmg.f replaced by the compact
 * scheme of the 2D codes.  Compiled with -ffp-contract=off: the GPU kernels
 * evaluate the same expressions and must match bit for bit.
 */
#include "vvt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define X(a, i, j, k) (a)[((size_t)(k) * P->ldj + (size_t)(j)) * P->ldi + (size_t)(i)]

/* ------------------------------------------------------- compact lines */

/* rows of a line of m unknowns (1-based): a (sub), b (diagonal), c (super) */
static void rows(int type, int m, double *a, double *b, double *c)
{
    for (int i = 1; i <= m; ++i) {
        a[i] = 1.0;
        b[i] = 4.0;
        c[i] = 1.0;
    }
    if (type == 1) {   /* velvor2dperiodicrbconv.f:1085-1092 */
        b[1] = 1.0;
        c[1] = 2.0;
        a[m] = 2.0;
        b[m] = 1.0;
    }
}

/* tdma's elimination (velvor2dperiodicrbconv.f:3050-3054) on the constant
 * rows: fact(j) = a(j)/b'(j-1), b'(j) = b(j) - fact(j) c(j-1) */
void vt_compact_tables(int type, int m, double *fact, double *bp, double *c)
{
    double *a = (double *)malloc((size_t)(m + 1) * sizeof(double));
    rows(type, m, a, bp, c);
    fact[1] = 0.0;
    for (int j = 2; j <= m; ++j) {
        fact[j] = a[j] / bp[j - 1];
        bp[j] = bp[j] - fact[j] * c[j - 1];
    }
    free(a);
}

/* right-hand side of row i of a line (nodes f[0..m+1]); dri = 1/h */
static double line_rhs(int type, int m, int i, double dri, const double *f)
{
    if (type == 0) {   /* velvor2dperiodicrbconv.f:613-620 */
        if (i == 1) return 3.0 * (f[2] - f[0]) * dri - (-1.5 * f[0] + 2.0 * f[1] - 0.5 * f[2]) * dri;
        if (i == m) return 3.0 * (f[m + 1] - f[m - 1]) * dri - (0.5 * f[m - 1] - 2.0 * f[m] + 1.5 * f[m + 1]) * dri;
        return 3.0 * (f[i + 1] - f[i - 1]) * dri;
    }
    /* velvor2dperiodicrbconv.f:1085-1097 */
    if (i == 1) return dri * (0.533333333 * f[0] - 3.6 * f[1] + 2.966666667 * f[2] + 0.1 * f[4]);
    if (i == m) return dri * (-0.1 * f[m - 3] - 2.966666667 * f[m - 1] + 3.6 * f[m] - 0.533333333 * f[m + 1]);
    return 3.0 * (f[i + 1] - f[i - 1]) * dri;
}

/* the solve of tdma (velvor2dperiodicrbconv.f:3050-3062) with the tables */
static void line_solve(int m, const double *fact, const double *bp, const double *c, double *r, double *y)
{
    for (int j = 2; j <= m; ++j) r[j] = r[j] - fact[j] * r[j - 1];
    y[m] = r[m] / bp[m];
    for (int j = m - 1; j >= 1; --j) y[j] = (r[j] - c[j] * y[j + 1]) / bp[j];
}

void vt_compact_line(int type, int m, double h, const double *f, double *out)
{
    double *fact = (double *)malloc((size_t)(m + 1) * 4 * sizeof(double));
    double *bp = fact + (m + 1), *c = bp + (m + 1), *r = c + (m + 1);
    vt_compact_tables(type, m, fact, bp, c);
    const double dri = 1.0 / h;
    for (int i = 1; i <= m; ++i) r[i] = line_rhs(type, m, i, dri, f);
    line_solve(m, fact, bp, c, r, out);
    free(fact);
}

/* derivative of field F along axis (0: x/i, 1: y/k, 2: z/j) over the lines
 * whose other two indices run over [lo1, hi1] x [lo2, hi2]; unknowns 1..m,
 * nodes 0..m+1 */
static void deriv_axis(const vt_params *P, const double *F, double *D, int axis, int type, int m, double h,
                       int lo1, int hi1, int lo2, int hi2)
{
    double *buf = (double *)malloc((size_t)(m + 2) * 6 * sizeof(double));
    double *f = buf, *out = f + (m + 2), *fact = out + (m + 2), *bp = fact + (m + 2), *c = bp + (m + 2),
           *r = c + (m + 2);
    vt_compact_tables(type, m, fact, bp, c);
    const double dri = 1.0 / h;
    for (int p = lo1; p <= hi1; ++p)
        for (int q = lo2; q <= hi2; ++q) {
            for (int t = 0; t <= m + 1; ++t)
                f[t] = axis == 0 ? X(F, t, p, q) : axis == 1 ? X(F, p, q, t) : X(F, p, t, q);
            for (int t = 1; t <= m; ++t) r[t] = line_rhs(type, m, t, dri, f);
            line_solve(m, fact, bp, c, r, out);
            for (int t = 1; t <= m; ++t) {
                if (axis == 0) X(D, t, p, q) = out[t];
                else if (axis == 1) X(D, p, q, t) = out[t];
                else X(D, p, t, q) = out[t];
            }
        }
    free(buf);
}

/* x lines are indexed (j, k), y lines (i, j), z lines (i, k) */
void vt_derivatives(const vt_params *P, const vt_fields *F, vt_derivs *D)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    /* A: x cells (walls half a spacing away), y faces, z faces */
    deriv_axis(P, F->A, D->d[0], 0, 1, nr, P->dr, 1, nz - 1, 1, np - 1);
    deriv_axis(P, F->A, D->d[1], 1, 0, np - 1, P->dp, 1, nr, 1, nz - 1);
    deriv_axis(P, F->A, D->d[2], 2, 0, nz - 1, P->dz, 1, nr, 1, np - 1);
    /* B: x faces, y cells (ghost planes one spacing away), z faces */
    deriv_axis(P, F->B, D->d[3], 0, 0, nr - 1, P->dr, 1, nz - 1, 1, np);
    deriv_axis(P, F->B, D->d[4], 1, 0, np, P->dp, 1, nr - 1, 1, nz - 1);
    deriv_axis(P, F->B, D->d[5], 2, 0, nz - 1, P->dz, 1, nr - 1, 1, np);
    /* C: x faces, y faces, z cells (walls half a spacing away) */
    deriv_axis(P, F->C, D->d[6], 0, 0, nr - 1, P->dr, 1, nz, 1, np - 1);
    deriv_axis(P, F->C, D->d[7], 1, 0, np - 1, P->dp, 1, nr - 1, 1, nz);
    deriv_axis(P, F->C, D->d[8], 2, 1, nz, P->dz, 1, nr - 1, 1, np - 1);
    /* th: x cells and z cells (walls half a spacing away), y cells (ghosts) */
    deriv_axis(P, F->th, D->d[9], 0, 1, nr, P->dr, 1, nz, 1, np);
    deriv_axis(P, F->th, D->d[10], 1, 0, np, P->dp, 1, nr, 1, nz);
    deriv_axis(P, F->th, D->d[11], 2, 1, nz, P->dz, 1, nr, 1, np);
}

/* ------------------------------------------------------- explicit step */

/* Avorticity (rot3dmg.f:359-465, Cartesian), at i = 1..nr, j = 1..nz-1,
 * k = 1..np-1 (the y-faces inside the side walls) */
void vt_avorticity(const vt_params *P, vt_fields *F, const vt_derivs *D)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp, dt = P->dt, Ptl = P->Ptl, Re = P->Re, Gr = P->Gr;
    const double *A = F->A, *B = F->B, *C = F->C, *th = F->th, *uel = F->uel, *vel = F->vel, *wel = F->wel;
    for (int k = 1; k <= np - 1; ++k)
        for (int j = 1; j <= nz - 1; ++j)
            for (int i = 1; i <= nr; ++i) {
                const int im = i - 1, jp = j + 1, ka = k + 1, km = k - 1;
                const double u = (X(uel, i, j, k) + X(uel, i, jp, k) + X(uel, i, j, ka) + X(uel, i, jp, ka) +
                                  X(uel, im, j, k) + X(uel, im, jp, k) + X(uel, im, j, ka) + X(uel, im, jp, ka)) *
                                 0.125;
                const double v = (X(vel, i, jp, k) + X(vel, i, j, k)) * 0.5;
                const double w = (X(wel, i, j, ka) + X(wel, i, j, k)) * 0.5;
                const double BB = (X(B, i, j, k) + X(B, im, j, k) + X(B, i, j, ka) + X(B, im, j, ka)) * 0.25;
                const double CC = (X(C, i, j, k) + X(C, i, jp, k) + X(C, im, j, k) + X(C, im, jp, k)) * 0.25;
                double Arra, cnr;
                if (i == 1) {
                    Arra = (3.2 * X(A, i - 1, j, k) + 2.0 * X(A, i + 1, j, k) - 0.2 * X(A, i + 2, j, k)) / (dr * dr);
                    cnr = -2.5 / (dr * dr);
                } else if (i == nr) {
                    Arra = (3.2 * X(A, i + 1, j, k) + 2.0 * X(A, i - 1, j, k) - 0.2 * X(A, i - 2, j, k)) / (dr * dr);
                    cnr = -2.5 / (dr * dr);
                } else {
                    Arra = (X(A, i - 1, j, k) + X(A, i + 1, j, k)) / (dr * dr);
                    cnr = -1.0 / (dr * dr);
                }
                const double uelr = ((X(uel, i, j, k) + X(uel, i, jp, k) + X(uel, i, j, ka) + X(uel, i, jp, ka)) * 0.25 -
                                     (X(uel, im, j, k) + X(uel, im, jp, k) + X(uel, im, j, ka) + X(uel, im, jp, ka)) *
                                         0.25) / dr;
                const double Azza = (X(A, i, j - 1, k) + X(A, i, j + 1, k)) / (dz * dz);
                const double uelz = ((X(uel, i, jp, k) + X(uel, i, jp, ka) + X(uel, im, jp, k) + X(uel, im, jp, ka)) *
                                         0.25 -
                                     (X(uel, i, j, k) + X(uel, i, j, ka) + X(uel, im, j, k) + X(uel, im, j, ka)) *
                                         0.25) / dz;
                const double thh = (X(th, i, j, ka) + X(th, i, jp, ka) - X(th, i, j, k) - X(th, i, jp, k)) * 0.5 / dp;
                const double Ahha = (X(A, i, j, ka) + X(A, i, j, km)) / (dp * dp);
                const double uelh = ((X(uel, i, j, ka) + X(uel, i, jp, ka) + X(uel, im, j, ka) + X(uel, im, jp, ka)) *
                                         0.25 -
                                     (X(uel, i, j, k) + X(uel, i, jp, k) + X(uel, im, j, k) + X(uel, im, jp, k)) *
                                         0.25) / dp;
                const double cnz = -1.0 / (dz * dz);
                const double cnh = -1.0 / (dp * dp);
                const double Ax = X(D->d[0], i, j, k), Ay = X(D->d[1], i, j, k), Az = X(D->d[2], i, j, k);
                const double RHS1 = Arra + Ahha + Azza + X(A, i, j, k) / (Ptl * dt);
                const double ULT = (u * Ax + v * Ay + w * Az - uelh * BB - uelz * CC) / Ptl;
                const double RHS2 = -2.0 * Re * uelz - Gr * Ptl * thh;
                const double cf1 = 1.0 / (Ptl * dt) - uelr / (2.0 * Ptl) - cnr - cnz - cnh;
                const double cf2 = uelr / (2.0 * Ptl) + cnr + cnz + cnh;
                X(F->An, i, j, k) = (cf2 * X(F->Ap, i, j, k) - ULT - RHS2 + RHS1) / cf1;
            }
}

/* Bvorticity (rot3dmg.f:495-590, Cartesian), at i = 1..nr-1, j = 1..nz-1,
 * k = 1..np */
void vt_bvorticity(const vt_params *P, vt_fields *F, const vt_derivs *D)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp, dt = P->dt, Ptl = P->Ptl, Re = P->Re, Gr = P->Gr;
    const double *A = F->A, *B = F->B, *C = F->C, *th = F->th, *uel = F->uel, *vel = F->vel, *wel = F->wel;
    for (int k = 1; k <= np; ++k)
        for (int j = 1; j <= nz - 1; ++j)
            for (int i = 1; i <= nr - 1; ++i) {
                const int ip = i + 1, jp = j + 1, km = k - 1, ka = k + 1;
                const double u = (X(uel, i, j, k) + X(uel, i, jp, k)) * 0.5;
                const double v = (X(vel, i, j, km) + X(vel, i, jp, km) + X(vel, i, j, k) + X(vel, i, jp, k) +
                                  X(vel, ip, j, km) + X(vel, ip, jp, km) + X(vel, ip, j, k) + X(vel, ip, jp, k)) *
                                 0.125;
                const double w = (X(wel, i, j, k) + X(wel, ip, j, k)) * 0.5;
                const double AA = (X(A, i, j, k) + X(A, ip, j, k) + X(A, i, j, km) + X(A, ip, j, km)) * 0.25;
                const double CC = (X(C, i, j, k) + X(C, i, jp, k) + X(C, i, j, km) + X(C, i, jp, km)) * 0.25;
                const double Brra = (X(B, i - 1, j, k) + X(B, i + 1, j, k)) / (dr * dr);
                const double velr = ((X(vel, ip, j, km) + X(vel, ip, jp, km) + X(vel, ip, j, k) + X(vel, ip, jp, k)) *
                                         0.25 -
                                     (X(vel, i, j, km) + X(vel, i, jp, km) + X(vel, i, j, k) + X(vel, i, jp, k)) *
                                         0.25) / dr;
                const double Bzza = (X(B, i, j - 1, k) + X(B, i, j + 1, k)) / (dz * dz);
                const double velz = ((X(vel, i, jp, km) + X(vel, ip, jp, km) + X(vel, i, jp, k) + X(vel, ip, jp, k)) *
                                         0.25 -
                                     (X(vel, i, j, km) + X(vel, ip, j, km) + X(vel, i, j, k) + X(vel, ip, j, k)) *
                                         0.25) / dz;
                const double thr = (X(th, ip, jp, k) + X(th, ip, j, k) - X(th, i, jp, k) - X(th, i, j, k)) * 0.5 / dr;
                const double Bhha = (X(B, i, j, ka) + X(B, i, j, km)) / (dp * dp);
                const double velh = ((X(vel, i, j, k) + X(vel, ip, j, k) + X(vel, i, jp, k) + X(vel, ip, jp, k)) * 0.25 -
                                     (X(vel, i, j, km) + X(vel, ip, j, km) + X(vel, i, jp, km) + X(vel, ip, jp, km)) *
                                         0.25) / dp;
                const double cnr = -1.0 / (dr * dr);
                const double cnz = -1.0 / (dz * dz);
                const double cnh = -1.0 / (dp * dp);
                const double Bx = X(D->d[3], i, j, k), By = X(D->d[4], i, j, k), Bz = X(D->d[5], i, j, k);
                const double RHS1 = Brra + Bhha + Bzza + X(B, i, j, k) / (Ptl * dt);
                const double ULT = (u * Bx + v * By + w * Bz - velr * AA - velz * CC) / Ptl;
                const double RHS2 = -2.0 * Re * velz + Gr * Ptl * thr;
                const double cf1 = 1.0 / (Ptl * dt) - velh / (2.0 * Ptl) - cnr - cnz - cnh;
                const double cf2 = velh / (2.0 * Ptl) + cnr + cnz + cnh;
                X(F->Bn, i, j, k) = (cf2 * X(F->Bp, i, j, k) - ULT - RHS2 + RHS1) / cf1;
            }
}

/* Cvorticity (rot3dmg.f:619-725, Cartesian), at i = 1..nr-1, j = 1..nz,
 * k = 1..np-1 */
void vt_cvorticity(const vt_params *P, vt_fields *F, const vt_derivs *D)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp, dt = P->dt, Ptl = P->Ptl, Re = P->Re;
    const double *A = F->A, *B = F->B, *C = F->C, *uel = F->uel, *vel = F->vel, *wel = F->wel;
    for (int k = 1; k <= np - 1; ++k)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nr - 1; ++i) {
                const int ip = i + 1, jm = j - 1, km = k - 1, ka = k + 1;
                const double u = (X(uel, i, j, ka) + X(uel, i, j, k)) * 0.5;
                const double v = (X(vel, i, j, k) + X(vel, ip, j, k)) * 0.5;
                const double w = (X(wel, i, j, k) + X(wel, ip, j, k) + X(wel, i, j, ka) + X(wel, ip, j, ka) +
                                  X(wel, i, jm, k) + X(wel, ip, jm, k) + X(wel, i, jm, ka) + X(wel, ip, jm, ka)) *
                                 0.125;
                const double AA = (X(A, i, j, k) + X(A, i, jm, k) + X(A, ip, j, k) + X(A, ip, jm, k)) * 0.25;
                const double BB = (X(B, i, j, k) + X(B, i, jm, k) + X(B, i, j, ka) + X(B, i, jm, ka)) * 0.25;
                const double Crra = (X(C, i - 1, j, k) + X(C, i + 1, j, k)) / (dr * dr);
                const double welr = ((X(wel, ip, j, k) + X(wel, ip, j, ka) + X(wel, ip, jm, k) + X(wel, ip, jm, ka)) *
                                         0.25 -
                                     (X(wel, i, j, k) + X(wel, i, j, ka) + X(wel, i, jm, k) + X(wel, i, jm, ka)) *
                                         0.25) / dr;
                double Czza, cnz;
                if (j == 1) {
                    Czza = (3.2 * X(C, i, j - 1, k) + 2.0 * X(C, i, j + 1, k) - 0.2 * X(C, i, j + 2, k)) / (dz * dz);
                    cnz = -2.5 / (dz * dz);
                } else if (j == nz) {
                    Czza = (3.2 * X(C, i, j + 1, k) + 2.0 * X(C, i, j - 1, k) - 0.2 * X(C, i, j - 2, k)) / (dz * dz);
                    cnz = -2.5 / (dz * dz);
                } else {
                    Czza = (X(C, i, j - 1, k) + X(C, i, j + 1, k)) / (dz * dz);
                    cnz = -1.0 / (dz * dz);
                }
                const double welz = ((X(wel, i, j, k) + X(wel, ip, j, k) + X(wel, i, j, ka) + X(wel, ip, j, ka)) * 0.25 -
                                     (X(wel, i, jm, k) + X(wel, ip, jm, k) + X(wel, i, jm, ka) + X(wel, ip, jm, ka)) *
                                         0.25) / dz;
                const double Chha = (X(C, i, j, ka) + X(C, i, j, km)) / (dp * dp);
                const double welh = ((X(wel, i, j, ka) + X(wel, ip, j, ka) + X(wel, i, jm, ka) + X(wel, ip, jm, ka)) *
                                         0.25 -
                                     (X(wel, i, j, k) + X(wel, ip, j, k) + X(wel, i, jm, k) + X(wel, ip, jm, k)) *
                                         0.25) / dp;
                const double cnr = -1.0 / (dr * dr);
                const double cnh = -1.0 / (dp * dp);
                const double Cx = X(D->d[6], i, j, k), Cy = X(D->d[7], i, j, k), Cz = X(D->d[8], i, j, k);
                const double RHS1 = Crra + Chha + Czza + X(C, i, j, k) / (Ptl * dt);
                const double ULT = (u * Cx + v * Cy + w * Cz - welr * AA - welh * BB) / Ptl;
                const double RHS2 = -2.0 * Re * welz;
                const double cf1 = 1.0 / (Ptl * dt) - welz / (2.0 * Ptl) - cnr - cnz - cnh;
                const double cf2 = welz / (2.0 * Ptl) + cnr + cnz + cnh;
                X(F->Cn, i, j, k) = (cf2 * X(F->Cp, i, j, k) - ULT - RHS2 + RHS1) / cf1;
            }
}

/* temperature (rot3dmg.f:754-871, Cartesian), at the interior cells
 * i = 1..nr, j = 1..nz, k = 1..np; the walls and ghosts are set by
 * vt_update */
void vt_temperature(const vt_params *P, vt_fields *F, const vt_derivs *D)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp, dt = P->dt;
    const double *th = F->th, *uel = F->uel, *vel = F->vel, *wel = F->wel;
    for (int k = 1; k <= np; ++k)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nr; ++i) {
                const int im = i - 1, jm = j - 1, km = k - 1, ka = k + 1;
                const double u = (X(uel, i, j, k) + X(uel, im, j, k)) * 0.5;
                const double v = (X(vel, i, j, k) + X(vel, i, j, km)) * 0.5;
                const double w = (X(wel, i, j, k) + X(wel, i, jm, k)) * 0.5;
                double thrra, cnr, thzza, cnz;
                if (i == 1) {
                    thrra = (3.2 * X(th, i - 1, j, k) + 2.0 * X(th, i + 1, j, k) - 0.2 * X(th, i + 2, j, k)) / (dr * dr);
                    cnr = -2.5 / (dr * dr);
                } else if (i == nr) {
                    thrra = (3.2 * X(th, i + 1, j, k) + 2.0 * X(th, i - 1, j, k) - 0.2 * X(th, i - 2, j, k)) / (dr * dr);
                    cnr = -2.5 / (dr * dr);
                } else {
                    thrra = (X(th, i - 1, j, k) + X(th, i + 1, j, k)) / (dr * dr);
                    cnr = -1.0 / (dr * dr);
                }
                if (j == 1) {
                    thzza = (3.2 * X(th, i, j - 1, k) + 2.0 * X(th, i, j + 1, k) - 0.2 * X(th, i, j + 2, k)) / (dz * dz);
                    cnz = -2.5 / (dz * dz);
                } else if (j == nz) {
                    thzza = (3.2 * X(th, i, j + 1, k) + 2.0 * X(th, i, j - 1, k) - 0.2 * X(th, i, j - 2, k)) / (dz * dz);
                    cnz = -2.5 / (dz * dz);
                } else {
                    thzza = (X(th, i, j - 1, k) + X(th, i, j + 1, k)) / (dz * dz);
                    cnz = -1.0 / (dz * dz);
                }
                const double thhha = (X(th, i, j, ka) + X(th, i, j, km)) / (dp * dp);
                const double cnp = -1.0 / (dp * dp);
                const double Tx = X(D->d[9], i, j, k), Ty = X(D->d[10], i, j, k), Tz = X(D->d[11], i, j, k);
                const double RHS1 = thrra + thhha + thzza + X(th, i, j, k) / dt;
                const double ULT = u * Tx + v * Ty + w * Tz;
                const double cf1 = 1.0 / dt - cnr - cnz - cnp;
                const double cf2 = cnr + cnz + cnp;
                X(F->thn, i, j, k) = (cf2 * X(F->thp, i, j, k) - ULT + RHS1) / cf1;
            }
}

/* heat(trans1,trans2), velvor2dhorconvtank.f:1574-1622 with its wall
 * rows at the bottom (j = 0..4) and the surface (j = nz+1..nz-3), each x
 * column summed over k = np/2..np-1 in the routine's order, then the
 * columns summed in order of i with the weight dr */
void vt_heat(const vt_params *P, const vt_fields *F, double out[2])
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dp = P->dp, dzi = 1.0 / P->dz;
    const double *th = F->th;
    double s1 = 0.0, s2 = 0.0;
    for (int i = 1; i <= nr; ++i) {
        double trans1 = 0.0, trans2 = 0.0;
        for (int k = np / 2; k <= np - 1; ++k) {
            const int ka = k + 1;
            trans1 = trans1 +
                     0.5 * dp *
                         (-3.35238095 * X(th, i, 0, k) + 4.375 * X(th, i, 1, k) - 1.45833333 * X(th, i, 2, k) +
                          0.525 * X(th, i, 3, k) - 0.08928571 * X(th, i, 4, k)) *
                         dzi +
                     0.5 * dp *
                         (-3.35238095 * X(th, i, 0, ka) + 4.375 * X(th, i, 1, ka) - 1.45833333 * X(th, i, 2, ka) +
                          0.525 * X(th, i, 3, ka) - 0.08928571 * X(th, i, 4, ka)) *
                         dzi;
            trans2 = trans2 +
                     0.5 * dp *
                         (-3.35238095 * X(th, i, nz + 1, k) + 4.375 * X(th, i, nz, k) -
                          1.45833333 * X(th, i, nz - 1, k) + 0.525 * X(th, i, nz - 2, k) -
                          0.08928571 * X(th, i, nz - 3, k)) *
                         dzi +
                     0.5 * dp *
                         (-3.35238095 * X(th, i, nz + 1, ka) + 4.375 * X(th, i, nz, ka) -
                          1.45833333 * X(th, i, nz - 1, ka) + 0.525 * X(th, i, nz - 2, ka) -
                          0.08928571 * X(th, i, nz - 3, ka)) *
                         dzi;
        }
        s1 = s1 + trans1 * P->dr;
        s2 = s2 + trans2 * P->dr;
    }
    out[0] = s1;
    out[1] = s2;
}

/* rhsw (rot3dmg.f:1026-1048, Cartesian): i = 1..nr, j = 1..nz-1, k = 1..np */
void vt_rhsw(const vt_params *P, vt_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dp = P->dp;
    const double *An = F->An, *Bn = F->Bn;
    for (int k = 1; k <= np; ++k)
        for (int j = 1; j <= nz - 1; ++j)
            for (int i = 1; i <= nr; ++i) {
                const double Ah = (X(An, i, j, k) - X(An, i, j, k - 1)) / dp;
                const double Br = (X(Bn, i, j, k) - X(Bn, i - 1, j, k)) / dr;
                double v;
                if (i == 1) v = Ah - X(Bn, 1, j, k) / dr;
                else if (i == nr) v = Ah + X(Bn, nr - 1, j, k) / dr;
                else v = Ah - Br;
                X(F->wrh, i, j, k) = v;
            }
}

/* rhsv (rot3dmg.f:1066-1093, Cartesian): i = 1..nr, j = 1..nz, k = 1..np-1 */
void vt_rhsv(const vt_params *P, vt_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz;
    const double *An = F->An, *Cn = F->Cn;
    for (int k = 1; k <= np - 1; ++k)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nr; ++i) {
                const double Az = (X(An, i, j, k) - X(An, i, j - 1, k)) / dz;
                const double Cr = (X(Cn, i, j, k) - X(Cn, i - 1, j, k)) / dr;
                double v;
                if (i == 1) v = X(Cn, 1, j, k) / dr;
                else if (i == nr) v = -X(Cn, nr - 1, j, k) / dr;
                else v = Cr;
                if (j == 1) v = v - X(An, i, 1, k) / dz;
                else if (j == nz) v = v + X(An, i, nz - 1, k) / dz;
                else v = v - Az;
                X(F->vrh, i, j, k) = v;
            }
}

/* rhsu (rot3dmg.f:1111-1134, Cartesian): i = 1..nr-1, j = 1..nz, k = 1..np */
void vt_rhsu(const vt_params *P, vt_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dz = P->dz, dp = P->dp;
    const double *Bn = F->Bn, *Cn = F->Cn;
    for (int k = 1; k <= np; ++k)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nr - 1; ++i) {
                const double Bz = (X(Bn, i, j, k) - X(Bn, i, j - 1, k)) / dz;
                const double Ch = (X(Cn, i, j, k) - X(Cn, i, j, k - 1)) / dp;
                double v;
                if (j == 1) v = X(Bn, i, 1, k) / dz - Ch;
                else if (j == nz) v = -X(Bn, i, nz - 1, k) / dz - Ch;
                else v = Bz - Ch;
                X(F->urh, i, j, k) = v;
            }
}

/* update: rot3dmg.f:980-1015 (wall vorticity at the x and z walls, the
 * rotation of the time levels) and velvor2dhorconvtank.f:1720-1745 (zero
 * vorticity at the side walls, the insulated walls, the linear surface
 * temperature) */
void vt_update(const vt_params *P, vt_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz;
    const double *uel = F->uel, *vel = F->vel, *wel = F->wel;
    double *An = F->An, *Bn = F->Bn, *Cn = F->Cn, *thn = F->thn;
    /* z walls (rot3dmg.f:980-987) */
    for (int k = 1; k <= np - 1; ++k)
        for (int i = 1; i <= nr; ++i) {
            X(An, i, 0, k) = -(3.0 * X(vel, i, 1, k) - X(vel, i, 2, k) / 3.0) / dz;
            X(An, i, nz, k) = -(X(vel, i, nz - 1, k) / 3.0 - 3.0 * X(vel, i, nz, k)) / dz;
        }
    for (int k = 1; k <= np; ++k)
        for (int i = 1; i <= nr - 1; ++i) {
            X(Bn, i, 0, k) = (3.0 * X(uel, i, 1, k) - X(uel, i, 2, k) / 3.0) / dz;
            X(Bn, i, nz, k) = (-3.0 * X(uel, i, nz, k) + X(uel, i, nz - 1, k) / 3.0) / dz;
        }
    /* x walls (rot3dmg.f:989-996) */
    for (int k = 1; k <= np; ++k)
        for (int j = 1; j <= nz - 1; ++j) {
            X(Bn, 0, j, k) = -(3.0 * X(wel, 1, j, k) - X(wel, 2, j, k) / 3.0) / dr;
            X(Bn, nr, j, k) = -(X(wel, nr - 1, j, k) / 3.0 - 3.0 * X(wel, nr, j, k)) / dr;
        }
    for (int k = 1; k <= np - 1; ++k)
        for (int j = 1; j <= nz; ++j) {
            X(Cn, 0, j, k) = (3.0 * X(vel, 1, j, k) - X(vel, 2, j, k) / 3.0) / dr;
            X(Cn, nr, j, k) = (X(vel, nr - 1, j, k) / 3.0 - 3.0 * X(vel, nr, j, k)) / dr;
        }
    /* side walls, stress free: A = C = 0 (velvor2dhorconvtank.f:1720-1727
     * for C), B mirrored into the ghost planes */
    for (int j = 0; j <= nz + 1; ++j)
        for (int i = 0; i <= nr + 1; ++i) {
            X(An, i, j, 0) = 0.0;
            X(An, i, j, np) = 0.0;
            X(Cn, i, j, 0) = 0.0;
            X(Cn, i, j, np) = 0.0;
            X(Bn, i, j, 0) = X(Bn, i, j, 1);
            X(Bn, i, j, np + 1) = X(Bn, i, j, np);
        }
    /* temperature: insulated bottom and the linear surface
     * (velvor2dhorconvtank.f:1733-1737), insulated x walls (the same
     * insulated row), then the insulated side walls (1741-1744) */
    for (int k = 1; k <= np; ++k)
        for (int i = 1; i <= nr; ++i) {
            X(thn, i, 0, k) = -30.0 / 92.0 *
                              (-3.75 * X(thn, i, 1, k) + 5.0 / 6.0 * X(thn, i, 2, k) - 3.0 / 20.0 * X(thn, i, 3, k));
            X(thn, i, nz + 1, k) = (double)k / (double)np * 1.0;
        }
    for (int k = 1; k <= np; ++k)
        for (int j = 0; j <= nz + 1; ++j) {
            X(thn, 0, j, k) = -30.0 / 92.0 *
                              (-3.75 * X(thn, 1, j, k) + 5.0 / 6.0 * X(thn, 2, j, k) - 3.0 / 20.0 * X(thn, 3, j, k));
            X(thn, nr + 1, j, k) =
                -30.0 / 92.0 *
                (-3.75 * X(thn, nr, j, k) + 5.0 / 6.0 * X(thn, nr - 1, j, k) - 3.0 / 20.0 * X(thn, nr - 2, j, k));
        }
    for (int j = 0; j <= nz + 1; ++j)
        for (int i = 0; i <= nr + 1; ++i) {
            X(thn, i, j, 0) = 2.0 / 1.5 * X(thn, i, j, 1) - 1.0 / 3.0 * X(thn, i, j, 2);
            X(thn, i, j, np + 1) = 2.0 / 1.5 * X(thn, i, j, np) - 1.0 / 3.0 * X(thn, i, j, np - 1);
        }
    /* rot3dmg.f:1003-1015 */
    const size_t n = (size_t)P->ldi * P->ldj * (size_t)(np + 2);
    for (size_t q = 0; q < n; ++q) {
        F->Ap[q] = F->A[q];
        F->A[q] = F->An[q];
        F->Bp[q] = F->B[q];
        F->B[q] = F->Bn[q];
        F->Cp[q] = F->C[q];
        F->C[q] = F->Cn[q];
        F->thp[q] = F->th[q];
        F->th[q] = F->thn[q];
    }
}

/* ------------------------------------------------------------- Poisson */

void vt_ranges(const vt_params *P, int which, int *i1, int *j1, int *k1)
{
    *i1 = (which == 3) ? P->nr - 1 : P->nr;
    *j1 = (which == 1) ? P->nz - 1 : P->nz;
    *k1 = (which == 2) ? P->np - 1 : P->np;
}

/* one direction's rows at row t of n (f(s) reads row s): kind 0 Dirichlet
 * (1,-2,1) with zero walls; kind 1 the no-slip closure (-2, 5/3, -0.2) at
 * the two end rows (the rows of rot3dmg.f's residuals next to a wall half a
 * spacing away); kind 2 Neumann (mirrored ghost: end rows (-1, 1)) */
#define ROW_OP(kind, n, t, f, h2, out)                                                \
    do {                                                                              \
        if ((kind) == 1 && (t) == 1) (out) = (-2.0 * f(1) + 5.0 / 3.0 * f(2) - 0.2 * f(3)) / (h2); \
        else if ((kind) == 1 && (t) == (n))                                          \
            (out) = (-2.0 * f(n) + 5.0 / 3.0 * f((n) - 1) - 0.2 * f((n) - 2)) / (h2);  \
        else if ((kind) == 2 && (t) == 1) (out) = (f(2) - f(1)) / (h2);              \
        else if ((kind) == 2 && (t) == (n)) (out) = (f((n) - 1) - f(n)) / (h2);      \
        else {                                                                        \
            const double lo_ = ((t) == 1) ? 0.0 : f((t) - 1), hi_ = ((t) == (n)) ? 0.0 : f((t) + 1); \
            (out) = (lo_ - 2.0 * f(t) + hi_) / (h2);                                  \
        }                                                                             \
    } while (0)

/* kinds per direction (x, z, y) of each operator */
static void kinds(int which, int *kx, int *kz, int *ky)
{
    *kx = (which == 3) ? 0 : 1;
    *kz = (which == 1) ? 0 : 1;
    *ky = (which == 2) ? 0 : 2;
}

void vt_apply(const vt_params *P, int which, const double *psi, double *out)
{
    int i1, j1, k1, kx, kz, ky;
    vt_ranges(P, which, &i1, &j1, &k1);
    kinds(which, &kx, &kz, &ky);
    const double dr2 = P->dr * P->dr, dz2 = P->dz * P->dz, dp2 = P->dp * P->dp;
    for (int k = 1; k <= k1; ++k)
        for (int j = 1; j <= j1; ++j)
            for (int i = 1; i <= i1; ++i) {
                double lx, lz, ly;
#define FX(s) X(psi, (s), j, k)
#define FZ(s) X(psi, i, (s), k)
#define FY(s) X(psi, i, j, (s))
                ROW_OP(kx, i1, i, FX, dr2, lx);
                ROW_OP(kz, j1, j, FZ, dz2, lz);
                ROW_OP(ky, k1, k, FY, dp2, ly);
#undef FX
#undef FZ
#undef FY
                X(out, i, j, k) = lx + lz + ly;
            }
}

double vt_residual(const vt_params *P, int which, const double *psi, const double *rhs, double *res)
{
    int i1, j1, k1;
    vt_ranges(P, which, &i1, &j1, &k1);
    const size_t n = (size_t)P->ldi * P->ldj * (size_t)(P->np + 2);
    double *lp = (double *)calloc(n, sizeof(double));
    vt_apply(P, which, psi, lp);
    if (res) memset(res, 0, n * sizeof(double));
    double m = 0.0;
    for (int k = 1; k <= k1; ++k)
        for (int j = 1; j <= j1; ++j)
            for (int i = 1; i <= i1; ++i) {
                const double r = X(rhs, i, j, k) - X(lp, i, j, k);
                if (res) X(res, i, j, k) = r;
                if (fabs(r) > m) m = fabs(r);
            }
    free(lp);
    return m;
}

void vt_psi_bounds(const vt_params *P, int which, double *psi)
{
    int i1, j1, k1;
    vt_ranges(P, which, &i1, &j1, &k1);
    const int np = P->np;
    for (int k = 0; k <= np + 1; ++k)
        for (int j = 0; j <= P->nz + 1; ++j)
            for (int i = 0; i <= P->nr + 1; ++i) {
                const int in = i >= 1 && i <= i1 && j >= 1 && j <= j1;
                if (!in) X(psi, i, j, k) = 0.0;
            }
    if (which == 2) {
        for (int j = 0; j <= P->nz + 1; ++j)
            for (int i = 0; i <= P->nr + 1; ++i) {
                X(psi, i, j, 0) = 0.0;
                X(psi, i, j, np) = 0.0;
                X(psi, i, j, np + 1) = 0.0;
            }
    } else {
        for (int j = 0; j <= P->nz + 1; ++j)
            for (int i = 0; i <= P->nr + 1; ++i) {
                X(psi, i, j, 0) = X(psi, i, j, 1);
                X(psi, i, j, np + 1) = X(psi, i, j, np);
            }
    }
}

double vt_energy(const vt_params *P, const vt_fields *F)
{
    double s = 0.0;
    for (int which = 1; which <= 3; ++which) {
        int i1, j1, k1;
        vt_ranges(P, which, &i1, &j1, &k1);
        const double *a = (which == 1) ? F->wel : (which == 2) ? F->vel : F->uel;
        for (int k = 1; k <= k1; ++k)
            for (int j = 1; j <= j1; ++j)
                for (int i = 1; i <= i1; ++i) s += X(a, i, j, k) * X(a, i, j, k);
    }
    return 0.5 * P->dr * P->dz * P->dp * s;
}
