/* vv3_oracle.c -- CPU oracle (test infrastructure only) for the 3D time step
 * of rot3dmg.f (3D rotating annulus, periodic phi) and of its stationary
 * horizontal-annulus variant rot3dmghor.f (P->model = 1).  Each routine restates
 * the reference subroutine's loops and expressions in the reference's
 * operation order (compiled with -ffp-contract=off, it reproduces the
 * reference bit for bit); the file:line of each follows its name.  The exact
 * Poisson solve is this oracle's own (a DFT in phi and a dense LU per mode)
 * and is checked against the reference's residual operators in the tests.
 */
#include "vv3_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IX(i, j, k) (((size_t)(k) * P->ldj + (size_t)(j)) * P->ldi + (size_t)(i))
#define X(a, i, j, k) ((a)[IX(i, j, k)])

/* rot3dmghor.f's azimuth of plane k, theta=6.283185307d0/np*k
 * (rot3dmghor.f:374, 515, 644) */
static double hor_theta(const v3_params *P, int k) { return 6.283185307 / P->np * k; }

/* Avorticity, rot3dmg.f:341-468 (hybrid upwind 397-406, 421-430, 444-453) */
void v3_avorticity(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp, dt = P->dt;
    const double Ptl = P->Ptl, Ra = P->Ra, Re = P->Re, Gr = P->Gr;
    const double *A = F->A, *B = F->B, *C = F->C, *Ap = F->Ap, *th = F->th, *uel = F->uel, *vel = F->vel,
                 *wel = F->wel;
    for (int k = 0; k <= np; ++k)
        for (int j = 1; j <= nz - 1; ++j)
            for (int i = 1; i <= nr; ++i) {
                const int im = i - 1, jp = j + 1;
                const int km = (k == 0) ? np - 1 : k - 1;
                const int ka = (k == np) ? 1 : k + 1;
                const double r = P->r0 + ((double)i - 0.5) * dr;
                const double u = (X(uel, i, j, k) + X(uel, i, jp, k) + X(uel, i, j, ka) + X(uel, i, jp, ka) +
                                  X(uel, im, j, k) + X(uel, im, jp, k) + X(uel, im, j, ka) + X(uel, im, jp, ka)) *
                                 0.125;
                const double v = (X(vel, i, jp, k) + X(vel, i, j, k)) * 0.5;
                const double w = (X(wel, i, j, ka) + X(wel, i, j, k)) * 0.5;
                const double tem = (X(th, i, j, k) + X(th, i, j, ka) + X(th, i, jp, k) + X(th, i, jp, ka)) * 0.25;
                const double BB = (X(B, i, j, k) + X(B, im, j, k) + X(B, i, j, ka) + X(B, im, j, ka)) * 0.25;
                const double CC = (X(C, i, j, k) + X(C, i, jp, k) + X(C, im, j, k) + X(C, im, jp, k)) * 0.25;
                double Arra, Ara, Ar, cnr;
                if (i == 1) {
                    Arra = (3.2 * X(A, i - 1, j, k) + 2.0 * X(A, i + 1, j, k) - 0.2 * X(A, i + 2, j, k)) / (dr * dr);
                    Ara = (-4.0 * X(A, i - 1, j, k) / 3.0 + X(A, i + 1, j, k) / 3.0) / dr;
                    Ar = (-4.0 * X(A, i - 1, j, k) / 3.0 + X(A, i, j, k) + X(A, i + 1, j, k) / 3.0) / dr;
                    cnr = -2.5 / (dr * dr) + 0.5 / dr / r;
                } else if (i == nr) {
                    Arra = (3.2 * X(A, i + 1, j, k) + 2.0 * X(A, i - 1, j, k) - 0.2 * X(A, i - 2, j, k)) / (dr * dr);
                    Ara = (-X(A, i - 1, j, k) / 3.0 + 4.0 * X(A, i + 1, j, k) / 3.0) / dr;
                    Ar = (-4.0 * X(A, i - 1, j, k) / 3.0 + X(A, i, j, k) + X(A, i + 1, j, k) / 3.0) / dr;
                    cnr = -2.5 / (dr * dr) - 0.5 / dr / r;
                } else {
                    Arra = (X(A, i - 1, j, k) + X(A, i + 1, j, k)) / (dr * dr);
                    Ara = (X(A, i + 1, j, k) - X(A, i - 1, j, k)) * 0.5 / dr;
                    Ar = Ara;
                    cnr = -1.0 / (dr * dr);
                    const double Pe = fabs(u * dr / Ptl);
                    if (Pe > 2) {
                        if (u > 0) Ar = (X(A, i, j, k) - X(A, i - 1, j, k)) / dr;
                        else if (u < 0) Ar = (X(A, i + 1, j, k) - X(A, i, j, k)) / dr;
                    }
                }
                const double uelr = ((X(uel, i, j, k) + X(uel, i, jp, k) + X(uel, i, j, ka) + X(uel, i, jp, ka)) * 0.25 -
                                     (X(uel, im, j, k) + X(uel, im, jp, k) + X(uel, im, j, ka) + X(uel, im, jp, ka)) *
                                         0.25) /
                                    dr;
                const double Azza = (X(A, i, j - 1, k) + X(A, i, j + 1, k)) / (dz * dz);
                double Az = (X(A, i, j + 1, k) - X(A, i, j - 1, k)) * 0.5 / dz;
                {
                    const double Pe = fabs(w * dz / Ptl);
                    if (Pe > 2) {
                        if (w > 0) Az = (X(A, i, j, k) - X(A, i, j - 1, k)) / dz;
                        else if (w < 0) Az = (X(A, i, j + 1, k) - X(A, i, j, k)) / dz;
                    }
                }
                const double uelz = ((X(uel, i, jp, k) + X(uel, i, jp, ka) + X(uel, im, jp, k) + X(uel, im, jp, ka)) * 0.25 -
                                     (X(uel, i, j, k) + X(uel, i, j, ka) + X(uel, im, j, k) + X(uel, im, j, ka)) * 0.25) /
                                    dz;
                const double thz = (X(th, i, jp, k) + X(th, i, jp, ka) - X(th, i, j, k) - X(th, i, j, ka)) * 0.5 / dz;
                const double thh = (X(th, i, j, ka) + X(th, i, jp, ka) - X(th, i, j, k) - X(th, i, jp, k)) * 0.5 / dp;
                const double Ahha = (X(A, i, j, ka) + X(A, i, j, km)) / (dp * dp);
                double Ah = (X(A, i, j, ka) - X(A, i, j, km)) * 0.5 / dp;
                {
                    const double Pe = fabs(v * dp * r / Ptl);
                    if (Pe > 2) {
                        if (v > 0) Ah = (X(A, i, j, k) - X(A, i, j, km)) / dp;
                        else if (v < 0) Ah = (X(A, i, j, ka) - X(A, i, j, k)) / dp;
                    }
                }
                const double Bh = (X(B, i, j, ka) + X(B, im, j, ka) - X(B, i, j, k) - X(B, im, j, k)) * 0.5 / dp;
                const double uelh = ((X(uel, i, j, ka) + X(uel, i, jp, ka) + X(uel, im, j, ka) + X(uel, im, jp, ka)) * 0.25 -
                                     (X(uel, i, j, k) + X(uel, i, jp, k) + X(uel, im, j, k) + X(uel, im, jp, k)) * 0.25) /
                                    dp;
                const double cnz = -1.0 / (dz * dz);
                const double cnh = -1.0 / (dp * dp) / (r * r);
                const double RHS1 = Arra + Ara / r + Ahha / (r * r) + Azza - 2.0 * Bh / (r * r) + X(A, i, j, k) / (Ptl * dt);
                const double ULT = (u * Ar + v * Ah / r + w * Az - uelh * BB / r - uelz * CC) / Ptl;
                /* rot3dmghor.f:479 replaces the Coriolis and centrifugal terms */
                const double RHS2 = (P->model == 1)
                                        ? sin(hor_theta(P, k)) * Gr * Ptl * thz
                                        : -2.0 * Re * uelz + 2.0 * Ra * (tem * uelz + thz * u) / Re / Ptl - Gr * Ptl / r * thh;
                const double cf1 = 1.0 / (Ptl * dt) - uelr / (2.0 * Ptl) + 1.0 / (2.0 * (r * r)) - cnr - cnz - cnh;
                const double cf2 = uelr / (2.0 * Ptl) - 1.0 / (2.0 * (r * r)) + cnr + cnz + cnh;
                X(F->An, i, j, k) = (cf2 * X(Ap, i, j, k) - ULT - RHS2 + RHS1) / cf1;
            }
}

/* Bvorticity, rot3dmg.f:477-593 (hybrid upwind 530-539, 547-556, 564-573) */
void v3_bvorticity(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp, dt = P->dt;
    const double Ptl = P->Ptl, Ra = P->Ra, Re = P->Re, Gr = P->Gr;
    const double *A = F->A, *B = F->B, *C = F->C, *Bp = F->Bp, *th = F->th, *uel = F->uel, *vel = F->vel,
                 *wel = F->wel;
    for (int k = 0; k <= np; ++k)
        for (int j = 1; j <= nz - 1; ++j)
            for (int i = 1; i <= nr - 1; ++i) {
                const int ip = i + 1, jp = j + 1;
                const int km = (k == 0) ? np - 1 : k - 1;
                const int ka = (k == np) ? 1 : k + 1;
                const double r = P->r0 + (double)i * dr;
                const double u = (X(uel, i, j, k) + X(uel, i, jp, k)) * 0.5;
                const double v = (X(vel, i, j, km) + X(vel, i, jp, km) + X(vel, i, j, k) + X(vel, i, jp, k) +
                                  X(vel, ip, j, km) + X(vel, ip, jp, km) + X(vel, ip, j, k) + X(vel, ip, jp, k)) *
                                 0.125;
                const double w = (X(wel, i, j, k) + X(wel, ip, j, k)) * 0.5;
                const double tem = (X(th, i, jp, k) + X(th, ip, jp, k) + X(th, i, j, k) + X(th, ip, j, k)) * 0.25;
                const double AA = (X(A, i, j, k) + X(A, ip, j, k) + X(A, i, j, km) + X(A, ip, j, km)) * 0.25;
                const double CC = (X(C, i, j, k) + X(C, i, jp, k) + X(C, i, j, km) + X(C, i, jp, km)) * 0.25;
                const double Brra = (X(B, i - 1, j, k) + X(B, i + 1, j, k)) / (dr * dr);
                const double Bra = (X(B, i + 1, j, k) - X(B, i - 1, j, k)) * 0.5 / dr;
                double Br = Bra;
                {
                    const double Pe = fabs(u * dr / Ptl);
                    if (Pe > 2) {
                        if (u > 0) Br = (X(B, i, j, k) - X(B, i - 1, j, k)) / dr;
                        else if (u < 0) Br = (X(B, i + 1, j, k) - X(B, i, j, k)) / dr;
                    }
                }
                const double velr = ((X(vel, ip, j, km) + X(vel, ip, jp, km) + X(vel, ip, j, k) + X(vel, ip, jp, k)) * 0.25 -
                                     (X(vel, i, j, km) + X(vel, i, jp, km) + X(vel, i, j, k) + X(vel, i, jp, k)) * 0.25) /
                                    dr;
                const double Bzza = (X(B, i, j - 1, k) + X(B, i, j + 1, k)) / (dz * dz);
                double Bz = (X(B, i, j + 1, k) - X(B, i, j - 1, k)) * 0.5 / dz;
                {
                    const double Pe = fabs(w * dz / Ptl);
                    if (Pe > 2) {
                        if (w > 0) Bz = (X(B, i, j, k) - X(B, i, j - 1, k)) / dz;
                        else if (w < 0) Bz = (X(B, i, j + 1, k) - X(B, i, j, k)) / dz;
                    }
                }
                const double velz = ((X(vel, i, jp, km) + X(vel, ip, jp, km) + X(vel, i, jp, k) + X(vel, ip, jp, k)) * 0.25 -
                                     (X(vel, i, j, km) + X(vel, ip, j, km) + X(vel, i, j, k) + X(vel, ip, j, k)) * 0.25) /
                                    dz;
                const double thz = (X(th, i, jp, k) + X(th, ip, jp, k) - X(th, i, j, k) - X(th, ip, j, k)) * 0.5 / dz;
                const double thr = (X(th, ip, jp, k) + X(th, ip, j, k) - X(th, i, jp, k) - X(th, i, j, k)) * 0.5 / dr;
                const double Bhha = (X(B, i, j, ka) + X(B, i, j, km)) / (dp * dp);
                double Bh = (X(B, i, j, ka) - X(B, i, j, km)) * 0.5 / dp;
                {
                    const double Pe = fabs(v * dp * r / Ptl);
                    if (Pe > 2) {
                        if (v > 0) Bh = (X(B, i, j, k) - X(B, i, j, km)) / dp;
                        else if (v < 0) Bh = (X(B, i, j, ka) - X(B, i, j, k)) / dp;
                    }
                }
                const double Ah = (X(A, i, j, k) + X(A, ip, j, k) - X(A, i, j, km) - X(A, ip, j, km)) * 0.5 / dp;
                const double velh = ((X(vel, i, j, k) + X(vel, ip, j, k) + X(vel, i, jp, k) + X(vel, ip, jp, k)) * 0.25 -
                                     (X(vel, i, j, km) + X(vel, ip, j, km) + X(vel, i, jp, km) + X(vel, ip, jp, km)) * 0.25) /
                                    dp;
                const double cnr = -1.0 / (dr * dr);
                const double cnz = -1.0 / (dz * dz);
                const double cnh = -1.0 / (dp * dp) / (r * r);
                const double RHS1 = Brra + Bra / r + Bhha / (r * r) + Bzza + 2.0 * Ah / (r * r) + X(B, i, j, k) / (Ptl * dt);
                const double ULT = (u * Br + v * Bh / r + w * Bz - (velr - v / r) * AA - velz * CC) / Ptl;
                /* rot3dmghor.f:606 */
                const double RHS2 = (P->model == 1) ? cos(hor_theta(P, k)) * Gr * Ptl * thz
                                                    : -2.0 * Re * velz + r * Ra * thz +
                                                          2.0 * Ra * (tem * velz + thz * v) / Re / Ptl + Gr * Ptl * thr;
                const double cf1 = 1.0 / (Ptl * dt) - (u / r + velh / r) / (2.0 * Ptl) + 1.0 / (2.0 * (r * r)) - cnr - cnz - cnh;
                const double cf2 = (u / r + velh / r) / (2.0 * Ptl) - 1.0 / (2.0 * (r * r)) + cnr + cnz + cnh;
                X(F->Bn, i, j, k) = (cf2 * X(Bp, i, j, k) - ULT - RHS2 + RHS1) / cf1;
            }
}

/* Cvorticity, rot3dmg.f:601-728 (hybrid upwind 648-657, 684-693, 701-710) */
void v3_cvorticity(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp, dt = P->dt;
    const double Ptl = P->Ptl, Ra = P->Ra, Re = P->Re, Gr = P->Gr;
    const double *A = F->A, *B = F->B, *C = F->C, *Cp = F->Cp, *th = F->th, *uel = F->uel, *vel = F->vel,
                 *wel = F->wel;
    for (int k = 0; k <= np; ++k)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nr - 1; ++i) {
                const int ip = i + 1, im = i - 1, jm = j - 1;
                const int km = (k == 0) ? np - 1 : k - 1;
                const int ka = (k == np) ? 1 : k + 1;
                const double r = P->r0 + (double)i * dr;
                const double u = (X(uel, i, j, ka) + X(uel, i, j, k)) * 0.5;
                const double v = (X(vel, i, j, k) + X(vel, ip, j, k)) * 0.5;
                const double w = (X(wel, i, j, k) + X(wel, ip, j, k) + X(wel, i, j, ka) + X(wel, ip, j, ka) +
                                  X(wel, i, jm, k) + X(wel, ip, jm, k) + X(wel, i, jm, ka) + X(wel, ip, jm, ka)) *
                                 0.125;
                const double AA = (X(A, i, j, k) + X(A, i, jm, k) + X(A, ip, j, k) + X(A, ip, jm, k)) * 0.25;
                const double BB = (X(B, i, j, k) + X(B, i, jm, k) + X(B, i, j, ka) + X(B, i, jm, ka)) * 0.25;
                const double tem = (X(th, i, j, k) + X(th, ip, j, k) + X(th, ip, j, ka) + X(th, i, j, ka)) * 0.25;
                const double Crra = (X(C, i - 1, j, k) + X(C, i + 1, j, k)) / (dr * dr);
                const double Cra = (X(C, i + 1, j, k) - X(C, i - 1, j, k)) * 0.5 / dr;
                double Cr = Cra;
                {
                    const double Pe = fabs(u * dr / Ptl);
                    if (Pe > 2) {
                        if (u > 0) Cr = (X(C, i, j, k) - X(C, i - 1, j, k)) / dr;
                        else if (u < 0) Cr = (X(C, i + 1, j, k) - X(C, i, j, k)) / dr;
                    }
                }
                const double welr = ((X(wel, ip, j, k) + X(wel, ip, j, ka) + X(wel, ip, jm, k) + X(wel, ip, jm, ka)) * 0.25 -
                                     (X(wel, i, j, k) + X(wel, i, j, ka) + X(wel, i, jm, k) + X(wel, i, jm, ka)) * 0.25) /
                                    dr;
                const double uelr = (X(uel, ip, j, k) + X(uel, ip, j, ka) - X(uel, im, j, k) - X(uel, im, j, ka)) * 0.25 / dr;
                const double thr = (X(th, ip, j, k) + X(th, ip, j, ka) - X(th, i, j, k) - X(th, i, j, ka)) * 0.5 / dr;
                (void)uelr;   /* computed by the reference (rot3dmg.f:663-664), unused */
                double Czza, Cz, cnz;
                if (j == 1) {
                    Czza = (3.2 * X(C, i, j - 1, k) + 2.0 * X(C, i, j + 1, k) - 0.2 * X(C, i, j + 2, k)) / (dz * dz);
                    Cz = (-4.0 * X(C, i, j - 1, k) / 3.0 + X(C, i, j, k) + X(C, i, j + 1, k) / 3.0) / dz;
                    cnz = -2.5 / (dz * dz);
                } else if (j == nz) {
                    Czza = (3.2 * X(C, i, j + 1, k) + 2.0 * X(C, i, j - 1, k) - 0.2 * X(C, i, j - 2, k)) / (dz * dz);
                    Cz = (-X(C, i, j - 1, k) / 3.0 - X(C, i, j, k) + 4.0 * X(C, i, j + 1, k) / 3.0) / dz;
                    cnz = -2.5 / (dz * dz);
                } else {
                    Czza = (X(C, i, j - 1, k) + X(C, i, j + 1, k)) / (dz * dz);
                    Cz = (X(C, i, j + 1, k) - X(C, i, j - 1, k)) * 0.5 / dz;
                    const double Pe = fabs(w * dz / Ptl);
                    if (Pe > 2) {
                        if (w > 0) Cz = (X(C, i, j, k) - X(C, i, j - 1, k)) / dz;
                        else if (w < 0) Cz = (X(C, i, j + 1, k) - X(C, i, j, k)) / dz;
                    }
                    cnz = -1.0 / (dz * dz);
                }
                const double welz = ((X(wel, i, j, k) + X(wel, ip, j, k) + X(wel, i, j, ka) + X(wel, ip, j, ka)) * 0.25 -
                                     (X(wel, i, jm, k) + X(wel, ip, jm, k) + X(wel, i, jm, ka) + X(wel, ip, jm, ka)) * 0.25) /
                                    dz;
                const double Chha = (X(C, i, j, ka) + X(C, i, j, km)) / (dp * dp);
                double Ch = (X(C, i, j, ka) - X(C, i, j, km)) * 0.5 / dp;
                {
                    const double Pe = fabs(v * dp * r / Ptl);
                    if (Pe > 2) {
                        if (v > 0) Ch = (X(C, i, j, k) - X(C, i, j, km)) / dp;
                        else if (v < 0) Ch = (X(C, i, j, ka) - X(C, i, j, k)) / dp;
                    }
                }
                const double welh = ((X(wel, i, j, ka) + X(wel, ip, j, ka) + X(wel, i, jm, ka) + X(wel, ip, jm, ka)) * 0.25 -
                                     (X(wel, i, j, k) + X(wel, ip, j, k) + X(wel, i, jm, k) + X(wel, ip, jm, k)) * 0.25) /
                                    dp;
                const double velh = (X(vel, i, j, ka) + X(vel, ip, j, ka) - X(vel, i, j, km) - X(vel, ip, j, km)) * 0.25 / dp;
                (void)velh;   /* computed by the reference (rot3dmg.f:714-715), unused */
                const double thh = (X(th, i, j, ka) + X(th, ip, j, ka) - X(th, i, j, k) - X(th, ip, j, k)) * 0.5 / dp;
                const double cnr = -1.0 / (dr * dr);
                const double cnh = -1.0 / (dp * dp) / (r * r);
                const double RHS1 = Crra + Cra / r + Chha / (r * r) + Czza + X(C, i, j, k) / (Ptl * dt);
                const double ULT = (u * Cr + v * Ch / r + w * Cz - welr * AA - welh * BB / r) / Ptl;
                /* rot3dmghor.f:749 */
                const double RHS2 = (P->model == 1)
                                        ? -1.0 / r * Gr * Ptl * cos(hor_theta(P, k)) * thh -
                                              Gr * Ptl * sin(hor_theta(P, k)) * thr
                                        : -2.0 * Re * welz - Ra * thh +
                                              2.0 * Ra * (tem * welz - thr * u - thh * v / r) / Re / Ptl;
                const double cf1 = 1.0 / (Ptl * dt) - welz / (2.0 * Ptl) - cnr - cnz - cnh;
                const double cf2 = welz / (2.0 * Ptl) + cnr + cnz + cnh;
                X(F->Cn, i, j, k) = (cf2 * X(Cp, i, j, k) - ULT - RHS2 + RHS1) / cf1;
            }
}

/* temperature, rot3dmg.f:736-874 (hybrid upwind 797-806, 834-843, 851-860) */
void v3_temperature(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp, dt = P->dt;
    const double Ptl = P->Ptl;
    const double *th = F->th, *thp = F->thp, *uel = F->uel, *vel = F->vel, *wel = F->wel;
    for (int k = 0; k <= np; ++k)
        for (int j = 0; j <= nz + 1; ++j)
            for (int i = 1; i <= nr; ++i) {
                const int im = i - 1, jm = j - 1;
                const int km = (k == 0) ? np - 1 : k - 1;
                const int ka = (k == np) ? 1 : k + 1;
                const double r = P->r0 + ((double)i - 0.5) * dr;
                double u, v, w;
                if (j == 0 || j == nz + 1) {
                    u = 0.0;
                    v = 0.0;
                    w = 0.0;
                } else {
                    u = (X(uel, i, j, k) + X(uel, im, j, k)) * 0.5;
                    v = (X(vel, i, j, k) + X(vel, i, j, km)) * 0.5;
                    w = (X(wel, i, j, k) + X(wel, i, jm, k)) * 0.5;
                }
                double thrra, thra, thr, cnr;
                if (i == 1) {
                    thrra = (3.2 * X(th, i - 1, j, k) + 2.0 * X(th, i + 1, j, k) - 0.2 * X(th, i + 2, j, k)) / (dr * dr);
                    thra = (-4.0 * X(th, i - 1, j, k) / 3.0 + X(th, i + 1, j, k) / 3.0) / dr;
                    thr = (-4.0 * X(th, i - 1, j, k) / 3.0 + X(th, i, j, k) + X(th, i + 1, j, k) / 3.0) / dr;
                    cnr = -2.5 / (dr * dr) + 0.5 / dr / r;
                } else if (i == nr) {
                    thrra = (3.2 * X(th, i + 1, j, k) + 2.0 * X(th, i - 1, j, k) - 0.2 * X(th, i - 2, j, k)) / (dr * dr);
                    thra = (-X(th, i - 1, j, k) / 3.0 + 4.0 * X(th, i + 1, j, k) / 3.0) / dr;
                    thr = (-X(th, i - 1, j, k) / 3.0 - X(th, i, j, k) + 4.0 * X(th, i + 1, j, k) / 3.0) / dr;
                    cnr = -2.5 / (dr * dr) - 0.5 / dr / r;
                } else {
                    thrra = (X(th, i - 1, j, k) + X(th, i + 1, j, k)) / (dr * dr);
                    thra = (X(th, i + 1, j, k) - X(th, i - 1, j, k)) * 0.5 / dr;
                    thr = thra;
                    const double Pe = fabs(u * dr / Ptl);
                    if (Pe > 2) {
                        if (u > 0) thr = (X(th, i, j, k) - X(th, i - 1, j, k)) / dr;
                        else if (u < 0) thr = (X(th, i + 1, j, k) - X(th, i, j, k)) / dr;
                    }
                    cnr = -1.0 / (dr * dr);
                }
                double thzza, thz, cnz;
                if (j == 1) {
                    thzza = (3.2 * X(th, i, j - 1, k) + 2.0 * X(th, i, j + 1, k) - 0.2 * X(th, i, j + 2, k)) / (dz * dz);
                    thz = (-4.0 * X(th, i, j - 1, k) / 3.0 + X(th, i, j, k) + X(th, i, j + 1, k) / 3.0) / dz;
                    cnz = -2.5 / (dz * dz);
                } else if (j == nz) {
                    thzza = (3.2 * X(th, i, j + 1, k) + 2.0 * X(th, i, j - 1, k) - 0.2 * X(th, i, j - 2, k)) / (dz * dz);
                    thz = (-X(th, i, j - 1, k) / 3.0 - X(th, i, j, k) + 4.0 * X(th, i, j + 1, k) / 3.0) / dz;
                    cnz = -2.5 / (dz * dz);
                } else if (j == 0) {
                    thzza = 8.0 * X(th, i, j + 1, k) / (dz * dz);
                    thz = 0.0;
                    cnz = -4.0 / (dz * dz);
                } else if (j == nz + 1) {
                    thzza = 8.0 * X(th, i, j - 1, k) / (dz * dz);
                    thz = 0.0;
                    cnz = -4.0 / (dz * dz);
                } else {
                    thzza = (X(th, i, j - 1, k) + X(th, i, j + 1, k)) / (dz * dz);
                    thz = (X(th, i, j + 1, k) - X(th, i, j - 1, k)) * 0.5 / dz;
                    const double Pe = fabs(w * dz / Ptl);
                    if (Pe > 2) {
                        if (w > 0) thz = (X(th, i, j, k) - X(th, i, j - 1, k)) / dz;
                        else if (w < 0) thz = (X(th, i, j + 1, k) - X(th, i, j, k)) / dz;
                    }
                    cnz = -1.0 / (dz * dz);
                }
                const double thhha = (X(th, i, j, ka) + X(th, i, j, km)) / (dp * dp);
                double thh = (X(th, i, j, ka) - X(th, i, j, km)) * 0.5 / dp;
                {
                    const double Pe = fabs(v * dp * r / Ptl);
                    if (Pe > 2) {
                        if (v > 0) thh = (X(th, i, j, k) - X(th, i, j, km)) / dp;
                        else if (v < 0) thh = (X(th, i, j, ka) - X(th, i, j, k)) / dp;
                    }
                }
                const double cnp = -1.0 / (dp * dp) / (r * r);
                const double RHS1 = thrra + thra / r + thhha / (r * r) + thzza + X(th, i, j, k) / dt;
                const double ULT = u * thr + v * thh / r + w * thz;
                const double cf1 = 1.0 / dt - cnr - cnz - cnp;
                const double cf2 = cnr + cnz + cnp;
                X(F->thn, i, j, k) = (cf2 * X(thp, i, j, k) - ULT + RHS1) / cf1;
            }
}

/* heat of rot3dmghor.f:913-946: the fluxes through the inner (trans1) and
 * outer (trans2) cylinders over nz/2-3 <= j <= nz/2+3 (weights 1/120 at the
 * two ends, 1/60 between) and phi */
static void heat_hor(const v3_params *P, const v3_fields *F, double out[2])
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp;
    const double *th = F->th;
    const int nza = nz / 2 - 3, nzb = nz / 2 + 3;
    double trans1 = 0.0, trans2 = 0.0;
    for (int k = 1; k <= np; ++k) {
        double transa = dz * (-184.0 * X(th, 0, nza, k) + 225.0 * X(th, 1, nza, k) - 50.0 * X(th, 2, nza, k) +
                              9.0 * X(th, 3, nza, k)) / (120.0 * dr);
        double transb = dz * (184.0 * X(th, nr + 1, nza, k) - 225.0 * X(th, nr, nza, k) +
                              50.0 * X(th, nr - 1, nza, k) - 9.0 * X(th, nr - 2, nza, k)) / (120.0 * dr);
        transa = transa + dz * (-184.0 * X(th, 0, nzb, k) + 225.0 * X(th, 1, nzb, k) - 50.0 * X(th, 2, nzb, k) +
                                9.0 * X(th, 3, nzb, k)) / (120.0 * dr);
        transb = transb + dz * (184.0 * X(th, nr + 1, nzb, k) - 225.0 * X(th, nr, nzb, k) +
                                50.0 * X(th, nr - 1, nzb, k) - 9.0 * X(th, nr - 2, nzb, k)) / (120.0 * dr);
        for (int j = nza + 1; j <= nzb - 1; ++j) {
            transa = transa + dz * (-184.0 * X(th, 0, j, k) + 225.0 * X(th, 1, j, k) - 50.0 * X(th, 2, j, k) +
                                    9.0 * X(th, 3, j, k)) / (60.0 * dr);
            transb = transb + dz * (184.0 * X(th, nr + 1, j, k) - 225.0 * X(th, nr, j, k) +
                                    50.0 * X(th, nr - 1, j, k) - 9.0 * X(th, nr - 2, j, k)) / (60.0 * dr);
        }
        trans1 = trans1 + transa * dp;
        trans2 = trans2 + transb * dp;
    }
    out[0] = trans1;
    out[1] = trans2;
}

void v3_heat2(const v3_params *P, const v3_fields *F, double out[2])
{
    if (P->model == 1) {
        heat_hor(P, F, out);
    } else {
        out[0] = v3_heat(P, F);
        out[1] = 0.0;
    }
}

/* heat, rot3dmg.f:882-905: inner-wall flux integrated over z (trapezoid-
 * like weights 1/240, 1/80, 1/60) and phi; rot3dmghor.f's trans1 for
 * model 1 */
double v3_heat(const v3_params *P, const v3_fields *F)
{
    if (P->model == 1) {
        double out[2];
        heat_hor(P, F, out);
        return out[0];
    }
    const int nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp;
    const double *th = F->th;
    double trans1 = 0.0;
    for (int k = 1; k <= np; ++k) {
        double trans = dz * (-184.0 * X(th, 0, 0, k) + 225.0 * X(th, 1, 0, k) - 50.0 * X(th, 2, 0, k) +
                             9.0 * X(th, 3, 0, k)) / (240.0 * dr);
        trans = trans + dz * (-184.0 * X(th, 0, 1, k) + 225.0 * X(th, 1, 1, k) - 50.0 * X(th, 2, 1, k) +
                              9.0 * X(th, 3, 1, k)) / (80.0 * dr);
        trans = trans + dz * (-184.0 * X(th, 0, nz, k) + 225.0 * X(th, 1, nz, k) - 50.0 * X(th, 2, nz, k) +
                              9.0 * X(th, 3, nz, k)) / (80.0 * dr);
        trans = trans + dz * (-184.0 * X(th, 0, nz + 1, k) + 225.0 * X(th, 1, nz + 1, k) -
                              50.0 * X(th, 2, nz + 1, k) + 9.0 * X(th, 3, nz + 1, k)) / (240.0 * dr);
        for (int j = 2; j <= nz - 1; ++j)
            trans = trans + dz * (-184.0 * X(th, 0, j, k) + 225.0 * X(th, 1, j, k) - 50.0 * X(th, 2, j, k) +
                                  9.0 * X(th, 3, j, k)) / (60.0 * dr);
        trans1 = trans1 + trans * dp;
    }
    return trans1;
}

/* update, rot3dmg.f:964-1018: wall vorticity from the new velocities, wall
 * temperatures, then the time-level rotation of A, B, C, th */
void v3_update(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz;
    const double *uel = F->uel, *vel = F->vel, *wel = F->wel;
    for (int k = 0; k <= np; ++k)
        for (int i = 1; i <= nr; ++i) {
            X(F->An, i, 0, k) = -(3.0 * X(vel, i, 1, k) - X(vel, i, 2, k) / 3.0) / dz;
            X(F->An, i, nz, k) = -(X(vel, i, nz - 1, k) / 3.0 - 3.0 * X(vel, i, nz, k)) / dz;
        }
    for (int k = 0; k <= np; ++k)
        for (int i = 1; i <= nr - 1; ++i) {
            X(F->Bn, i, 0, k) = (3.0 * X(uel, i, 1, k) - X(uel, i, 2, k) / 3.0) / dz;
            X(F->Bn, i, nz, k) = (-3.0 * X(uel, i, nz, k) + X(uel, i, nz - 1, k) / 3.0) / dz;
        }
    for (int k = 0; k <= np; ++k)
        for (int j = 1; j <= nz - 1; ++j) {
            X(F->Bn, 0, j, k) = -(3.0 * X(wel, 1, j, k) - X(wel, 2, j, k) / 3.0) / dr;
            X(F->Bn, nr, j, k) = -(X(wel, nr - 1, j, k) / 3.0 - 3.0 * X(wel, nr, j, k)) / dr;
        }
    for (int k = 0; k <= np; ++k)
        for (int j = 1; j <= nz; ++j) {
            X(F->Cn, 0, j, k) = (3.0 * X(vel, 1, j, k) - X(vel, 2, j, k) / 3.0) / dr;
            X(F->Cn, nr, j, k) = (X(vel, nr - 1, j, k) / 3.0 - 3.0 * X(vel, nr, j, k)) / dr;
        }
    for (int k = 0; k <= np; ++k)
        for (int j = 0; j <= nz + 1; ++j) {
            /* rot3dmghor.f:1042-1043 swaps the wall temperatures */
            X(F->thn, 0, j, k) = (P->model == 1) ? 1.0 : 0.0;
            X(F->thn, nr + 1, j, k) = (P->model == 1) ? 0.0 : 1.0;
        }
    for (int k = 0; k <= np; ++k)
        for (int j = 0; j <= nz + 1; ++j)
            for (int i = 0; i <= nr + 1; ++i) {
                X(F->Ap, i, j, k) = X(F->A, i, j, k);
                X(F->A, i, j, k) = X(F->An, i, j, k);
                X(F->Bp, i, j, k) = X(F->B, i, j, k);
                X(F->B, i, j, k) = X(F->Bn, i, j, k);
                X(F->Cp, i, j, k) = X(F->C, i, j, k);
                X(F->C, i, j, k) = X(F->Cn, i, j, k);
                X(F->thp, i, j, k) = X(F->th, i, j, k);
                X(F->th, i, j, k) = X(F->thn, i, j, k);
            }
}

/* kinetic energy, the build's definition (vv3_oracle.h) */
double v3_energy(const v3_params *P, const v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    double s = 0.0;
    for (int k = 0; k < np; ++k) {
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nr - 1; ++i) {
                const double u = X(F->uel, i, j, k);
                s += (P->r0 + (double)i * P->dr) * u * u;
            }
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nr; ++i) {
                const double v = X(F->vel, i, j, k);
                s += (P->r0 + ((double)i - 0.5) * P->dr) * v * v;
            }
        for (int j = 1; j <= nz - 1; ++j)
            for (int i = 1; i <= nr; ++i) {
                const double w = X(F->wel, i, j, k);
                s += (P->r0 + ((double)i - 0.5) * P->dr) * w * w;
            }
    }
    return 0.5 * P->dr * P->dz * P->dp * s;
}

/* local Nusselt numbers, sect360v2mg.f:318-325 (vv3_oracle.h) */
void v3_local_nu(const v3_params *P, const v3_fields *F, double *inner, double *outer)
{
    const int nr = P->nr, jm = P->nz / 2;
    const double dr = P->dr, r0 = P->r0, rmax = P->rmax, L = log(r0);
    const double *th = F->th;
    for (int k = 0; k <= P->np; ++k) {
        inner[k] = -(r0 * L * (-184.0 * X(th, 0, jm, k) + 225.0 * X(th, 1, jm, k) - 50.0 * X(th, 2, jm, k) +
                               9.0 * X(th, 3, jm, k)) / (60.0 * dr));
        outer[k] = -(rmax * L * (184.0 * X(th, nr + 1, jm, k) - 225.0 * X(th, nr, jm, k) +
                                 50.0 * X(th, nr - 1, jm, k) - 9.0 * X(th, nr - 2, jm, k)) / (60.0 * dr));
    }
}

/* Fourier-mode integrals after rot3dmghor.f:954-991 (vv3_oracle.h) */
void v3_modes(const v3_params *P, const v3_fields *F, int nmax, double *c, double *s)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp;
    for (int n = 1; n <= nmax; ++n) {
        double sc = 0.0, ss = 0.0;
        for (int k = 1; k <= np; ++k) {
            const double ang = (double)k * dp;
            double sumc = 0.0, sums = 0.0;
            for (int i = 0; i <= nr + 1; ++i) {
                const double r = P->r0 + (double)i * dr;
                for (int j = 0; j <= nz + 1; ++j) {
                    double fact = 1.0;
                    if (i == 0 || i == nr + 1) fact = fact / 2.0;
                    if (j == 0 || j == nz + 1) fact = fact / 2.0;
                    sumc = sumc + r * X(F->th, i, j, k) * cos((double)n * ang) * fact;
                    sums = sums + r * X(F->th, i, j, k) * sin((double)n * ang) * fact;
                }
            }
            sc = sc + sumc * dr * dz * dp;
            ss = ss + sums * dr * dz * dp;
        }
        c[n - 1] = sc;
        s[n - 1] = ss;
    }
}

/* rhsw, rot3dmg.f:1026-1048 */
void v3_rhsw(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dp = P->dp;
    const double *An = F->An, *Bn = F->Bn;
    for (int k = 0; k <= np; ++k)
        for (int j = 1; j <= nz - 1; ++j)
            for (int i = 1; i <= nr; ++i) {
                const int km = (k == 0) ? np - 1 : k - 1;
                const double r = P->r0 + ((double)i - 0.5) * dr;
                const double Ah = (X(An, i, j, k) - X(An, i, j, km)) / dp;
                const double Br = (X(Bn, i, j, k) - X(Bn, i - 1, j, k)) / dr;
                double v;
                if (i == 1) v = Ah / r - X(Bn, 1, j, k) / dr - 0.5 * X(Bn, 1, j, k) / r;
                else if (i == nr) v = Ah / r + X(Bn, nr - 1, j, k) / dr - 0.5 * X(Bn, nr - 1, j, k) / r;
                else v = Ah / r - Br - 0.5 * (X(Bn, i, j, k) + X(Bn, i - 1, j, k)) / r;
                X(F->wrh, i, j, k) = v;
            }
}

/* rhsv, rot3dmg.f:1066-1093 */
void v3_rhsv(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz;
    const double *An = F->An, *Cn = F->Cn;
    for (int k = 0; k <= np; ++k)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nr; ++i) {
                const double r = P->r0 + ((double)i - 0.5) * dr;
                const double Az = (X(An, i, j, k) - X(An, i, j - 1, k)) / dz;
                const double Cr = (X(Cn, i, j, k) - X(Cn, i - 1, j, k)) / dr;
                double v;
                if (i == 1) v = X(Cn, 1, j, k) / dr + X(Cn, 1, j, k) / r;
                else if (i == nr) v = -X(Cn, nr - 1, j, k) / dr + X(Cn, nr - 1, j, k) / r;
                else v = Cr + (X(Cn, i, j, k) + X(Cn, i - 1, j, k)) / r;
                if (j == 1) v = v - X(An, i, 1, k) / dz;
                else if (j == nz) v = v + X(An, i, nz - 1, k) / dz;
                else v = v - Az;
                X(F->vrh, i, j, k) = v;
            }
}

/* rhsu, rot3dmg.f:1111-1134 (uses the new w of poisson1) */
void v3_rhsu(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp;
    const double *Bn = F->Bn, *Cn = F->Cn, *wel = F->wel;
    for (int k = 0; k <= np; ++k)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nr - 1; ++i) {
                const int km = (k == 0) ? np - 1 : k - 1;
                const double r = P->r0 + (double)i * dr;
                const double Bz = (X(Bn, i, j, k) - X(Bn, i, j - 1, k)) / dz;
                const double welz = 0.5 * (X(wel, i, j, k) + X(wel, i + 1, j, k) - X(wel, i, j - 1, k) -
                                           X(wel, i + 1, j - 1, k)) / dz;
                const double Ch = (X(Cn, i, j, k) - X(Cn, i, j, km)) / dp;
                double v;
                if (j == 1) v = X(Bn, i, 1, k) / dz - Ch / r - 2.0 * welz / r;
                else if (j == nz) v = -X(Bn, i, nz - 1, k) / dz - Ch / r - 2.0 * welz / r;
                else v = Bz - Ch / r - 2.0 * welz / r;
                X(F->urh, i, j, k) = v;
            }
}

/* unknown ranges of the three operators: i in [i0, i1], j in [j0, j1],
 * k in [1, np] (k = 0 is k = np) */
static void ranges(const v3_params *P, int which, int *i1, int *j1)
{
    if (which == 1) { *i1 = P->nr; *j1 = P->nz - 1; }
    else if (which == 2) { *i1 = P->nr; *j1 = P->nz; }
    else { *i1 = P->nr - 1; *j1 = P->nz; }
}

/* L psi at (i, j, k) for operator `which` at the finest level, in the
 * reference residual's operation order */
static double apply_L(const v3_params *P, int which, const double *ps, int i, int j, int k)
{
    const int n1 = P->nr, n2 = P->nz, n3 = P->np;
    const double hr = P->dr, hz = P->dz, hp = P->dp;
    const double hhr = hr * hr, hhz = hz * hz, hhp = hp * hp;
    const int ka = (k == n3) ? 1 : k + 1;
    const int km = k - 1;
    double temp;
    if (which == 1) {   /* rot3dmg.f:1698-1719 */
        const double rval = P->r0 + ((double)i - 0.5) * hr;
        temp = (X(ps, i, j + 1, k) + X(ps, i, j - 1, k)) / hhz + (X(ps, i, j, ka) + X(ps, i, j, km)) / hhp / (rval * rval) -
               X(ps, i, j, k) * (2.0 / hhz + 2.0 / hhp / (rval * rval));
        if (i == 1)
            temp = temp + (-2.0 * X(ps, i, j, k) + 5.0 * X(ps, i + 1, j, k) / 3.0 - 0.2 * X(ps, i + 2, j, k)) / hhr +
                   0.5 * (-X(ps, i, j, k) + X(ps, i + 1, j, k)) / hr / rval;
        else if (i == n1)
            temp = temp + (-2.0 * X(ps, i, j, k) + 5.0 * X(ps, i - 1, j, k) / 3.0 - 0.2 * X(ps, i - 2, j, k)) / hhr +
                   0.5 * (X(ps, i, j, k) - X(ps, i - 1, j, k)) / hr / rval;
        else
            temp = temp + (X(ps, i - 1, j, k) - 2.0 * X(ps, i, j, k) + X(ps, i + 1, j, k)) / hhr +
                   0.5 * (X(ps, i + 1, j, k) - X(ps, i - 1, j, k)) / hr / rval;
    } else if (which == 2) {   /* rot3dmg.f:1935-1961 */
        const double rval = P->r0 + ((double)i - 0.5) * hr;
        temp = (X(ps, i, j, ka) - 2.0 * X(ps, i, j, k) + X(ps, i, j, km)) / hhp / (rval * rval) +
               X(ps, i, j, k) / (rval * rval);
        if (i == 1)
            temp = temp + (-2.0 * X(ps, i, j, k) + 5.0 / 3.0 * X(ps, i + 1, j, k) - 0.2 * X(ps, i + 2, j, k)) / hhr +
                   4.0 / 3.0 * X(ps, i + 1, j, k) / hr / rval;
        else if (i == n1)
            temp = temp + (-2.0 * X(ps, i, j, k) + 5.0 / 3.0 * X(ps, i - 1, j, k) - 0.2 * X(ps, i - 2, j, k)) / hhr -
                   4.0 / 3.0 * X(ps, i - 1, j, k) / hr / rval;
        else
            temp = temp + (X(ps, i - 1, j, k) - 2.0 * X(ps, i, j, k) + X(ps, i + 1, j, k)) / hhr +
                   1.5 * (X(ps, i + 1, j, k) - X(ps, i - 1, j, k)) / hr / rval;
        if (j == 1)
            temp = temp + (-2.0 * X(ps, i, j, k) + 5.0 / 3.0 * X(ps, i, j + 1, k) - 0.2 * X(ps, i, j + 2, k)) / hhz;
        else if (j == n2)
            temp = temp + (-2.0 * X(ps, i, j, k) + 5.0 / 3.0 * X(ps, i, j - 1, k) - 0.2 * X(ps, i, j - 2, k)) / hhz;
        else
            temp = temp + (X(ps, i, j + 1, k) - 2.0 * X(ps, i, j, k) + X(ps, i, j - 1, k)) / hhz;
    } else {   /* rot3dmg.f:2133-2154 */
        const double rval = (double)i * hr + P->r0;
        temp = (X(ps, i + 1, j, k) + X(ps, i - 1, j, k)) / hhr + 1.5 * (X(ps, i + 1, j, k) - X(ps, i - 1, j, k)) / hr / rval +
               X(ps, i, j, k) / (rval * rval) + (X(ps, i, j, ka) + X(ps, i, j, km)) / hhp / (rval * rval) -
               X(ps, i, j, k) * (2.0 / hhr + 2.0 / hhp / (rval * rval));
        if (j == 1)
            temp = temp + (-2.0 * X(ps, i, j, k) + 5.0 * X(ps, i, j + 1, k) / 3.0 - 0.2 * X(ps, i, j + 2, k)) / hhz;
        else if (j == n2)
            temp = temp + (-2.0 * X(ps, i, j, k) + 5.0 * X(ps, i, j - 1, k) / 3.0 - 0.2 * X(ps, i, j - 2, k)) / hhz;
        else
            temp = temp + (X(ps, i, j + 1, k) - 2.0 * X(ps, i, j, k) + X(ps, i, j - 1, k)) / hhz;
    }
    return temp;
}

double v3_residual(const v3_params *P, int which, const double *psi, const double *rhs, double *res)
{
    int i1, j1;
    ranges(P, which, &i1, &j1);
    const size_t n = (size_t)P->ldi * P->ldj * (P->np + 1);
    if (res) memset(res, 0, n * sizeof(double));
    double m = 0.0;
    for (int k = 1; k <= P->np; ++k)
        for (int j = 1; j <= j1; ++j)
            for (int i = 1; i <= i1; ++i) {
                const double t = X(rhs, i, j, k) - apply_L(P, which, psi, i, j, k);
                if (res) X(res, i, j, k) = t;
                if (fabs(t) > m) m = fabs(t);
            }
    return m;
}

/* dense LU with partial pivoting of an n x n row-major matrix, in place */
static int lu_factor(double *a, int *piv, int n)
{
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(a[(size_t)r * n + c]) > fabs(a[(size_t)p * n + c])) p = r;
        if (a[(size_t)p * n + c] == 0.0) return -1;
        piv[c] = p;
        if (p != c)
            for (int q = 0; q < n; ++q) {
                const double t = a[(size_t)c * n + q];
                a[(size_t)c * n + q] = a[(size_t)p * n + q];
                a[(size_t)p * n + q] = t;
            }
        for (int r = c + 1; r < n; ++r) {
            const double f = a[(size_t)r * n + c] / a[(size_t)c * n + c];
            a[(size_t)r * n + c] = f;
            for (int q = c + 1; q < n; ++q) a[(size_t)r * n + q] -= f * a[(size_t)c * n + q];
        }
    }
    return 0;
}

static void lu_solve(const double *a, const int *piv, int n, double *b)
{
    for (int c = 0; c < n; ++c) {
        const int p = piv[c];
        if (p != c) { const double t = b[c]; b[c] = b[p]; b[p] = t; }
        for (int r = c + 1; r < n; ++r) b[r] -= a[(size_t)r * n + c] * b[c];
    }
    for (int c = n - 1; c >= 0; --c) {
        double s = b[c];
        for (int q = c + 1; q < n; ++q) s -= a[(size_t)c * n + q] * b[q];
        b[c] = s / a[(size_t)c * n + c];
    }
}

int v3_poisson_exact(const v3_params *P, int which, double *psi, const double *rhs)
{
    int i1, j1;
    ranges(P, which, &i1, &j1);
    const int np = P->np, ni = i1, nj = j1, m = ni * nj;
    const size_t nfield = (size_t)P->ldi * P->ldj * (np + 1);
    const double pi = acos(-1.0);
    memset(psi, 0, nfield * sizeof(double));
    double *a = (double *)malloc((size_t)m * m * sizeof(double));
    int *piv = (int *)malloc((size_t)m * sizeof(int));
    double *bc = (double *)malloc((size_t)m * sizeof(double));
    double *bs = (double *)malloc((size_t)m * sizeof(double));
    double *e = (double *)calloc(nfield, sizeof(double));
    double *spec = (double *)calloc((size_t)m * (np + 1) * 2, sizeof(double));
    if (!a || !piv || !bc || !bs || !e || !spec) { free(a); free(piv); free(bc); free(bs); free(e); free(spec); return -1; }
    int st = 0;
    for (int mode = 0; mode <= np / 2 && !st; ++mode) {
        /* the operator applied to unit vectors e_{i,j} * cos(mode phi_k)
         * gives column (i,j) of the mode's (r,z) matrix times cos(mode phi) */
        for (int c = 0; c < m; ++c) {
            const int ic = 1 + c % ni, jc = 1 + c / ni;
            memset(e, 0, nfield * sizeof(double));
            for (int k = 0; k <= np; ++k) X(e, ic, jc, k) = cos(2.0 * pi * (double)((long)mode * k % np) / np);
            for (int rr = 0; rr < m; ++rr) {
                const int ir = 1 + rr % ni, jr = 1 + rr / ni;
                /* evaluate at k = np (phase 1): coefficient of the mode */
                a[(size_t)rr * m + c] = apply_L(P, which, e, ir, jr, np);
            }
        }
        if (lu_factor(a, piv, m)) { st = -1; break; }
        /* projections of rhs on cos / sin of the mode (k = 1..np) */
        for (int rr = 0; rr < m; ++rr) {
            const int ir = 1 + rr % ni, jr = 1 + rr / ni;
            long double sc = 0.0L, ss = 0.0L;
            for (int k = 1; k <= np; ++k) {
                const long double ang = 2.0L * (long double)pi * (long double)((long)mode * k % np) / np;
                sc += (long double)X(rhs, ir, jr, k) * cosl(ang);
                ss += (long double)X(rhs, ir, jr, k) * sinl(ang);
            }
            bc[rr] = (double)sc;
            bs[rr] = (double)ss;
        }
        lu_solve(a, piv, m, bc);
        lu_solve(a, piv, m, bs);
        for (int rr = 0; rr < m; ++rr) {
            spec[((size_t)mode * m + rr) * 2] = bc[rr];
            spec[((size_t)mode * m + rr) * 2 + 1] = bs[rr];
        }
    }
    if (!st) {
        for (int rr = 0; rr < m; ++rr) {
            const int ir = 1 + rr % ni, jr = 1 + rr / ni;
            for (int k = 1; k <= np; ++k) {
                long double s = 0.0L;
                for (int mode = 0; mode <= np / 2; ++mode) {
                    const long double ang = 2.0L * (long double)pi * (long double)((long)mode * k % np) / np;
                    const long double w = (mode == 0 || 2 * mode == np) ? 1.0L : 2.0L;
                    s += w * ((long double)spec[((size_t)mode * m + rr) * 2] * cosl(ang) +
                              (long double)spec[((size_t)mode * m + rr) * 2 + 1] * sinl(ang));
                }
                X(psi, ir, jr, k) = (double)(s / np);
            }
            X(psi, ir, jr, 0) = X(psi, ir, jr, np);
        }
    }
    free(a); free(piv); free(bc); free(bs); free(e); free(spec);
    return st;
}
