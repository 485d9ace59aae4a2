/* vvs_oracle.c -- CPU oracle (test infrastructure only) of sect360v2mg.f's
 * 3D step: the rotating annulus with a radial barrier (vvs_oracle.h).  Each
 * routine restates the reference subroutine from its text, in its
 * operation order; the exact Poisson solve is this oracle's own (a dense LU
 * of the whole operator) and is checked against the restated residual
 * operators, which the tests pin by conservation, self-adjointness, closed-
 * form eigenvectors and agreement with rot3dmg.f's operators away from the
 * barrier.
 */
#include "vvs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IX(i, j, k) (((size_t)(k) * P->ldj + (size_t)(j)) * P->ldi + (size_t)(i))
#define X(a, i, j, k) ((a)[IX(i, j, k)])

/* Avorticity, sect360v2mg.f:343-452: k = 1..np-1 (no wrap), the outer-wall
 * row i = nr with the one-sided Ar of dum15 (sect360v2mg.f:1268-1275) */
void vs_avorticity(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp, dt = P->dt;
    const double Ptl = P->Ptl, Ra = P->Ra, Re = P->Re, Gr = P->Gr;
    const double *A = F->A, *B = F->B, *C = F->C, *Ap = F->Ap, *th = F->th, *uel = F->uel, *vel = F->vel,
                 *wel = F->wel;
    for (int k = 1; k <= np - 1; ++k)
        for (int j = 1; j <= nz - 1; ++j)
            for (int i = 1; i <= nr; ++i) {
                const int im = i - 1, jp = j + 1, km = k - 1, ka = k + 1;
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
                    Ar = (-X(A, i - 1, j, k) / 3.0 - X(A, i, j, k) + 4.0 * X(A, i + 1, j, k) / 3.0) / dr;
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
                const double RHS2 = -2.0 * Re * uelz + 2.0 * Ra * (tem * uelz + thz * u) / Re / Ptl - Gr * Ptl / r * thh;
                const double cf1 = 1.0 / (Ptl * dt) - uelr / (2.0 * Ptl) + 1.0 / (2.0 * (r * r)) - cnr - cnz - cnh;
                const double cf2 = uelr / (2.0 * Ptl) - 1.0 / (2.0 * (r * r)) + cnr + cnz + cnh;
                X(F->An, i, j, k) = (cf2 * X(Ap, i, j, k) - ULT - RHS2 + RHS1) / cf1;
            }
}

/* Bvorticity, sect360v2mg.f:460-569: k = 1..np, the barrier rows k = 1 and
 * k = np with the wall closures of sect360v2mg.f:523-532 */
void vs_bvorticity(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp, dt = P->dt;
    const double Ptl = P->Ptl, Ra = P->Ra, Re = P->Re, Gr = P->Gr;
    const double *A = F->A, *B = F->B, *C = F->C, *Bp = F->Bp, *th = F->th, *uel = F->uel, *vel = F->vel,
                 *wel = F->wel;
    for (int k = 1; k <= np; ++k)
        for (int j = 1; j <= nz - 1; ++j)
            for (int i = 1; i <= nr - 1; ++i) {
                const int ip = i + 1, jp = j + 1, km = k - 1, ka = k + 1;
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
                double Bhha, Bh, cnh;
                if (k == 1) {
                    Bhha = (3.2 * X(B, i, j, 0) + 2.0 * X(B, i, j, 2) - 0.2 * X(B, i, j, 3)) / (dp * dp);
                    Bh = (-4.0 / 3.0 * X(B, i, j, 0) + X(B, i, j, 1) + X(B, i, j, 2) / 3.0) / dp;
                    cnh = -2.5 / (dp * dp) / (r * r);
                } else if (k == np) {
                    Bhha = (-0.2 * X(B, i, j, np - 2) + 2.0 * X(B, i, j, np - 1) + 3.2 * X(B, i, j, np + 1)) / (dp * dp);
                    Bh = (-X(B, i, j, np - 1) / 3.0 - X(B, i, j, np) + 4.0 / 3.0 * X(B, i, j, np + 1)) / dp;
                    cnh = -2.5 / (dp * dp) / (r * r);
                } else {
                    Bhha = (X(B, i, j, ka) + X(B, i, j, km)) / (dp * dp);
                    Bh = (X(B, i, j, ka) - X(B, i, j, km)) * 0.5 / dp;
                    cnh = -1.0 / (dp * dp) / (r * r);
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
                const double RHS1 = Brra + Bra / r + Bhha / (r * r) + Bzza + 2.0 * Ah / (r * r) + X(B, i, j, k) / (Ptl * dt);
                const double ULT = (u * Br + v * Bh / r + w * Bz - (velr - v / r) * AA - velz * CC) / Ptl;
                const double RHS2 = -2.0 * Re * velz + r * Ra * thz + 2.0 * Ra * (tem * velz + thz * v) / Re / Ptl +
                                    Gr * Ptl * thr;
                const double cf1 = 1.0 / (Ptl * dt) - (u / r + velh / r) / (2.0 * Ptl) + 1.0 / (2.0 * (r * r)) - cnr - cnz - cnh;
                const double cf2 = (u / r + velh / r) / (2.0 * Ptl) - 1.0 / (2.0 * (r * r)) + cnr + cnz + cnh;
                X(F->Bn, i, j, k) = (cf2 * X(Bp, i, j, k) - ULT - RHS2 + RHS1) / cf1;
            }
}

/* Cvorticity, sect360v2mg.f:577-684: k = 1..np-1 (no wrap) */
void vs_cvorticity(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp, dt = P->dt;
    const double Ptl = P->Ptl, Ra = P->Ra, Re = P->Re;
    const double *A = F->A, *B = F->B, *C = F->C, *Cp = F->Cp, *th = F->th, *uel = F->uel, *vel = F->vel,
                 *wel = F->wel;
    for (int k = 1; k <= np - 1; ++k)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nr - 1; ++i) {
                const int ip = i + 1, im = i - 1, jm = j - 1, km = k - 1, ka = k + 1;
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
                /* uelr (sect360v2mg.f:623-624) is computed by the reference and unused */
                const double thr = (X(th, ip, j, k) + X(th, ip, j, ka) - X(th, i, j, k) - X(th, i, j, ka)) * 0.5 / dr;
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
                /* velh (sect360v2mg.f:667-668) is computed by the reference and unused */
                const double thh = (X(th, i, j, ka) + X(th, ip, j, ka) - X(th, i, j, k) - X(th, ip, j, k)) * 0.5 / dp;
                const double cnr = -1.0 / (dr * dr);
                const double cnh = -1.0 / (dp * dp) / (r * r);
                const double RHS1 = Crra + Cra / r + Chha / (r * r) + Czza + X(C, i, j, k) / (Ptl * dt);
                const double ULT = (u * Cr + v * Ch / r + w * Cz - welr * AA - welh * BB / r) / Ptl;
                const double RHS2 = -2.0 * Re * welz - Ra * thh + 2.0 * Ra * (tem * welz - thr * u - thh * v / r) / Re / Ptl;
                const double cf1 = 1.0 / (Ptl * dt) - welz / (2.0 * Ptl) - cnr - cnz - cnh;
                const double cf2 = welz / (2.0 * Ptl) + cnr + cnz + cnh;
                X(F->Cn, i, j, k) = (cf2 * X(Cp, i, j, k) - ULT - RHS2 + RHS1) / cf1;
            }
}

/* temperature, sect360v2mg.f:692-831: i = 1..nr, j = 0..nz+1 and
 * k = 0..np+1, the barrier planes k = 0, np+1 insulated like the end walls
 * j = 0, nz+1.  On the row k = np the reference assigns its one-sided phi
 * derivative to thz (sect360v2mg.f:787), so that row advects with the
 * phi-derivative in place of thz and with the thh left over from the row
 * k = np-1 of the same (i, j): the loop below keeps that. */
void vs_temperature(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp, dt = P->dt;
    const double Ptl = P->Ptl;
    const double *th = F->th, *thp = F->thp, *uel = F->uel, *vel = F->vel, *wel = F->wel;
    for (int i = 1; i <= nr; ++i)
        for (int j = 0; j <= nz + 1; ++j) {
            double thh = 0.0;   /* carried from one k to the next, as the reference's local */
            for (int k = 0; k <= np + 1; ++k) {
                const int im = i - 1, jm = j - 1, km = k - 1, ka = k + 1;
                const double r = P->r0 + ((double)i - 0.5) * dr;
                double u, v, w;
                if (j == 0 || j == nz + 1 || k == 0 || k == np + 1) {
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
                double thhha, cnp;
                if (k == 1) {
                    thhha = (3.2 * X(th, i, j, 0) + 2.0 * X(th, i, j, 2) - 0.2 * X(th, i, j, 3)) / (dp * dp);
                    thh = (-4.0 * X(th, i, j, 0) / 3.0 + X(th, i, j, 1) + X(th, i, j, 2) / 3.0) / dp;
                    cnp = -2.5 / (dp * dp) / (r * r);
                } else if (k == np) {
                    thhha = (-0.2 * X(th, i, j, np - 2) + 2.0 * X(th, i, j, np - 1) + 3.2 * X(th, i, j, np + 1)) / (dp * dp);
                    thz = (-X(th, i, j, np - 1) / 3.0 - X(th, i, j, np) + 4.0 * X(th, i, j, np + 1) / 3.0) / dp;
                    cnp = -2.5 / (dp * dp) / (r * r);
                } else if (k == 0) {
                    thhha = 8.0 * X(th, i, j, 1) / (dp * dp);
                    thh = 0.0;
                    cnp = -4.0 / (dp * dp) / (r * r);
                } else if (k == np + 1) {
                    thhha = 8.0 * X(th, i, j, np) / (dp * dp);
                    thh = 0.0;
                    cnp = -4.0 / (dp * dp) / (r * r);
                } else {
                    thhha = (X(th, i, j, ka) + X(th, i, j, km)) / (dp * dp);
                    thh = (X(th, i, j, ka) - X(th, i, j, km)) / (2.0 * dp);
                    const double Pe = fabs(v * dp * r / Ptl);
                    if (Pe > 2) {
                        if (v > 0) thh = (X(th, i, j, k) - X(th, i, j, km)) / dp;
                        else if (v < 0) thh = (X(th, i, j, ka) - X(th, i, j, k)) / dp;
                    }
                    cnp = -1.0 / (dp * dp) / (r * r);
                }
                const double RHS1 = thrra + thra / r + thhha / (r * r) + thzza + X(th, i, j, k) / dt;
                const double ULT = u * thr + v * thh / r + w * thz;
                const double cf1 = 1.0 / dt - cnr - cnz - cnp;
                const double cf2 = cnr + cnz + cnp;
                X(F->thn, i, j, k) = (cf2 * X(thp, i, j, k) - ULT + RHS1) / cf1;
            }
        }
}

/* heat, sect360v2mg.f:840-881: the inner-wall flux over z (weights 1/240,
 * 1/80, 1/60) summed over every plane k = 0..np+1 times dp */
double vs_heat(const v3_params *P, const v3_fields *F)
{
    const int nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp;
    const double *th = F->th;
    double trans1 = 0.0;
    for (int k = 0; k <= np + 1; ++k) {
        double trans = dz * (-184.0 * X(th, 0, 0, k) + 225.0 * X(th, 1, 0, k) - 50.0 * X(th, 2, 0, k) +
                             9.0 * X(th, 3, 0, k)) / (240.0 * dr);
        trans = trans + dz * (-184.0 * X(th, 0, 1, k) + 225.0 * X(th, 1, 1, k) - 50.0 * X(th, 2, 1, k) +
                              9.0 * X(th, 3, 1, k)) / (80.0 * dr);
        for (int j = 2; j <= nz - 1; ++j)
            trans = trans + dz * (-184.0 * X(th, 0, j, k) + 225.0 * X(th, 1, j, k) - 50.0 * X(th, 2, j, k) +
                                  9.0 * X(th, 3, j, k)) / (60.0 * dr);
        trans = trans + dz * (-184.0 * X(th, 0, nz, k) + 225.0 * X(th, 1, nz, k) - 50.0 * X(th, 2, nz, k) +
                              9.0 * X(th, 3, nz, k)) / (80.0 * dr);
        trans = trans + dz * (-184.0 * X(th, 0, nz + 1, k) + 225.0 * X(th, 1, nz + 1, k) -
                              50.0 * X(th, 2, nz + 1, k) + 9.0 * X(th, 3, nz + 1, k)) / (240.0 * dr);
        trans1 = trans1 + trans * dp;
    }
    return trans1;
}

/* update, sect360v2mg.f:936-996: wall vorticity on the end walls, the
 * cylinders and the barrier's faces (sect360v2mg.f:964-974), wall
 * temperatures, then the time-level rotation over k = 0..np+1 */
void vs_update(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp;
    const double *uel = F->uel, *vel = F->vel, *wel = F->wel;
    for (int i = 1; i <= nr; ++i)
        for (int k = 1; k <= np - 1; ++k) {
            X(F->An, i, 0, k) = -(3.0 * X(vel, i, 1, k) - X(vel, i, 2, k) / 3.0) / dz;
            X(F->An, i, nz, k) = -(X(vel, i, nz - 1, k) / 3.0 - 3.0 * X(vel, i, nz, k)) / dz;
        }
    for (int i = 1; i <= nr - 1; ++i)
        for (int k = 1; k <= np; ++k) {
            X(F->Bn, i, 0, k) = 2.0 * X(uel, i, 1, k) / dz;
            X(F->Bn, i, nz, k) = -2.0 * X(uel, i, nz, k) / dz;
        }
    for (int j = 1; j <= nz - 1; ++j)
        for (int k = 1; k <= np; ++k) {
            X(F->Bn, 0, j, k) = -2.0 * X(wel, 1, j, k) / dr;
            X(F->Bn, nr, j, k) = 2.0 * X(wel, nr, j, k) / dr;
        }
    for (int j = 1; j <= nz; ++j)
        for (int k = 1; k <= np - 1; ++k) {
            X(F->Cn, 0, j, k) = (3.0 * X(vel, 1, j, k) - X(vel, 2, j, k) / 3.0) / dr;
            X(F->Cn, nr, j, k) = (X(vel, nr - 1, j, k) / 3.0 - 3.0 * X(vel, nr, j, k)) / dr;
        }
    for (int i = 1; i <= nr; ++i)
        for (int j = 1; j <= nz; ++j) {
            X(F->An, i, j, 0) = 1.0 / (P->r0 + ((double)i - 0.5) * dr) * (2.0 * X(wel, i, j, 1)) / dp;
            X(F->An, i, j, np) = 1.0 / (P->r0 + ((double)i - 0.5) * dr) * (-2.0 * X(wel, i, j, np)) / dp;
        }
    for (int i = 1; i <= nr - 1; ++i)
        for (int j = 1; j <= nz; ++j) {
            X(F->Cn, i, j, 0) = 1.0 / (P->r0 + ((double)i - 0.5) * dr) * (-2.0 * X(uel, i, j, 1)) / dp;
            X(F->Cn, i, j, np) = 1.0 / (P->r0 + ((double)i - 0.5) * dr) * (2.0 * X(uel, i, j, np)) / dp;
        }
    for (int j = 0; j <= nz + 1; ++j)
        for (int k = 0; k <= np + 1; ++k) {
            X(F->thn, 0, j, k) = 0.0;
            X(F->thn, nr + 1, j, k) = 1.0;
        }
    for (int k = 0; k <= np + 1; ++k)
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

/* rhsw, sect360v2mg.f:999-1031 */
void vs_rhsw(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dp = P->dp;
    const double *An = F->An, *Bn = F->Bn;
    for (int k = 1; k <= np; ++k)
        for (int j = 1; j <= nz - 1; ++j)
            for (int i = 1; i <= nr; ++i) {
                const int km = k - 1;
                const double r = P->r0 + ((double)i - 0.5) * dr;
                const double Ah = (X(An, i, j, k) - X(An, i, j, km)) / dp;
                const double Br = (X(Bn, i, j, k) - X(Bn, i - 1, j, k)) / dr;
                double v;
                if (i == 1) v = -X(Bn, 1, j, k) / dr - 0.5 * X(Bn, 1, j, k) / r;
                else if (i == nr) v = X(Bn, nr - 1, j, k) / dr - 0.5 * X(Bn, nr - 1, j, k) / r;
                else v = -Br - 0.5 * (X(Bn, i, j, k) + X(Bn, i - 1, j, k)) / r;
                if (k == 1) v = v + X(An, i, j, k) / r / dp;
                else if (k == np) v = v - X(An, i, j, km) / r / dp;
                else v = v + Ah / r;
                X(F->wrh, i, j, k) = v;
            }
}

/* rhsv, sect360v2mg.f:1034-1063 */
void vs_rhsv(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz;
    const double *An = F->An, *Cn = F->Cn;
    for (int k = 1; k <= np - 1; ++k)
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

/* rhsu, sect360v2mg.f:1066-1100 (uses the new w of poisson1) */
void vs_rhsu(const v3_params *P, v3_fields *F)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    const double dr = P->dr, dz = P->dz, dp = P->dp;
    const double *Bn = F->Bn, *Cn = F->Cn, *wel = F->wel;
    for (int k = 1; k <= np; ++k)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nr - 1; ++i) {
                const int km = k - 1;
                const double r = P->r0 + (double)i * dr;
                const double Bz = (X(Bn, i, j, k) - X(Bn, i, j - 1, k)) / dz;
                const double welz = 0.5 * (X(wel, i, j, k) + X(wel, i + 1, j, k) - X(wel, i, j - 1, k) -
                                           X(wel, i + 1, j - 1, k)) / dz;
                const double Ch = (X(Cn, i, j, k) - X(Cn, i, j, km)) / dp;
                double v;
                if (j == 1) v = X(Bn, i, 1, k) / dz - 2.0 * welz / r;
                else if (j == nz) v = -X(Bn, i, nz - 1, k) / dz - 2.0 * welz / r;
                else v = Bz - 2.0 * welz / r;
                if (k == 1) v = v - X(Cn, i, j, k) / r / dp;
                else if (k == np) v = v + X(Cn, i, j, k - 1) / r / dp;
                else v = v - Ch / r;
                X(F->urh, i, j, k) = v;
            }
}

void vs_ranges(const v3_params *P, int which, int *i1, int *j1, int *k1)
{
    if (which == 1) { *i1 = P->nr; *j1 = P->nz - 1; *k1 = P->np; }
    else if (which == 2) { *i1 = P->nr; *j1 = P->nz; *k1 = P->np - 1; }
    else { *i1 = P->nr - 1; *j1 = P->nz; *k1 = P->np; }
}

/* value of psi at (i, j, k) as the residual reads it: the unknowns, and zero
 * on the walls the operator reaches */
static double pv(const v3_params *P, int which, const double *ps, int i, int j, int k)
{
    int i1, j1, k1;
    vs_ranges(P, which, &i1, &j1, &k1);
    if (i < 1 || i > i1 || j < 1 || j > j1 || k < 1 || k > k1) return 0.0;
    return X(ps, i, j, k);
}

/* L psi at (i, j, k) for operator `which` at the finest level, in the
 * reference residual's operation order */
static double apply_L(const v3_params *P, int which, const double *ps, int i, int j, int k)
{
    const int n1 = P->nr, n2 = P->nz, n3 = P->np;
    const double hr = P->dr, hz = P->dz, hp = P->dp;
    const double hhr = hr * hr, hhz = hz * hz, hhp = hp * hp;
#define Q(a, b, c) pv(P, which, ps, a, b, c)
    double temp;
    if (which == 1) {   /* sect360v2mg.f:1819-1848 */
        const double rval = P->r0 + ((double)i - 0.5) * hr;
        if (i == 1)
            temp = (Q(i, j + 1, k) + Q(i, j - 1, k)) / hhz - Q(i, j, k) * 2.0 / hhz + (-Q(i, j, k) + Q(i + 1, j, k)) / hhr +
                   0.5 * (-Q(i, j, k) + Q(i + 1, j, k)) / hr / rval;
        else if (i == n1)
            temp = (Q(i, j + 1, k) + Q(i, j - 1, k)) / hhz - Q(i, j, k) * 2.0 / hhz + (-Q(i, j, k) + Q(i - 1, j, k)) / hhr +
                   0.5 * (Q(i, j, k) - Q(i - 1, j, k)) / hr / rval;
        else
            temp = (Q(i, j + 1, k) + Q(i, j - 1, k)) / hhz - Q(i, j, k) * 2.0 / hhz +
                   (Q(i - 1, j, k) - 2.0 * Q(i, j, k) + Q(i + 1, j, k)) / hhr +
                   0.5 * (Q(i + 1, j, k) - Q(i - 1, j, k)) / hr / rval;
        if (k == 1) temp = temp + (-Q(i, j, k) + Q(i, j, k + 1)) / hhp / (rval * rval);
        else if (k == n3) temp = temp + (Q(i, j, k - 1) - Q(i, j, k)) / hhp / (rval * rval);
        else
            temp = temp + (Q(i, j, k - 1) + Q(i, j, k + 1)) / hhp / (rval * rval) - Q(i, j, k) * 2.0 / hhp / (rval * rval);
    } else if (which == 2) {   /* sect360v2mg.f:2036-2064 */
        const double rval = P->r0 + ((double)i - 0.5) * hr;
        temp = (Q(i, j, k + 1) - 2.0 * Q(i, j, k) + Q(i, j, k - 1)) / hhp / (rval * rval) + Q(i, j, k) / (rval * rval);
        if (i == 1)
            temp = temp + (-2.0 * Q(i, j, k) + 5.0 / 3.0 * Q(i + 1, j, k) - 0.2 * Q(i + 2, j, k)) / hhr +
                   4.0 / 3.0 * Q(i + 1, j, k) / hr / rval;
        else if (i == n1)
            temp = temp + (-2.0 * Q(i, j, k) + 5.0 / 3.0 * Q(i - 1, j, k) - 0.2 * Q(i - 2, j, k)) / hhr -
                   4.0 / 3.0 * Q(i - 1, j, k) / hr / rval;
        else
            temp = temp + (Q(i - 1, j, k) - 2.0 * Q(i, j, k) + Q(i + 1, j, k)) / hhr +
                   1.5 * (Q(i + 1, j, k) - Q(i - 1, j, k)) / hr / rval;
        if (j == 1) temp = temp + (-2.0 * Q(i, j, k) + 5.0 / 3.0 * Q(i, j + 1, k) - 0.2 * Q(i, j + 2, k)) / hhz;
        else if (j == n2) temp = temp + (-2.0 * Q(i, j, k) + 5.0 / 3.0 * Q(i, j - 1, k) - 0.2 * Q(i, j - 2, k)) / hhz;
        else temp = temp + (Q(i, j + 1, k) - 2.0 * Q(i, j, k) + Q(i, j - 1, k)) / hhz;
    } else {   /* sect360v2mg.f:2235-2257 */
        const double rval = (double)i * hr + P->r0;
        temp = (Q(i + 1, j, k) - Q(i, j, k) * 2.0 + Q(i - 1, j, k)) / hhr + 1.5 * (Q(i + 1, j, k) - Q(i - 1, j, k)) / hr / rval +
               Q(i, j, k) / (rval * rval);
        if (j == 1) temp = temp + (-Q(i, j, k) + Q(i, j + 1, k)) / hhz;
        else if (j == n2) temp = temp + (-Q(i, j, k) + Q(i, j - 1, k)) / hhz;
        else temp = temp + (Q(i, j + 1, k) - 2.0 * Q(i, j, k) + Q(i, j - 1, k)) / hhz;
        if (k == 1) temp = temp + (-Q(i, j, k) + Q(i, j, k + 1)) / (rval * rval) / hhp;
        else if (k == n3) temp = temp + (-Q(i, j, k) + Q(i, j, k - 1)) / (rval * rval) / hhp;
        else temp = temp + (-2.0 * Q(i, j, k) + Q(i, j, k - 1) + Q(i, j, k + 1)) / (rval * rval) / hhp;
    }
#undef Q
    return temp;
}

double vs_residual(const v3_params *P, int which, const double *psi, const double *rhs, double *res)
{
    int i1, j1, k1;
    vs_ranges(P, which, &i1, &j1, &k1);
    if (res) memset(res, 0, (size_t)P->ldi * P->ldj * (P->np + 2) * sizeof(double));
    double m = 0.0;
    for (int k = 1; k <= k1; ++k)
        for (int j = 1; j <= j1; ++j)
            for (int i = 1; i <= i1; ++i) {
                const double t = X(rhs, i, j, k) - apply_L(P, which, psi, i, j, k);
                if (res) X(res, i, j, k) = t;
                if (fabs(t) > m) m = fabs(t);
            }
    return m;
}

/* dense LU with partial pivoting, row-major n x n, in place */
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
            if (f != 0.0)
                for (int q = c + 1; q < n; ++q) a[(size_t)r * n + q] -= f * a[(size_t)c * n + q];
        }
    }
    return 0;
}

int vs_poisson_exact(const v3_params *P, int which, double *x, const double *rhs)
{
    int i1, j1, k1;
    vs_ranges(P, which, &i1, &j1, &k1);
    const int m = i1 * j1 * k1;
    const size_t nf = (size_t)P->ldi * P->ldj * (P->np + 2);
    double *a = (double *)malloc((size_t)m * m * sizeof(double));
    double *e = (double *)calloc(nf, sizeof(double));
    double *b = (double *)malloc((size_t)m * sizeof(double));
    int *piv = (int *)malloc((size_t)m * sizeof(int));
    if (!a || !e || !b || !piv) { free(a); free(e); free(b); free(piv); return -1; }
#define U(i, j, k) (((size_t)((k) - 1) * j1 + (size_t)((j) - 1)) * i1 + (size_t)((i) - 1))
    /* column c of the matrix: L applied to the unit vector of unknown c,
     * evaluated on the unknowns its stencil reaches */
    for (int kc = 1; kc <= k1; ++kc)
        for (int jc = 1; jc <= j1; ++jc)
            for (int ic = 1; ic <= i1; ++ic) {
                const size_t c = U(ic, jc, kc);
                X(e, ic, jc, kc) = 1.0;
                for (int k = 1; k <= k1; ++k)
                    for (int j = 1; j <= j1; ++j)
                        for (int i = 1; i <= i1; ++i) {
                            double v = 0.0;
                            if (abs(i - ic) <= 2 && abs(j - jc) <= 2 && abs(k - kc) <= 1) v = apply_L(P, which, e, i, j, k);
                            a[U(i, j, k) * m + c] = v;
                        }
                X(e, ic, jc, kc) = 0.0;
            }
    int st = lu_factor(a, piv, m);
    if (!st) {
        for (int k = 1; k <= k1; ++k)
            for (int j = 1; j <= j1; ++j)
                for (int i = 1; i <= i1; ++i) b[U(i, j, k)] = X(rhs, i, j, k);
        for (int c = 0; c < m; ++c) {
            const int p = piv[c];
            if (p != c) { const double t = b[c]; b[c] = b[p]; b[p] = t; }
            for (int r = c + 1; r < m; ++r) b[r] -= a[(size_t)r * m + c] * b[c];
        }
        for (int c = m - 1; c >= 0; --c) {
            double s = b[c];
            for (int q = c + 1; q < m; ++q) s -= a[(size_t)c * m + q] * b[q];
            b[c] = s / a[(size_t)c * m + c];
        }
        for (int k = 1; k <= k1; ++k)
            for (int j = 1; j <= j1; ++j)
                for (int i = 1; i <= i1; ++i) X(x, i, j, k) = b[U(i, j, k)];
    }
#undef U
    free(a); free(e); free(b); free(piv);
    return st;
}

int vs_poisson(const v3_params *P, int which, double *psi, double *psip, const double *rhs)
{
    const int nr = P->nr, nz = P->nz, np = P->np;
    /* psip <- psi on the range the reference copies (sect360v2mg.f:1427-1432,
     * 1519-1524, 1619-1624) */
    const int ie = (which == 3) ? nr : nr + 1, je = (which == 1) ? nz : nz + 1, ke = (which == 2) ? np : np + 1;
    for (int k = 0; k <= ke; ++k)
        for (int j = 0; j <= je; ++j)
            for (int i = 0; i <= ie; ++i) X(psip, i, j, k) = X(psi, i, j, k);
    if (vs_poisson_exact(P, which, psi, rhs)) return -1;
    if (which == 1) {   /* sect360v2mg.f:1439-1457 */
        for (int k = 0; k <= np + 1; ++k)
            for (int i = 0; i <= nr + 1; ++i) {
                X(psi, i, 0, k) = 0.0;
                X(psi, i, nz, k) = 0.0;
            }
        for (int j = 1; j <= nz - 1; ++j)
            for (int k = 0; k <= np + 1; ++k) {
                const int kk = (k == 0) ? 1 : (k == np + 1) ? np : k;
                X(psi, 0, j, k) = -X(psi, 1, j, kk);
                X(psi, nr + 1, j, k) = -X(psi, nr, j, kk);
            }
        for (int i = 1; i <= nr; ++i)
            for (int j = 1; j <= nz - 1; ++j) {
                X(psi, i, j, 0) = -X(psi, i, j, 1);
                X(psi, i, j, np + 1) = -X(psi, i, j, np);
            }
    } else if (which == 2) {   /* sect360v2mg.f:1530-1547 */
        for (int j = 0; j <= nz + 1; ++j)
            for (int i = 0; i <= nr + 1; ++i) {
                X(psi, i, j, 0) = 0.0;
                X(psi, i, j, np) = 0.0;
            }
        for (int j = 0; j <= nz + 1; ++j)
            for (int k = 1; k <= np - 1; ++k) {
                X(psi, 0, j, k) = 0.0;
                X(psi, nr + 1, j, k) = 0.0;
            }
        for (int i = 1; i <= nr; ++i)
            for (int k = 1; k <= np - 1; ++k) {
                X(psi, i, 0, k) = 0.0;
                X(psi, i, nz + 1, k) = 0.0;
            }
    } else {   /* sect360v2mg.f:1630-1654 */
        for (int j = 0; j <= nz + 1; ++j)
            for (int k = 0; k <= np + 1; ++k) {
                X(psi, 0, j, k) = 0.0;
                X(psi, nr, j, k) = 0.0;
            }
        for (int i = 1; i <= nr - 1; ++i)
            for (int k = 0; k <= np + 1; ++k) {
                const int kk = (k == 0) ? 1 : (k == np + 1) ? np : k;
                X(psi, i, 0, k) = -X(psi, i, 1, kk);
                X(psi, i, nz + 1, k) = -X(psi, i, nz, kk);
            }
        for (int i = 1; i <= nr - 1; ++i)
            for (int j = 1; j <= nz; ++j) {
                X(psi, i, j, 0) = -X(psi, i, j, 1);
                X(psi, i, j, np + 1) = -X(psi, i, j, np);
            }
    }
    return 0;
}
