// vv3.hip -- the 3D rotating-annulus time step of rot3dmg.f on gfx950.
//
// Explicit part: one thread per point for Avorticity / Bvorticity /
// Cvorticity / temperature, the Poisson right-hand sides and the wall
// update, evaluating the reference's expressions in its operation order
// (the library is compiled with -ffp-contract=off), so the results equal
// the reference bit for bit.
//
// The three Poisson solves are exact separable solves (vv3_poisson.hip).

#include "vv3_internal.h"

#include <cmath>
#include <vector>

struct vv3_run {
    vv3_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    int nr = 0, nz = 0, np = 0, ldi = 0, ldj = 0;
    double dr = 0, dz = 0, dp = 0;
    size_t nf = 0;                       // doubles per field
    double *f[VV3_NFIELD] = {nullptr};   // device fields (rotated by pointer swap)
    vv3::Poisson3 pois;                  // the three Poisson operators
    double *heat = nullptr;              // per-k partial fluxes (np+1) and the total
};

namespace vv3 {

namespace {

constexpr int WG = 256;

#define X(a, i, j, k) (a)[((size_t)(k) * g.ldj + (size_t)(j)) * g.ldi + (size_t)(i)]

struct Fields {
    const double *A, *B, *C, *Ap, *Bp, *Cp, *th, *thp, *uel, *vel, *wel;
    double *An, *Bn, *Cn, *thn;
};

// Avorticity at (i, j, k), rot3dmg.f:359-465
__device__ double avort(const Geo &g, const Fields &F, int i, int j, int k)
{
    const double *A = F.A, *B = F.B, *C = F.C, *th = F.th, *uel = F.uel, *vel = F.vel, *wel = F.wel;
    const int nr = g.nr, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, dt = g.dt, Ptl = g.Ptl, Ra = g.Ra, Re = g.Re, Gr = g.Gr;
    const int im = i - 1, jp = j + 1;
    const int km = (k == 0) ? np - 1 : k - 1;
    const int ka = (k == np) ? 1 : k + 1;
    const double r = g.r0 + ((double)i - 0.5) * dr;
    const double u = (X(uel, i, j, k) + X(uel, i, jp, k) + X(uel, i, j, ka) + X(uel, i, jp, ka) + X(uel, im, j, k) +
                      X(uel, im, jp, k) + X(uel, im, j, ka) + X(uel, im, jp, ka)) * 0.125;
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
        if (fabs(u * dr / Ptl) > 2) {          // hybrid upwind, rot3dmg.f:397-406
            if (u > 0) Ar = (X(A, i, j, k) - X(A, i - 1, j, k)) / dr;
            else if (u < 0) Ar = (X(A, i + 1, j, k) - X(A, i, j, k)) / dr;
        }
    }
    const double uelr = ((X(uel, i, j, k) + X(uel, i, jp, k) + X(uel, i, j, ka) + X(uel, i, jp, ka)) * 0.25 -
                         (X(uel, im, j, k) + X(uel, im, jp, k) + X(uel, im, j, ka) + X(uel, im, jp, ka)) * 0.25) / dr;
    const double Azza = (X(A, i, j - 1, k) + X(A, i, j + 1, k)) / (dz * dz);
    double Az = (X(A, i, j + 1, k) - X(A, i, j - 1, k)) * 0.5 / dz;
    if (fabs(w * dz / Ptl) > 2) {
        if (w > 0) Az = (X(A, i, j, k) - X(A, i, j - 1, k)) / dz;
        else if (w < 0) Az = (X(A, i, j + 1, k) - X(A, i, j, k)) / dz;
    }
    const double uelz = ((X(uel, i, jp, k) + X(uel, i, jp, ka) + X(uel, im, jp, k) + X(uel, im, jp, ka)) * 0.25 -
                         (X(uel, i, j, k) + X(uel, i, j, ka) + X(uel, im, j, k) + X(uel, im, j, ka)) * 0.25) / dz;
    const double thz = (X(th, i, jp, k) + X(th, i, jp, ka) - X(th, i, j, k) - X(th, i, j, ka)) * 0.5 / dz;
    const double thh = (X(th, i, j, ka) + X(th, i, jp, ka) - X(th, i, j, k) - X(th, i, jp, k)) * 0.5 / dp;
    const double Ahha = (X(A, i, j, ka) + X(A, i, j, km)) / (dp * dp);
    double Ah = (X(A, i, j, ka) - X(A, i, j, km)) * 0.5 / dp;
    if (fabs(v * dp * r / Ptl) > 2) {
        if (v > 0) Ah = (X(A, i, j, k) - X(A, i, j, km)) / dp;
        else if (v < 0) Ah = (X(A, i, j, ka) - X(A, i, j, k)) / dp;
    }
    const double Bh = (X(B, i, j, ka) + X(B, im, j, ka) - X(B, i, j, k) - X(B, im, j, k)) * 0.5 / dp;
    const double uelh = ((X(uel, i, j, ka) + X(uel, i, jp, ka) + X(uel, im, j, ka) + X(uel, im, jp, ka)) * 0.25 -
                         (X(uel, i, j, k) + X(uel, i, jp, k) + X(uel, im, j, k) + X(uel, im, jp, k)) * 0.25) / dp;
    const double cnz = -1.0 / (dz * dz);
    const double cnh = -1.0 / (dp * dp) / (r * r);
    const double RHS1 = Arra + Ara / r + Ahha / (r * r) + Azza - 2.0 * Bh / (r * r) + X(A, i, j, k) / (Ptl * dt);
    const double ULT = (u * Ar + v * Ah / r + w * Az - uelh * BB / r - uelz * CC) / Ptl;
    const double RHS2 = -2.0 * Re * uelz + 2.0 * Ra * (tem * uelz + thz * u) / Re / Ptl - Gr * Ptl / r * thh;
    const double cf1 = 1.0 / (Ptl * dt) - uelr / (2.0 * Ptl) + 1.0 / (2.0 * (r * r)) - cnr - cnz - cnh;
    const double cf2 = uelr / (2.0 * Ptl) - 1.0 / (2.0 * (r * r)) + cnr + cnz + cnh;
    return (cf2 * X(F.Ap, i, j, k) - ULT - RHS2 + RHS1) / cf1;
}

// Bvorticity at (i, j, k), rot3dmg.f:495-590
__device__ double bvort(const Geo &g, const Fields &F, int i, int j, int k)
{
    const double *A = F.A, *B = F.B, *C = F.C, *th = F.th, *uel = F.uel, *vel = F.vel, *wel = F.wel;
    const int np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, dt = g.dt, Ptl = g.Ptl, Ra = g.Ra, Re = g.Re, Gr = g.Gr;
    const int ip = i + 1, jp = j + 1;
    const int km = (k == 0) ? np - 1 : k - 1;
    const int ka = (k == np) ? 1 : k + 1;
    const double r = g.r0 + (double)i * dr;
    const double u = (X(uel, i, j, k) + X(uel, i, jp, k)) * 0.5;
    const double v = (X(vel, i, j, km) + X(vel, i, jp, km) + X(vel, i, j, k) + X(vel, i, jp, k) + X(vel, ip, j, km) +
                      X(vel, ip, jp, km) + X(vel, ip, j, k) + X(vel, ip, jp, k)) * 0.125;
    const double w = (X(wel, i, j, k) + X(wel, ip, j, k)) * 0.5;
    const double tem = (X(th, i, jp, k) + X(th, ip, jp, k) + X(th, i, j, k) + X(th, ip, j, k)) * 0.25;
    const double AA = (X(A, i, j, k) + X(A, ip, j, k) + X(A, i, j, km) + X(A, ip, j, km)) * 0.25;
    const double CC = (X(C, i, j, k) + X(C, i, jp, k) + X(C, i, j, km) + X(C, i, jp, km)) * 0.25;
    const double Brra = (X(B, i - 1, j, k) + X(B, i + 1, j, k)) / (dr * dr);
    const double Bra = (X(B, i + 1, j, k) - X(B, i - 1, j, k)) * 0.5 / dr;
    double Br = Bra;
    if (fabs(u * dr / Ptl) > 2) {
        if (u > 0) Br = (X(B, i, j, k) - X(B, i - 1, j, k)) / dr;
        else if (u < 0) Br = (X(B, i + 1, j, k) - X(B, i, j, k)) / dr;
    }
    const double velr = ((X(vel, ip, j, km) + X(vel, ip, jp, km) + X(vel, ip, j, k) + X(vel, ip, jp, k)) * 0.25 -
                         (X(vel, i, j, km) + X(vel, i, jp, km) + X(vel, i, j, k) + X(vel, i, jp, k)) * 0.25) / dr;
    const double Bzza = (X(B, i, j - 1, k) + X(B, i, j + 1, k)) / (dz * dz);
    double Bz = (X(B, i, j + 1, k) - X(B, i, j - 1, k)) * 0.5 / dz;
    if (fabs(w * dz / Ptl) > 2) {
        if (w > 0) Bz = (X(B, i, j, k) - X(B, i, j - 1, k)) / dz;
        else if (w < 0) Bz = (X(B, i, j + 1, k) - X(B, i, j, k)) / dz;
    }
    const double velz = ((X(vel, i, jp, km) + X(vel, ip, jp, km) + X(vel, i, jp, k) + X(vel, ip, jp, k)) * 0.25 -
                         (X(vel, i, j, km) + X(vel, ip, j, km) + X(vel, i, j, k) + X(vel, ip, j, k)) * 0.25) / dz;
    const double thz = (X(th, i, jp, k) + X(th, ip, jp, k) - X(th, i, j, k) - X(th, ip, j, k)) * 0.5 / dz;
    const double thr = (X(th, ip, jp, k) + X(th, ip, j, k) - X(th, i, jp, k) - X(th, i, j, k)) * 0.5 / dr;
    const double Bhha = (X(B, i, j, ka) + X(B, i, j, km)) / (dp * dp);
    double Bh = (X(B, i, j, ka) - X(B, i, j, km)) * 0.5 / dp;
    if (fabs(v * dp * r / Ptl) > 2) {
        if (v > 0) Bh = (X(B, i, j, k) - X(B, i, j, km)) / dp;
        else if (v < 0) Bh = (X(B, i, j, ka) - X(B, i, j, k)) / dp;
    }
    const double Ah = (X(A, i, j, k) + X(A, ip, j, k) - X(A, i, j, km) - X(A, ip, j, km)) * 0.5 / dp;
    const double velh = ((X(vel, i, j, k) + X(vel, ip, j, k) + X(vel, i, jp, k) + X(vel, ip, jp, k)) * 0.25 -
                         (X(vel, i, j, km) + X(vel, ip, j, km) + X(vel, i, jp, km) + X(vel, ip, jp, km)) * 0.25) / dp;
    const double cnr = -1.0 / (dr * dr);
    const double cnz = -1.0 / (dz * dz);
    const double cnh = -1.0 / (dp * dp) / (r * r);
    const double RHS1 = Brra + Bra / r + Bhha / (r * r) + Bzza + 2.0 * Ah / (r * r) + X(B, i, j, k) / (Ptl * dt);
    const double ULT = (u * Br + v * Bh / r + w * Bz - (velr - v / r) * AA - velz * CC) / Ptl;
    const double RHS2 = -2.0 * Re * velz + r * Ra * thz + 2.0 * Ra * (tem * velz + thz * v) / Re / Ptl + Gr * Ptl * thr;
    const double cf1 = 1.0 / (Ptl * dt) - (u / r + velh / r) / (2.0 * Ptl) + 1.0 / (2.0 * (r * r)) - cnr - cnz - cnh;
    const double cf2 = (u / r + velh / r) / (2.0 * Ptl) - 1.0 / (2.0 * (r * r)) + cnr + cnz + cnh;
    return (cf2 * X(F.Bp, i, j, k) - ULT - RHS2 + RHS1) / cf1;
}

// Cvorticity at (i, j, k), rot3dmg.f:619-725
__device__ double cvort(const Geo &g, const Fields &F, int i, int j, int k)
{
    const double *A = F.A, *B = F.B, *C = F.C, *th = F.th, *uel = F.uel, *vel = F.vel, *wel = F.wel;
    const int nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, dt = g.dt, Ptl = g.Ptl, Ra = g.Ra, Re = g.Re;
    const int ip = i + 1, jm = j - 1;
    const int km = (k == 0) ? np - 1 : k - 1;
    const int ka = (k == np) ? 1 : k + 1;
    const double r = g.r0 + (double)i * dr;
    const double u = (X(uel, i, j, ka) + X(uel, i, j, k)) * 0.5;
    const double v = (X(vel, i, j, k) + X(vel, ip, j, k)) * 0.5;
    const double w = (X(wel, i, j, k) + X(wel, ip, j, k) + X(wel, i, j, ka) + X(wel, ip, j, ka) + X(wel, i, jm, k) +
                      X(wel, ip, jm, k) + X(wel, i, jm, ka) + X(wel, ip, jm, ka)) * 0.125;
    const double AA = (X(A, i, j, k) + X(A, i, jm, k) + X(A, ip, j, k) + X(A, ip, jm, k)) * 0.25;
    const double BB = (X(B, i, j, k) + X(B, i, jm, k) + X(B, i, j, ka) + X(B, i, jm, ka)) * 0.25;
    const double tem = (X(th, i, j, k) + X(th, ip, j, k) + X(th, ip, j, ka) + X(th, i, j, ka)) * 0.25;
    const double Crra = (X(C, i - 1, j, k) + X(C, i + 1, j, k)) / (dr * dr);
    const double Cra = (X(C, i + 1, j, k) - X(C, i - 1, j, k)) * 0.5 / dr;
    double Cr = Cra;
    if (fabs(u * dr / Ptl) > 2) {
        if (u > 0) Cr = (X(C, i, j, k) - X(C, i - 1, j, k)) / dr;
        else if (u < 0) Cr = (X(C, i + 1, j, k) - X(C, i, j, k)) / dr;
    }
    const double welr = ((X(wel, ip, j, k) + X(wel, ip, j, ka) + X(wel, ip, jm, k) + X(wel, ip, jm, ka)) * 0.25 -
                         (X(wel, i, j, k) + X(wel, i, j, ka) + X(wel, i, jm, k) + X(wel, i, jm, ka)) * 0.25) / dr;
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
        if (fabs(w * dz / Ptl) > 2) {
            if (w > 0) Cz = (X(C, i, j, k) - X(C, i, j - 1, k)) / dz;
            else if (w < 0) Cz = (X(C, i, j + 1, k) - X(C, i, j, k)) / dz;
        }
        cnz = -1.0 / (dz * dz);
    }
    const double welz = ((X(wel, i, j, k) + X(wel, ip, j, k) + X(wel, i, j, ka) + X(wel, ip, j, ka)) * 0.25 -
                         (X(wel, i, jm, k) + X(wel, ip, jm, k) + X(wel, i, jm, ka) + X(wel, ip, jm, ka)) * 0.25) / dz;
    const double Chha = (X(C, i, j, ka) + X(C, i, j, km)) / (dp * dp);
    double Ch = (X(C, i, j, ka) - X(C, i, j, km)) * 0.5 / dp;
    if (fabs(v * dp * r / Ptl) > 2) {
        if (v > 0) Ch = (X(C, i, j, k) - X(C, i, j, km)) / dp;
        else if (v < 0) Ch = (X(C, i, j, ka) - X(C, i, j, k)) / dp;
    }
    const double welh = ((X(wel, i, j, ka) + X(wel, ip, j, ka) + X(wel, i, jm, ka) + X(wel, ip, jm, ka)) * 0.25 -
                         (X(wel, i, j, k) + X(wel, ip, j, k) + X(wel, i, jm, k) + X(wel, ip, jm, k)) * 0.25) / dp;
    const double thh = (X(th, i, j, ka) + X(th, ip, j, ka) - X(th, i, j, k) - X(th, ip, j, k)) * 0.5 / dp;
    const double cnr = -1.0 / (dr * dr);
    const double cnh = -1.0 / (dp * dp) / (r * r);
    const double RHS1 = Crra + Cra / r + Chha / (r * r) + Czza + X(C, i, j, k) / (Ptl * dt);
    const double ULT = (u * Cr + v * Ch / r + w * Cz - welr * AA - welh * BB / r) / Ptl;
    const double RHS2 = -2.0 * Re * welz - Ra * thh + 2.0 * Ra * (tem * welz - thr * u - thh * v / r) / Re / Ptl;
    const double cf1 = 1.0 / (Ptl * dt) - welz / (2.0 * Ptl) - cnr - cnz - cnh;
    const double cf2 = welz / (2.0 * Ptl) + cnr + cnz + cnh;
    return (cf2 * X(F.Cp, i, j, k) - ULT - RHS2 + RHS1) / cf1;
}

// temperature at (i, j, k), rot3dmg.f:754-871
__device__ double temper(const Geo &g, const Fields &F, int i, int j, int k)
{
    const double *th = F.th, *uel = F.uel, *vel = F.vel, *wel = F.wel;
    const int nr = g.nr, nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, dt = g.dt, Ptl = g.Ptl;
    const int im = i - 1, jm = j - 1;
    const int km = (k == 0) ? np - 1 : k - 1;
    const int ka = (k == np) ? 1 : k + 1;
    const double r = g.r0 + ((double)i - 0.5) * dr;
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
        if (fabs(u * dr / Ptl) > 2) {
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
        if (fabs(w * dz / Ptl) > 2) {
            if (w > 0) thz = (X(th, i, j, k) - X(th, i, j - 1, k)) / dz;
            else if (w < 0) thz = (X(th, i, j + 1, k) - X(th, i, j, k)) / dz;
        }
        cnz = -1.0 / (dz * dz);
    }
    const double thhha = (X(th, i, j, ka) + X(th, i, j, km)) / (dp * dp);
    double thh = (X(th, i, j, ka) - X(th, i, j, km)) * 0.5 / dp;
    if (fabs(v * dp * r / Ptl) > 2) {
        if (v > 0) thh = (X(th, i, j, k) - X(th, i, j, km)) / dp;
        else if (v < 0) thh = (X(th, i, j, ka) - X(th, i, j, k)) / dp;
    }
    const double cnp = -1.0 / (dp * dp) / (r * r);
    const double RHS1 = thrra + thra / r + thhha / (r * r) + thzza + X(th, i, j, k) / dt;
    const double ULT = u * thr + v * thh / r + w * thz;
    const double cf1 = 1.0 / dt - cnr - cnz - cnp;
    const double cf2 = cnr + cnz + cnp;
    return (cf2 * X(F.thp, i, j, k) - ULT + RHS1) / cf1;
}

// One launch for A, B, C and T: blockIdx.y = 0..3, a thread per point of
// that routine's loop range.
__global__ void __launch_bounds__(WG) k3_explicit(Geo g, Fields F, int mask)
{
    const int which = blockIdx.y;
    if (!(mask & (1 << which))) return;
    const int nr = g.nr, nz = g.nz;
    int i0 = 1, ni, j0 = 1, nj;
    if (which == 0) { ni = nr; nj = nz - 1; }
    else if (which == 1) { ni = nr - 1; nj = nz - 1; }
    else if (which == 2) { ni = nr - 1; nj = nz; }
    else { ni = nr; j0 = 0; nj = nz + 2; }
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long cnt = (long long)ni * nj * (g.np + 1);
    if (t >= cnt) return;
    const int i = i0 + (int)(t % ni);
    const int j = j0 + (int)((t / ni) % nj);
    const int k = (int)(t / ((long long)ni * nj));
    if (which == 0) X(F.An, i, j, k) = avort(g, F, i, j, k);
    else if (which == 1) X(F.Bn, i, j, k) = bvort(g, F, i, j, k);
    else if (which == 2) X(F.Cn, i, j, k) = cvort(g, F, i, j, k);
    else X(F.thn, i, j, k) = temper(g, F, i, j, k);
}

// heat, rot3dmg.f:882-905: the flux of each azimuth k in parallel (with the
// reference's order of the z sum), then trans1 summed over k in order.
__global__ void __launch_bounds__(WG) k3_heat(Geo g, const double *th, double *part, double *out)
{
    const int nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp;
    for (int k = 1 + threadIdx.x; k <= np; k += blockDim.x) {
        double trans = dz * (-184.0 * X(th, 0, 0, k) + 225.0 * X(th, 1, 0, k) - 50.0 * X(th, 2, 0, k) +
                             9.0 * X(th, 3, 0, k)) / (240.0 * dr);
        trans = trans + dz * (-184.0 * X(th, 0, 1, k) + 225.0 * X(th, 1, 1, k) - 50.0 * X(th, 2, 1, k) +
                              9.0 * X(th, 3, 1, k)) / (80.0 * dr);
        trans = trans + dz * (-184.0 * X(th, 0, nz, k) + 225.0 * X(th, 1, nz, k) - 50.0 * X(th, 2, nz, k) +
                              9.0 * X(th, 3, nz, k)) / (80.0 * dr);
        trans = trans + dz * (-184.0 * X(th, 0, nz + 1, k) + 225.0 * X(th, 1, nz + 1, k) - 50.0 * X(th, 2, nz + 1, k) +
                              9.0 * X(th, 3, nz + 1, k)) / (240.0 * dr);
        for (int j = 2; j <= nz - 1; ++j)
            trans = trans + dz * (-184.0 * X(th, 0, j, k) + 225.0 * X(th, 1, j, k) - 50.0 * X(th, 2, j, k) +
                                  9.0 * X(th, 3, j, k)) / (60.0 * dr);
        part[k] = trans;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double trans1 = 0.0;
        for (int k = 1; k <= np; ++k) trans1 = trans1 + part[k] * dp;
        *out = trans1;
    }
}

struct RhsF {
    const double *An, *Bn, *Cn, *wel;
    double *wrh, *vrh, *urh;
};

// rhsw / rhsv / rhsu, rot3dmg.f:1026-1134 (which = 1 / 2 / 3)
__global__ void __launch_bounds__(WG) k3_rhs(Geo g, RhsF F, int which)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp;
    const int ni = (which == 3) ? nr - 1 : nr;
    const int nj = (which == 1) ? nz - 1 : nz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ni * nj * (np + 1)) return;
    const int i = 1 + (int)(t % ni);
    const int j = 1 + (int)((t / ni) % nj);
    const int k = (int)(t / ((long long)ni * nj));
    const int km = (k == 0) ? np - 1 : k - 1;
    const double *An = F.An, *Bn = F.Bn, *Cn = F.Cn, *wel = F.wel;
    if (which == 1) {
        const double r = g.r0 + ((double)i - 0.5) * dr;
        const double Ah = (X(An, i, j, k) - X(An, i, j, km)) / dp;
        const double Br = (X(Bn, i, j, k) - X(Bn, i - 1, j, k)) / dr;
        double v;
        if (i == 1) v = Ah / r - X(Bn, 1, j, k) / dr - 0.5 * X(Bn, 1, j, k) / r;
        else if (i == nr) v = Ah / r + X(Bn, nr - 1, j, k) / dr - 0.5 * X(Bn, nr - 1, j, k) / r;
        else v = Ah / r - Br - 0.5 * (X(Bn, i, j, k) + X(Bn, i - 1, j, k)) / r;
        X(F.wrh, i, j, k) = v;
    } else if (which == 2) {
        const double r = g.r0 + ((double)i - 0.5) * dr;
        const double Az = (X(An, i, j, k) - X(An, i, j - 1, k)) / dz;
        const double Cr = (X(Cn, i, j, k) - X(Cn, i - 1, j, k)) / dr;
        double v;
        if (i == 1) v = X(Cn, 1, j, k) / dr + X(Cn, 1, j, k) / r;
        else if (i == nr) v = -X(Cn, nr - 1, j, k) / dr + X(Cn, nr - 1, j, k) / r;
        else v = Cr + (X(Cn, i, j, k) + X(Cn, i - 1, j, k)) / r;
        if (j == 1) v = v - X(An, i, 1, k) / dz;
        else if (j == nz) v = v + X(An, i, nz - 1, k) / dz;
        else v = v - Az;
        X(F.vrh, i, j, k) = v;
    } else {
        const double r = g.r0 + (double)i * dr;
        const double Bz = (X(Bn, i, j, k) - X(Bn, i, j - 1, k)) / dz;
        const double welz = 0.5 * (X(wel, i, j, k) + X(wel, i + 1, j, k) - X(wel, i, j - 1, k) - X(wel, i + 1, j - 1, k)) / dz;
        const double Ch = (X(Cn, i, j, k) - X(Cn, i, j, km)) / dp;
        double v;
        if (j == 1) v = X(Bn, i, 1, k) / dz - Ch / r - 2.0 * welz / r;
        else if (j == nz) v = -X(Bn, i, nz - 1, k) / dz - Ch / r - 2.0 * welz / r;
        else v = Bz - Ch / r - 2.0 * welz / r;
        X(F.urh, i, j, k) = v;
    }
}

struct UpdF {
    const double *uel, *vel, *wel;
    double *An, *Bn, *Cn, *thn;
    double *Ap, *Bp, *Cp, *thp;
};

// update, rot3dmg.f:964-1018, one thread per point.  The wall vorticity and
// wall temperatures are written as the reference writes them.  The copies
// Ap <- A <- An (and B, C, th) become a rotation of the three buffers on the
// host: the old Ap buffer, which becomes the new An, receives An here (after
// the last read of the old Ap), so all three levels hold what the
// reference's copies leave.
__global__ void __launch_bounds__(WG) k3_update(Geo g, UpdF F)
{
    const int nr = g.nr, nz = g.nz;
    const double dr = g.dr, dz = g.dz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)(nr + 2) * (nz + 2) * (g.np + 1)) return;
    const int i = (int)(t % (nr + 2));
    const int j = (int)((t / (nr + 2)) % (nz + 2));
    const int k = (int)(t / ((long long)(nr + 2) * (nz + 2)));
    const size_t p = ((size_t)k * g.ldj + (size_t)j) * g.ldi + (size_t)i;
    const double *uel = F.uel, *vel = F.vel, *wel = F.wel;
    // A: endwalls j = 0, nz for i = 1..nr (rot3dmg.f:968-971)
    double a = F.An[p];
    if (i >= 1 && i <= nr) {
        if (j == 0) a = -(3.0 * X(vel, i, 1, k) - X(vel, i, 2, k) / 3.0) / dz;
        else if (j == nz) a = -(X(vel, i, nz - 1, k) / 3.0 - 3.0 * X(vel, i, nz, k)) / dz;
        if (j == 0 || j == nz) F.An[p] = a;
    }
    F.Ap[p] = a;
    // B: endwalls for i = 1..nr-1 (rot3dmg.f:973-976), cylinders i = 0, nr
    // for j = 1..nz-1 (rot3dmg.f:977-980)
    double b = F.Bn[p];
    bool bw = false;
    if (i >= 1 && i <= nr - 1 && (j == 0 || j == nz)) {
        bw = true;
        b = (j == 0) ? (3.0 * X(uel, i, 1, k) - X(uel, i, 2, k) / 3.0) / dz
                     : (-3.0 * X(uel, i, nz, k) + X(uel, i, nz - 1, k) / 3.0) / dz;
    } else if ((i == 0 || i == nr) && j >= 1 && j <= nz - 1) {
        bw = true;
        b = (i == 0) ? -(3.0 * X(wel, 1, j, k) - X(wel, 2, j, k) / 3.0) / dr
                     : -(X(wel, nr - 1, j, k) / 3.0 - 3.0 * X(wel, nr, j, k)) / dr;
    }
    if (bw) F.Bn[p] = b;
    F.Bp[p] = b;
    // C: cylinders i = 0, nr for j = 1..nz (rot3dmg.f:982-985)
    double c = F.Cn[p];
    if ((i == 0 || i == nr) && j >= 1 && j <= nz) {
        c = (i == 0) ? (3.0 * X(vel, 1, j, k) - X(vel, 2, j, k) / 3.0) / dr
                     : (X(vel, nr - 1, j, k) / 3.0 - 3.0 * X(vel, nr, j, k)) / dr;
        F.Cn[p] = c;
    }
    F.Cp[p] = c;
    // th: 0 on the inner, 1 on the outer cylinder (rot3dmg.f:987-990)
    double th = F.thn[p];
    if (i == 0 || i == nr + 1) {
        th = (i == 0) ? 0.0 : 1.0;
        F.thn[p] = th;
    }
    F.thp[p] = th;
}

inline int nblk(long long n, int wg) { return (int)((n + wg - 1) / wg); }

}  // namespace
}  // namespace vv3

using namespace vv3;

static Geo geo(const vv3_run *r)
{
    Geo g;
    g.nr = r->nr; g.nz = r->nz; g.np = r->np; g.ldi = r->ldi; g.ldj = r->ldj;
    g.dt = r->cfg.dt; g.dr = r->dr; g.dz = r->dz; g.dp = r->dp;
    g.Ra = r->cfg.Ra; g.Re = r->cfg.Re; g.Ptl = r->cfg.Ptl; g.Gr = r->cfg.Gr; g.r0 = r->cfg.r0;
    return g;
}

extern "C" {

int vv3_create(const vv3_config *cfg, int device, vv3_run **out)
{
    using vv::set_error;
    if (!cfg || !out) { set_error("vv3_create: null argument"); return VV_EINVAL; }
    *out = nullptr;
    if (cfg->nr < 4 || cfg->nz < 4 || cfg->np < 4 || (cfg->np & 1)) {
        set_error("vv3_create: need nr, nz >= 4 and an even np >= 4");
        return VV_EINVAL;
    }
    if (!(cfg->dt > 0) || !(cfg->rmax > cfg->r0) || !(cfg->r0 > 0) || !(cfg->zmax > 0) || !(cfg->Ptl > 0) ||
        cfg->Re == 0.0) {
        set_error("vv3_create: bad physical parameters");
        return VV_EINVAL;
    }
    VV_HIP(hipSetDevice(device));
    vv3_run *r = new vv3_run();
    r->cfg = *cfg;
    r->device = device;
    r->nr = cfg->nr; r->nz = cfg->nz; r->np = cfg->np;
    r->ldi = cfg->nr + 2; r->ldj = cfg->nz + 2;
    r->dr = (cfg->rmax - cfg->r0) / cfg->nr;
    r->dz = cfg->zmax / cfg->nz;
    r->dp = 2.0 * std::acos(-1.0) / cfg->np;           // rot3dmg.f:91-93
    r->nf = (size_t)r->ldi * r->ldj * (cfg->np + 1);
    auto fail = [&](int s) { vv3_destroy(r); return s; };
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) return fail(VV_EHIP);
    for (int q = 0; q < VV3_NFIELD; ++q) {
        if (hipMalloc(&r->f[q], r->nf * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
        if (hipMemsetAsync(r->f[q], 0, r->nf * sizeof(double), r->stream) != hipSuccess) return fail(VV_EHIP);
    }
    if (hipMalloc(&r->heat, (size_t)(cfg->np + 2) * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
    {
        const int st = poisson3_setup(r->pois, geo(r));
        if (st) return fail(st);
    }
    if (hipStreamSynchronize(r->stream) != hipSuccess) return fail(VV_EHIP);
    *out = r;
    return VV_OK;
}

int vv3_destroy(vv3_run *r)
{
    if (!r) return VV_OK;
    (void)hipSetDevice(r->device);
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    for (int q = 0; q < VV3_NFIELD; ++q)
        if (r->f[q]) (void)hipFree(r->f[q]);
    poisson3_free(r->pois);
    if (r->heat) (void)hipFree(r->heat);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
    return VV_OK;
}

int vv3_put(vv3_run *r, int field, const double *host)
{
    if (!r || !host || field < 0 || field >= VV3_NFIELD) { vv::set_error("vv3_put: bad argument"); return VV_EINVAL; }
    VV_HIP(hipSetDevice(r->device));
    VV_HIP(hipMemcpyAsync(r->f[field], host, r->nf * sizeof(double), hipMemcpyHostToDevice, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vv3_get(vv3_run *r, int field, double *host)
{
    if (!r || !host || field < 0 || field >= VV3_NFIELD) { vv::set_error("vv3_get: bad argument"); return VV_EINVAL; }
    VV_HIP(hipSetDevice(r->device));
    VV_HIP(hipMemcpyAsync(host, r->f[field], r->nf * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

static Fields fields(vv3_run *r)
{
    Fields F;
    F.A = r->f[VV3_A]; F.B = r->f[VV3_B]; F.C = r->f[VV3_C];
    F.Ap = r->f[VV3_AP]; F.Bp = r->f[VV3_BP]; F.Cp = r->f[VV3_CP];
    F.th = r->f[VV3_TH]; F.thp = r->f[VV3_THP];
    F.uel = r->f[VV3_UEL]; F.vel = r->f[VV3_VEL]; F.wel = r->f[VV3_WEL];
    F.An = r->f[VV3_AN]; F.Bn = r->f[VV3_BN]; F.Cn = r->f[VV3_CN]; F.thn = r->f[VV3_THN];
    return F;
}

static int launch_explicit(vv3_run *r, int mask)
{
    const long long cnt = (long long)(r->nr) * (r->nz + 2) * (r->np + 1);   // the largest range (temperature)
    hipLaunchKernelGGL(k3_explicit, dim3(nblk(cnt, WG), 4), dim3(WG), 0, r->stream, geo(r), fields(r), mask);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

static int launch_rhs(vv3_run *r, int which)
{
    RhsF F;
    F.An = r->f[VV3_AN]; F.Bn = r->f[VV3_BN]; F.Cn = r->f[VV3_CN]; F.wel = r->f[VV3_WEL];
    F.wrh = r->f[VV3_WRH]; F.vrh = r->f[VV3_VRH]; F.urh = r->f[VV3_URH];
    const long long cnt = (long long)r->nr * r->nz * (r->np + 1);
    hipLaunchKernelGGL(k3_rhs, dim3(nblk(cnt, WG)), dim3(WG), 0, r->stream, geo(r), F, which);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

static int launch_poisson(vv3_run *r, int which)
{
    const int rhs = (which == 1) ? VV3_WRH : (which == 2) ? VV3_VRH : VV3_URH;
    const int psi = (which == 1) ? VV3_WEL : (which == 2) ? VV3_VEL : VV3_UEL;
    const int psip = (which == 1) ? VV3_WELP : (which == 2) ? VV3_VELP : VV3_UELP;
    return poisson3_launch(r->pois, geo(r), which, r->f[psi], r->f[psip], r->f[rhs], r->stream);
}

static int launch_update(vv3_run *r)
{
    UpdF F;
    F.uel = r->f[VV3_UEL]; F.vel = r->f[VV3_VEL]; F.wel = r->f[VV3_WEL];
    F.An = r->f[VV3_AN]; F.Bn = r->f[VV3_BN]; F.Cn = r->f[VV3_CN]; F.thn = r->f[VV3_THN];
    F.Ap = r->f[VV3_AP]; F.Bp = r->f[VV3_BP]; F.Cp = r->f[VV3_CP]; F.thp = r->f[VV3_THP];
    hipLaunchKernelGGL(k3_update, dim3(nblk((long long)r->nf, WG)), dim3(WG), 0, r->stream, geo(r), F);
    VV_HIP(hipGetLastError());
    // rot3dmg.f:1003-1015: Ap <- A <- An as a rotation of the buffers
    const int trip[4][3] = {{VV3_AP, VV3_A, VV3_AN}, {VV3_BP, VV3_B, VV3_BN}, {VV3_CP, VV3_C, VV3_CN},
                            {VV3_THP, VV3_TH, VV3_THN}};
    for (auto &t : trip) {
        double *p = r->f[t[0]], *c = r->f[t[1]], *nn = r->f[t[2]];
        r->f[t[0]] = c;
        r->f[t[1]] = nn;
        r->f[t[2]] = p;
    }
    return VV_OK;
}

int vv3_vorticity(vv3_run *r, int which)
{
    if (!r || which < 0 || which > 3) { vv::set_error("vv3_vorticity: bad argument"); return VV_EINVAL; }
    VV_HIP(hipSetDevice(r->device));
    int st = launch_explicit(r, which == 0 ? 7 : 1 << (which - 1));
    if (!st) VV_HIP(hipStreamSynchronize(r->stream));
    return st;
}

int vv3_temperature(vv3_run *r)
{
    if (!r) { vv::set_error("vv3: null run"); return VV_EINVAL; }
    VV_HIP(hipSetDevice(r->device));
    int st = launch_explicit(r, 8);
    if (!st) VV_HIP(hipStreamSynchronize(r->stream));
    return st;
}

int vv3_heat(vv3_run *r, double *trans1)
{
    if (!r) { vv::set_error("vv3: null run"); return VV_EINVAL; }
    VV_HIP(hipSetDevice(r->device));
    hipLaunchKernelGGL(k3_heat, dim3(1), dim3(WG), 0, r->stream, geo(r), (const double *)r->f[VV3_TH], r->heat,
                       r->heat + r->np + 1);
    VV_HIP(hipGetLastError());
    double t = 0.0;
    VV_HIP(hipMemcpyAsync(&t, r->heat + r->np + 1, sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    if (trans1) *trans1 = t;
    return VV_OK;
}

int vv3_rhs(vv3_run *r, int which)
{
    if (!r || which < 1 || which > 3) { vv::set_error("vv3_rhs: bad argument"); return VV_EINVAL; }
    VV_HIP(hipSetDevice(r->device));
    int st = launch_rhs(r, which);
    if (!st) VV_HIP(hipStreamSynchronize(r->stream));
    return st;
}

int vv3_poisson(vv3_run *r, int which)
{
    if (!r || which < 1 || which > 3) { vv::set_error("vv3_poisson: bad argument"); return VV_EINVAL; }
    VV_HIP(hipSetDevice(r->device));
    int st = launch_poisson(r, which);
    if (!st) VV_HIP(hipStreamSynchronize(r->stream));
    return st;
}

int vv3_update(vv3_run *r)
{
    if (!r) { vv::set_error("vv3: null run"); return VV_EINVAL; }
    VV_HIP(hipSetDevice(r->device));
    int st = launch_update(r);
    if (!st) VV_HIP(hipStreamSynchronize(r->stream));
    return st;
}

int vv3_step(vv3_run *r, int n, double *trans1)
{
    if (!r || n < 0) { vv::set_error("vv3_step: bad argument"); return VV_EINVAL; }
    VV_HIP(hipSetDevice(r->device));
    for (int s = 0; s < n; ++s) {
        int st = launch_explicit(r, 15);
        if (!st) {
            hipLaunchKernelGGL(k3_heat, dim3(1), dim3(WG), 0, r->stream, geo(r), (const double *)r->f[VV3_TH],
                               r->heat, r->heat + r->np + 1);
            VV_HIP(hipGetLastError());
        }
        if (!st) st = launch_rhs(r, 1);
        if (!st) st = launch_poisson(r, 1);
        if (!st) st = launch_rhs(r, 2);
        if (!st) st = launch_poisson(r, 2);
        if (!st) st = launch_rhs(r, 3);
        if (!st) st = launch_poisson(r, 3);
        if (!st) st = launch_update(r);
        if (st) return st;
    }
    double t = 0.0;
    VV_HIP(hipMemcpyAsync(&t, r->heat + r->np + 1, sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    if (trans1) *trans1 = t;
    return VV_OK;
}

}  // extern "C"
