// vvs.hip -- the 3D step of sect360v2mg.f on gfx950: the rotating annulus
// of rot3dmg.f with a radial barrier, so that phi is bounded by walls (the
// barrier's two faces, the planes k = 0 and k = np).
//
// Staggering in phi (sect360v2mg.f:206-223, 964-980): th, u, w and B at the
// cells k = 1..np with the planes k = 0 and np+1 on the barrier's faces; v,
// A and C at the faces k = 0..np.  Fields hold np+2 planes.
//
// Explicit part: one thread per point for Avorticity / Bvorticity /
// Cvorticity / temperature, the Poisson right-hand sides and the wall
// update, evaluating the reference's expressions in its operation order
// (the library is compiled with -ffp-contract=off; divisions by constants
// and radii through their correctly rounded reciprocals, which returns the
// IEEE quotient): the results equal those
// of the C restatement of the reference's routines bit for bit (the
// reference itself does not compile without its include files).
//
// Poisson: the three operators of relax1/2/3 (sect360v2mg.f:1817-1848,
// 2034-2065, 2233-2261) are separable.  Their phi rows are the second
// difference with zero-flux faces (w, u: eigenvectors cos(m pi (k-1/2)/np),
// m = 0..np-1) or with zero values on the faces (v: sin(m pi k/np),
// m = 1..np-1), both with eigenvalues 2 cos(m pi/np) - 2, so a solve is
//   1. the phi transform: by FFT in LDS for np a power of two (8..2048, the
//      tank's y transforms, vv_ytrans.h), else a GEMM with the eigenvector
//      table,
//   2. the eigenbasis of the z rows (a batched GEMM),
//   3. one banded system in r per (phi mode, z eigenvalue), a thread each,
//   4. the inverse z and phi transforms,
// exact to rounding where the reference iterates multigrid cycles to an
// absolute residual of 1e-5 (tolcon, sect360v2mg.f:1668).

#include "vv3_internal.h"
#include "vv_gemm.h"
#include "vv_ytrans.h"
#include "vv_zsplit.h"
#include "../../include/vvsector.h"

#include <cmath>
#include <cstring>
#include <vector>

#define VVS_TRY(call)                 \
    do {                               \
        const int st_ = (call);        \
        if (st_ != VV_OK) return st_;  \
    } while (0)

namespace vvs {

using vv3::Geo;

struct SF {
    const double *A, *B, *C, *Ap, *Bp, *Cp, *th, *thp, *uel, *vel, *wel;
    double *An, *Bn, *Cn, *thn, *urh, *vrh, *wrh;
};

// A positive divisor with its correctly rounded reciprocal.  x / b is
// evaluated as q0 = x rb, t = fma(b, q0, -x) = b q0 - x (exact),
// q = fma(-t, rb, q0): for rb = RN(1/b) this is the correctly rounded
// quotient (Markstein's correction), the IEEE quotient of the reference's
// division, in three instructions.  Forming -t gives a zero quotient the
// sign IEEE division gives it.
struct Dv {
    double v, r;
};
__device__ __forceinline__ double operator/(double x, Dv b)
{
    const double q0 = x * b.r;
    const double t = __builtin_fma(b.v, q0, -x);
    return __builtin_fma(-t, b.r, q0);
}
// the same for a divisor of either sign (the exact case kept apart)
struct Dvs {
    double v, r;
};
__device__ __forceinline__ double operator/(double x, Dvs b)
{
    const double q0 = x * b.r;
    const double e = __builtin_fma(-b.v, q0, x);
    return (e == 0.0) ? q0 : __builtin_fma(e, b.r, q0);
}

// the constant divisors of the explicit routines, each formed on the host as
// the reference's expression forms it (dr*dr, Ptl*dt, 2*Ptl, 2*dp)
struct SK {
    Dv dr, dz, dp, dr2, dz2, dp2, three, ptl, ptl2, ptldt, dt, dp2x;
    Dvs re;
};

// a point's radius r with r*r and 2*(r*r) as divisors; RN(1/(2 q)) is
// RN(1/q)/2 exactly
struct Dr {
    Dv r, r2, r2h;
    __device__ explicit Dr(double rv)
    {
        const double q = rv * rv, iq = 1.0 / q;
        r = Dv{rv, 1.0 / rv};
        r2 = Dv{q, iq};
        r2h = Dv{2.0 * q, 0.5 * iq};
    }
};

// one sector Poisson operator (1: w, 2: v, 3: u): unknowns i = 1..i1,
// j = 1..j1, k = 1..k1
struct SOp {
    int i1 = 0, j1 = 0, k1 = 0, m = 0;
    double e1 = 0.0, en = 0.0;   // r rows 1 and i1: coefficient two points inward
    double *W = nullptr;         // k1 x k1: forward phi transform, row q = mode
    double *Wi = nullptr;        // k1 x k1: inverse, row = plane
    double *lphi = nullptr;      // k1 mode eigenvalues 2 cos(m pi/np) - 2
    double *V = nullptr, *Vinv = nullptr, *lz = nullptr;   // z eigenbasis
    bool zsplit = false;         // V, Vinv: the parity halves of vv3::z_split; lz even modes first
    double *rt = nullptr;        // 4 x i1: sub, diagonal, super, 1/(dp^2 r^2)
};

struct SPoisson {
    SOp op[3];
    double *S = nullptr, *T = nullptr, *cp = nullptr;
    double *S2 = nullptr, *T2 = nullptr, *cp2 = nullptr;   // a second set: v's solve beside w's and u's
    bool fft = false;     // phi transforms by FFT (np a power of two, 8..2048) instead of the W / Wi GEMMs
    vvt::YTrans yt;
};

namespace {

#define X(a, i, j, k) ((a)[((size_t)(k) * g.ldj + (size_t)(j)) * g.ldi + (size_t)(i)])
#define XW(a, i, j, k) X(a, i, j, k)

// thh of the row kk of temperature's interior branch (sect360v2mg.f:799-809):
// the value the reference's row k = np reuses
__device__ __forceinline__ double thh_row(const Geo &g, const SF &F, const SK &K, int i, int j, int kk)
{
    const double *th = F.th, *vel = F.vel;
    const double dp = g.dp;
    const double r = g.r0 + ((double)i - 0.5) * g.dr;
    const Dr R(r);
    const double v = (j == 0 || j == g.nz + 1) ? 0.0 : (X(vel, i, j, kk) + X(vel, i, j, kk - 1)) * 0.5;
    double thh = (X(th, i, j, kk + 1) - X(th, i, j, kk - 1)) / K.dp2x;
    const double Pe = fabs(v * dp * r / K.ptl);
    if (Pe > 2) {
        if (v > 0) thh = (X(th, i, j, kk) - X(th, i, j, kk - 1)) / K.dp;
        else if (v < 0) thh = (X(th, i, j, kk + 1) - X(th, i, j, kk)) / K.dp;
    }
    return thh;
}

__device__ __forceinline__ void d_avort(const Geo &g, const SF &F, const SK &K, int i, int j, int k)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, dt = g.dt;
    const double Ptl = g.Ptl, Ra = g.Ra, Re = g.Re, Gr = g.Gr;
    (void)nr; (void)nz; (void)np; (void)dz; (void)dt; (void)Ra; (void)Re; (void)Gr; (void)Ptl; (void)dp;
    const double *A = F.A, *B = F.B, *C = F.C, *Ap = F.Ap, *Bp = F.Bp, *Cp = F.Cp, *th = F.th, *thp = F.thp;
    const double *uel = F.uel, *vel = F.vel, *wel = F.wel, *An = F.An, *Bn = F.Bn, *Cn = F.Cn;
    (void)A; (void)B; (void)C; (void)Ap; (void)Bp; (void)Cp; (void)th; (void)thp; (void)uel; (void)vel; (void)wel;
    (void)An; (void)Bn; (void)Cn;

    const int im = i - 1, jp = j + 1, km = k - 1, ka = k + 1;
    const double r = g.r0 + ((double)i - 0.5) * dr;
    const Dr R(r);
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
        Arra = (3.2 * X(A, i - 1, j, k) + 2.0 * X(A, i + 1, j, k) - 0.2 * X(A, i + 2, j, k)) / K.dr2;
        Ara = (-4.0 * X(A, i - 1, j, k) / K.three + X(A, i + 1, j, k) / K.three) / K.dr;
        Ar = (-4.0 * X(A, i - 1, j, k) / K.three + X(A, i, j, k) + X(A, i + 1, j, k) / K.three) / K.dr;
        cnr = -2.5 / K.dr2 + 0.5 / K.dr / R.r;
    } else if (i == nr) {
        Arra = (3.2 * X(A, i + 1, j, k) + 2.0 * X(A, i - 1, j, k) - 0.2 * X(A, i - 2, j, k)) / K.dr2;
        Ara = (-X(A, i - 1, j, k) / K.three + 4.0 * X(A, i + 1, j, k) / K.three) / K.dr;
        Ar = (-X(A, i - 1, j, k) / K.three - X(A, i, j, k) + 4.0 * X(A, i + 1, j, k) / K.three) / K.dr;
        cnr = -2.5 / K.dr2 - 0.5 / K.dr / R.r;
    } else {
        Arra = (X(A, i - 1, j, k) + X(A, i + 1, j, k)) / K.dr2;
        Ara = (X(A, i + 1, j, k) - X(A, i - 1, j, k)) * 0.5 / K.dr;
        Ar = Ara;
        cnr = -1.0 / K.dr2;
        const double Pe = fabs(u * dr / K.ptl);
        if (Pe > 2) {
            if (u > 0) Ar = (X(A, i, j, k) - X(A, i - 1, j, k)) / K.dr;
            else if (u < 0) Ar = (X(A, i + 1, j, k) - X(A, i, j, k)) / K.dr;
        }
    }
    const double uelr = ((X(uel, i, j, k) + X(uel, i, jp, k) + X(uel, i, j, ka) + X(uel, i, jp, ka)) * 0.25 -
                         (X(uel, im, j, k) + X(uel, im, jp, k) + X(uel, im, j, ka) + X(uel, im, jp, ka)) *
                             0.25) /
                        dr;
    const double Azza = (X(A, i, j - 1, k) + X(A, i, j + 1, k)) / K.dz2;
    double Az = (X(A, i, j + 1, k) - X(A, i, j - 1, k)) * 0.5 / K.dz;
    {
        const double Pe = fabs(w * dz / K.ptl);
        if (Pe > 2) {
            if (w > 0) Az = (X(A, i, j, k) - X(A, i, j - 1, k)) / K.dz;
            else if (w < 0) Az = (X(A, i, j + 1, k) - X(A, i, j, k)) / K.dz;
        }
    }
    const double uelz = ((X(uel, i, jp, k) + X(uel, i, jp, ka) + X(uel, im, jp, k) + X(uel, im, jp, ka)) * 0.25 -
                         (X(uel, i, j, k) + X(uel, i, j, ka) + X(uel, im, j, k) + X(uel, im, j, ka)) * 0.25) /
                        dz;
    const double thz = (X(th, i, jp, k) + X(th, i, jp, ka) - X(th, i, j, k) - X(th, i, j, ka)) * 0.5 / K.dz;
    const double thh = (X(th, i, j, ka) + X(th, i, jp, ka) - X(th, i, j, k) - X(th, i, jp, k)) * 0.5 / K.dp;
    const double Ahha = (X(A, i, j, ka) + X(A, i, j, km)) / K.dp2;
    double Ah = (X(A, i, j, ka) - X(A, i, j, km)) * 0.5 / K.dp;
    {
        const double Pe = fabs(v * dp * r / K.ptl);
        if (Pe > 2) {
            if (v > 0) Ah = (X(A, i, j, k) - X(A, i, j, km)) / K.dp;
            else if (v < 0) Ah = (X(A, i, j, ka) - X(A, i, j, k)) / K.dp;
        }
    }
    const double Bh = (X(B, i, j, ka) + X(B, im, j, ka) - X(B, i, j, k) - X(B, im, j, k)) * 0.5 / K.dp;
    const double uelh = ((X(uel, i, j, ka) + X(uel, i, jp, ka) + X(uel, im, j, ka) + X(uel, im, jp, ka)) * 0.25 -
                         (X(uel, i, j, k) + X(uel, i, jp, k) + X(uel, im, j, k) + X(uel, im, jp, k)) * 0.25) /
                        dp;
    const double cnz = -1.0 / K.dz2;
    const double cnh = -1.0 / K.dp2 / R.r2;
    const double RHS1 = Arra + Ara / R.r + Ahha / R.r2 + Azza - 2.0 * Bh / R.r2 + X(A, i, j, k) / K.ptldt;
    const double ULT = (u * Ar + v * Ah / R.r + w * Az - uelh * BB / R.r - uelz * CC) / K.ptl;
    const double RHS2 = -2.0 * Re * uelz + 2.0 * Ra * (tem * uelz + thz * u) / K.re / K.ptl - Gr * Ptl / R.r * thh;
    const double cf1 = 1.0 / K.ptldt - uelr / K.ptl2 + 1.0 / R.r2h - cnr - cnz - cnh;
    const double cf2 = uelr / K.ptl2 - 1.0 / R.r2h + cnr + cnz + cnh;
    XW(F.An, i, j, k) = (cf2 * X(Ap, i, j, k) - ULT - RHS2 + RHS1) / cf1;
}

__device__ __forceinline__ void d_bvort(const Geo &g, const SF &F, const SK &K, int i, int j, int k)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, dt = g.dt;
    const double Ptl = g.Ptl, Ra = g.Ra, Re = g.Re, Gr = g.Gr;
    (void)nr; (void)nz; (void)np; (void)dz; (void)dt; (void)Ra; (void)Re; (void)Gr; (void)Ptl; (void)dp;
    const double *A = F.A, *B = F.B, *C = F.C, *Ap = F.Ap, *Bp = F.Bp, *Cp = F.Cp, *th = F.th, *thp = F.thp;
    const double *uel = F.uel, *vel = F.vel, *wel = F.wel, *An = F.An, *Bn = F.Bn, *Cn = F.Cn;
    (void)A; (void)B; (void)C; (void)Ap; (void)Bp; (void)Cp; (void)th; (void)thp; (void)uel; (void)vel; (void)wel;
    (void)An; (void)Bn; (void)Cn;

    const int ip = i + 1, jp = j + 1, km = k - 1, ka = k + 1;
    const double r = g.r0 + (double)i * dr;
    const Dr R(r);
    const double u = (X(uel, i, j, k) + X(uel, i, jp, k)) * 0.5;
    const double v = (X(vel, i, j, km) + X(vel, i, jp, km) + X(vel, i, j, k) + X(vel, i, jp, k) +
                      X(vel, ip, j, km) + X(vel, ip, jp, km) + X(vel, ip, j, k) + X(vel, ip, jp, k)) *
                     0.125;
    const double w = (X(wel, i, j, k) + X(wel, ip, j, k)) * 0.5;
    const double tem = (X(th, i, jp, k) + X(th, ip, jp, k) + X(th, i, j, k) + X(th, ip, j, k)) * 0.25;
    const double AA = (X(A, i, j, k) + X(A, ip, j, k) + X(A, i, j, km) + X(A, ip, j, km)) * 0.25;
    const double CC = (X(C, i, j, k) + X(C, i, jp, k) + X(C, i, j, km) + X(C, i, jp, km)) * 0.25;
    const double Brra = (X(B, i - 1, j, k) + X(B, i + 1, j, k)) / K.dr2;
    const double Bra = (X(B, i + 1, j, k) - X(B, i - 1, j, k)) * 0.5 / K.dr;
    double Br = Bra;
    {
        const double Pe = fabs(u * dr / K.ptl);
        if (Pe > 2) {
            if (u > 0) Br = (X(B, i, j, k) - X(B, i - 1, j, k)) / K.dr;
            else if (u < 0) Br = (X(B, i + 1, j, k) - X(B, i, j, k)) / K.dr;
        }
    }
    const double velr = ((X(vel, ip, j, km) + X(vel, ip, jp, km) + X(vel, ip, j, k) + X(vel, ip, jp, k)) * 0.25 -
                         (X(vel, i, j, km) + X(vel, i, jp, km) + X(vel, i, j, k) + X(vel, i, jp, k)) * 0.25) /
                        dr;
    const double Bzza = (X(B, i, j - 1, k) + X(B, i, j + 1, k)) / K.dz2;
    double Bz = (X(B, i, j + 1, k) - X(B, i, j - 1, k)) * 0.5 / K.dz;
    {
        const double Pe = fabs(w * dz / K.ptl);
        if (Pe > 2) {
            if (w > 0) Bz = (X(B, i, j, k) - X(B, i, j - 1, k)) / K.dz;
            else if (w < 0) Bz = (X(B, i, j + 1, k) - X(B, i, j, k)) / K.dz;
        }
    }
    const double velz = ((X(vel, i, jp, km) + X(vel, ip, jp, km) + X(vel, i, jp, k) + X(vel, ip, jp, k)) * 0.25 -
                         (X(vel, i, j, km) + X(vel, ip, j, km) + X(vel, i, j, k) + X(vel, ip, j, k)) * 0.25) /
                        dz;
    const double thz = (X(th, i, jp, k) + X(th, ip, jp, k) - X(th, i, j, k) - X(th, ip, j, k)) * 0.5 / K.dz;
    const double thr = (X(th, ip, jp, k) + X(th, ip, j, k) - X(th, i, jp, k) - X(th, i, j, k)) * 0.5 / K.dr;
    double Bhha, Bh, cnh;
    if (k == 1) {
        Bhha = (3.2 * X(B, i, j, 0) + 2.0 * X(B, i, j, 2) - 0.2 * X(B, i, j, 3)) / K.dp2;
        Bh = (-4.0 / K.three * X(B, i, j, 0) + X(B, i, j, 1) + X(B, i, j, 2) / K.three) / K.dp;
        cnh = -2.5 / K.dp2 / R.r2;
    } else if (k == np) {
        Bhha = (-0.2 * X(B, i, j, np - 2) + 2.0 * X(B, i, j, np - 1) + 3.2 * X(B, i, j, np + 1)) / K.dp2;
        Bh = (-X(B, i, j, np - 1) / K.three - X(B, i, j, np) + 4.0 / K.three * X(B, i, j, np + 1)) / K.dp;
        cnh = -2.5 / K.dp2 / R.r2;
    } else {
        Bhha = (X(B, i, j, ka) + X(B, i, j, km)) / K.dp2;
        Bh = (X(B, i, j, ka) - X(B, i, j, km)) * 0.5 / K.dp;
        cnh = -1.0 / K.dp2 / R.r2;
        const double Pe = fabs(v * dp * r / K.ptl);
        if (Pe > 2) {
            if (v > 0) Bh = (X(B, i, j, k) - X(B, i, j, km)) / K.dp;
            else if (v < 0) Bh = (X(B, i, j, ka) - X(B, i, j, k)) / K.dp;
        }
    }
    const double Ah = (X(A, i, j, k) + X(A, ip, j, k) - X(A, i, j, km) - X(A, ip, j, km)) * 0.5 / K.dp;
    const double velh = ((X(vel, i, j, k) + X(vel, ip, j, k) + X(vel, i, jp, k) + X(vel, ip, jp, k)) * 0.25 -
                         (X(vel, i, j, km) + X(vel, ip, j, km) + X(vel, i, jp, km) + X(vel, ip, jp, km)) * 0.25) /
                        dp;
    const double cnr = -1.0 / K.dr2;
    const double cnz = -1.0 / K.dz2;
    const double RHS1 = Brra + Bra / R.r + Bhha / R.r2 + Bzza + 2.0 * Ah / R.r2 + X(B, i, j, k) / K.ptldt;
    const double ULT = (u * Br + v * Bh / R.r + w * Bz - (velr - v / R.r) * AA - velz * CC) / K.ptl;
    const double RHS2 = -2.0 * Re * velz + r * Ra * thz + 2.0 * Ra * (tem * velz + thz * v) / K.re / K.ptl +
                        Gr * Ptl * thr;
    const double cf1 = 1.0 / K.ptldt - (u / R.r + velh / R.r) / K.ptl2 + 1.0 / R.r2h - cnr - cnz - cnh;
    const double cf2 = (u / R.r + velh / R.r) / K.ptl2 - 1.0 / R.r2h + cnr + cnz + cnh;
    XW(F.Bn, i, j, k) = (cf2 * X(Bp, i, j, k) - ULT - RHS2 + RHS1) / cf1;
}

__device__ __forceinline__ void d_cvort(const Geo &g, const SF &F, const SK &K, int i, int j, int k)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, dt = g.dt;
    const double Ptl = g.Ptl, Ra = g.Ra, Re = g.Re, Gr = g.Gr;
    (void)nr; (void)nz; (void)np; (void)dz; (void)dt; (void)Ra; (void)Re; (void)Gr; (void)Ptl; (void)dp;
    const double *A = F.A, *B = F.B, *C = F.C, *Ap = F.Ap, *Bp = F.Bp, *Cp = F.Cp, *th = F.th, *thp = F.thp;
    const double *uel = F.uel, *vel = F.vel, *wel = F.wel, *An = F.An, *Bn = F.Bn, *Cn = F.Cn;
    (void)A; (void)B; (void)C; (void)Ap; (void)Bp; (void)Cp; (void)th; (void)thp; (void)uel; (void)vel; (void)wel;
    (void)An; (void)Bn; (void)Cn;

    const int ip = i + 1, im = i - 1, jm = j - 1, km = k - 1, ka = k + 1;
    const double r = g.r0 + (double)i * dr;
    const Dr R(r);
    const double u = (X(uel, i, j, ka) + X(uel, i, j, k)) * 0.5;
    const double v = (X(vel, i, j, k) + X(vel, ip, j, k)) * 0.5;
    const double w = (X(wel, i, j, k) + X(wel, ip, j, k) + X(wel, i, j, ka) + X(wel, ip, j, ka) +
                      X(wel, i, jm, k) + X(wel, ip, jm, k) + X(wel, i, jm, ka) + X(wel, ip, jm, ka)) *
                     0.125;
    const double AA = (X(A, i, j, k) + X(A, i, jm, k) + X(A, ip, j, k) + X(A, ip, jm, k)) * 0.25;
    const double BB = (X(B, i, j, k) + X(B, i, jm, k) + X(B, i, j, ka) + X(B, i, jm, ka)) * 0.25;
    const double tem = (X(th, i, j, k) + X(th, ip, j, k) + X(th, ip, j, ka) + X(th, i, j, ka)) * 0.25;
    const double Crra = (X(C, i - 1, j, k) + X(C, i + 1, j, k)) / K.dr2;
    const double Cra = (X(C, i + 1, j, k) - X(C, i - 1, j, k)) * 0.5 / K.dr;
    double Cr = Cra;
    {
        const double Pe = fabs(u * dr / K.ptl);
        if (Pe > 2) {
            if (u > 0) Cr = (X(C, i, j, k) - X(C, i - 1, j, k)) / K.dr;
            else if (u < 0) Cr = (X(C, i + 1, j, k) - X(C, i, j, k)) / K.dr;
        }
    }
    const double welr = ((X(wel, ip, j, k) + X(wel, ip, j, ka) + X(wel, ip, jm, k) + X(wel, ip, jm, ka)) * 0.25 -
                         (X(wel, i, j, k) + X(wel, i, j, ka) + X(wel, i, jm, k) + X(wel, i, jm, ka)) * 0.25) /
                        dr;
    /* uelr (sect360v2mg.f:623-624) is computed by the reference and unused */
    const double thr = (X(th, ip, j, k) + X(th, ip, j, ka) - X(th, i, j, k) - X(th, i, j, ka)) * 0.5 / K.dr;
    double Czza, Cz, cnz;
    if (j == 1) {
        Czza = (3.2 * X(C, i, j - 1, k) + 2.0 * X(C, i, j + 1, k) - 0.2 * X(C, i, j + 2, k)) / K.dz2;
        Cz = (-4.0 * X(C, i, j - 1, k) / K.three + X(C, i, j, k) + X(C, i, j + 1, k) / K.three) / K.dz;
        cnz = -2.5 / K.dz2;
    } else if (j == nz) {
        Czza = (3.2 * X(C, i, j + 1, k) + 2.0 * X(C, i, j - 1, k) - 0.2 * X(C, i, j - 2, k)) / K.dz2;
        Cz = (-X(C, i, j - 1, k) / K.three - X(C, i, j, k) + 4.0 * X(C, i, j + 1, k) / K.three) / K.dz;
        cnz = -2.5 / K.dz2;
    } else {
        Czza = (X(C, i, j - 1, k) + X(C, i, j + 1, k)) / K.dz2;
        Cz = (X(C, i, j + 1, k) - X(C, i, j - 1, k)) * 0.5 / K.dz;
        const double Pe = fabs(w * dz / K.ptl);
        if (Pe > 2) {
            if (w > 0) Cz = (X(C, i, j, k) - X(C, i, j - 1, k)) / K.dz;
            else if (w < 0) Cz = (X(C, i, j + 1, k) - X(C, i, j, k)) / K.dz;
        }
        cnz = -1.0 / K.dz2;
    }
    const double welz = ((X(wel, i, j, k) + X(wel, ip, j, k) + X(wel, i, j, ka) + X(wel, ip, j, ka)) * 0.25 -
                         (X(wel, i, jm, k) + X(wel, ip, jm, k) + X(wel, i, jm, ka) + X(wel, ip, jm, ka)) * 0.25) /
                        dz;
    const double Chha = (X(C, i, j, ka) + X(C, i, j, km)) / K.dp2;
    double Ch = (X(C, i, j, ka) - X(C, i, j, km)) * 0.5 / K.dp;
    {
        const double Pe = fabs(v * dp * r / K.ptl);
        if (Pe > 2) {
            if (v > 0) Ch = (X(C, i, j, k) - X(C, i, j, km)) / K.dp;
            else if (v < 0) Ch = (X(C, i, j, ka) - X(C, i, j, k)) / K.dp;
        }
    }
    const double welh = ((X(wel, i, j, ka) + X(wel, ip, j, ka) + X(wel, i, jm, ka) + X(wel, ip, jm, ka)) * 0.25 -
                         (X(wel, i, j, k) + X(wel, ip, j, k) + X(wel, i, jm, k) + X(wel, ip, jm, k)) * 0.25) /
                        dp;
    /* velh (sect360v2mg.f:667-668) is computed by the reference and unused */
    const double thh = (X(th, i, j, ka) + X(th, ip, j, ka) - X(th, i, j, k) - X(th, ip, j, k)) * 0.5 / K.dp;
    const double cnr = -1.0 / K.dr2;
    const double cnh = -1.0 / K.dp2 / R.r2;
    const double RHS1 = Crra + Cra / R.r + Chha / R.r2 + Czza + X(C, i, j, k) / K.ptldt;
    const double ULT = (u * Cr + v * Ch / R.r + w * Cz - welr * AA - welh * BB / R.r) / K.ptl;
    const double RHS2 = -2.0 * Re * welz - Ra * thh + 2.0 * Ra * (tem * welz - thr * u - thh * v / R.r) / K.re / K.ptl;
    const double cf1 = 1.0 / K.ptldt - welz / K.ptl2 - cnr - cnz - cnh;
    const double cf2 = welz / K.ptl2 + cnr + cnz + cnh;
    XW(F.Cn, i, j, k) = (cf2 * X(Cp, i, j, k) - ULT - RHS2 + RHS1) / cf1;
}

__device__ __forceinline__ void d_temper(const Geo &g, const SF &F, const SK &K, int i, int j, int k)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, dt = g.dt;
    const double Ptl = g.Ptl, Ra = g.Ra, Re = g.Re, Gr = g.Gr;
    (void)nr; (void)nz; (void)np; (void)dz; (void)dt; (void)Ra; (void)Re; (void)Gr; (void)Ptl; (void)dp;
    const double *A = F.A, *B = F.B, *C = F.C, *Ap = F.Ap, *Bp = F.Bp, *Cp = F.Cp, *th = F.th, *thp = F.thp;
    const double *uel = F.uel, *vel = F.vel, *wel = F.wel, *An = F.An, *Bn = F.Bn, *Cn = F.Cn;
    (void)A; (void)B; (void)C; (void)Ap; (void)Bp; (void)Cp; (void)th; (void)thp; (void)uel; (void)vel; (void)wel;
    (void)An; (void)Bn; (void)Cn;
    // the row k = np advects with the thh of the row k = np-1 (see above)
    double thh = (k == np) ? thh_row(g, F, K, i, j, np - 1) : 0.0;

    const int im = i - 1, jm = j - 1, km = k - 1, ka = k + 1;
    const double r = g.r0 + ((double)i - 0.5) * dr;
    const Dr R(r);
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
        thrra = (3.2 * X(th, i - 1, j, k) + 2.0 * X(th, i + 1, j, k) - 0.2 * X(th, i + 2, j, k)) / K.dr2;
        thra = (-4.0 * X(th, i - 1, j, k) / K.three + X(th, i + 1, j, k) / K.three) / K.dr;
        thr = (-4.0 * X(th, i - 1, j, k) / K.three + X(th, i, j, k) + X(th, i + 1, j, k) / K.three) / K.dr;
        cnr = -2.5 / K.dr2 + 0.5 / K.dr / R.r;
    } else if (i == nr) {
        thrra = (3.2 * X(th, i + 1, j, k) + 2.0 * X(th, i - 1, j, k) - 0.2 * X(th, i - 2, j, k)) / K.dr2;
        thra = (-X(th, i - 1, j, k) / K.three + 4.0 * X(th, i + 1, j, k) / K.three) / K.dr;
        thr = (-X(th, i - 1, j, k) / K.three - X(th, i, j, k) + 4.0 * X(th, i + 1, j, k) / K.three) / K.dr;
        cnr = -2.5 / K.dr2 - 0.5 / K.dr / R.r;
    } else {
        thrra = (X(th, i - 1, j, k) + X(th, i + 1, j, k)) / K.dr2;
        thra = (X(th, i + 1, j, k) - X(th, i - 1, j, k)) * 0.5 / K.dr;
        thr = thra;
        const double Pe = fabs(u * dr / K.ptl);
        if (Pe > 2) {
            if (u > 0) thr = (X(th, i, j, k) - X(th, i - 1, j, k)) / K.dr;
            else if (u < 0) thr = (X(th, i + 1, j, k) - X(th, i, j, k)) / K.dr;
        }
        cnr = -1.0 / K.dr2;
    }
    double thzza, thz, cnz;
    if (j == 1) {
        thzza = (3.2 * X(th, i, j - 1, k) + 2.0 * X(th, i, j + 1, k) - 0.2 * X(th, i, j + 2, k)) / K.dz2;
        thz = (-4.0 * X(th, i, j - 1, k) / K.three + X(th, i, j, k) + X(th, i, j + 1, k) / K.three) / K.dz;
        cnz = -2.5 / K.dz2;
    } else if (j == nz) {
        thzza = (3.2 * X(th, i, j + 1, k) + 2.0 * X(th, i, j - 1, k) - 0.2 * X(th, i, j - 2, k)) / K.dz2;
        thz = (-X(th, i, j - 1, k) / K.three - X(th, i, j, k) + 4.0 * X(th, i, j + 1, k) / K.three) / K.dz;
        cnz = -2.5 / K.dz2;
    } else if (j == 0) {
        thzza = 8.0 * X(th, i, j + 1, k) / K.dz2;
        thz = 0.0;
        cnz = -4.0 / K.dz2;
    } else if (j == nz + 1) {
        thzza = 8.0 * X(th, i, j - 1, k) / K.dz2;
        thz = 0.0;
        cnz = -4.0 / K.dz2;
    } else {
        thzza = (X(th, i, j - 1, k) + X(th, i, j + 1, k)) / K.dz2;
        thz = (X(th, i, j + 1, k) - X(th, i, j - 1, k)) * 0.5 / K.dz;
        const double Pe = fabs(w * dz / K.ptl);
        if (Pe > 2) {
            if (w > 0) thz = (X(th, i, j, k) - X(th, i, j - 1, k)) / K.dz;
            else if (w < 0) thz = (X(th, i, j + 1, k) - X(th, i, j, k)) / K.dz;
        }
        cnz = -1.0 / K.dz2;
    }
    double thhha, cnp;
    if (k == 1) {
        thhha = (3.2 * X(th, i, j, 0) + 2.0 * X(th, i, j, 2) - 0.2 * X(th, i, j, 3)) / K.dp2;
        thh = (-4.0 * X(th, i, j, 0) / K.three + X(th, i, j, 1) + X(th, i, j, 2) / K.three) / K.dp;
        cnp = -2.5 / K.dp2 / R.r2;
    } else if (k == np) {
        thhha = (-0.2 * X(th, i, j, np - 2) + 2.0 * X(th, i, j, np - 1) + 3.2 * X(th, i, j, np + 1)) / K.dp2;
        thz = (-X(th, i, j, np - 1) / K.three - X(th, i, j, np) + 4.0 * X(th, i, j, np + 1) / K.three) / K.dp;
        cnp = -2.5 / K.dp2 / R.r2;
    } else if (k == 0) {
        thhha = 8.0 * X(th, i, j, 1) / K.dp2;
        thh = 0.0;
        cnp = -4.0 / K.dp2 / R.r2;
    } else if (k == np + 1) {
        thhha = 8.0 * X(th, i, j, np) / K.dp2;
        thh = 0.0;
        cnp = -4.0 / K.dp2 / R.r2;
    } else {
        thhha = (X(th, i, j, ka) + X(th, i, j, km)) / K.dp2;
        thh = (X(th, i, j, ka) - X(th, i, j, km)) / K.dp2x;
        const double Pe = fabs(v * dp * r / K.ptl);
        if (Pe > 2) {
            if (v > 0) thh = (X(th, i, j, k) - X(th, i, j, km)) / K.dp;
            else if (v < 0) thh = (X(th, i, j, ka) - X(th, i, j, k)) / K.dp;
        }
        cnp = -1.0 / K.dp2 / R.r2;
    }
    const double RHS1 = thrra + thra / R.r + thhha / R.r2 + thzza + X(th, i, j, k) / K.dt;
    const double ULT = u * thr + v * thh / R.r + w * thz;
    const double cf1 = 1.0 / K.dt - cnr - cnz - cnp;
    const double cf2 = cnr + cnz + cnp;
    XW(F.thn, i, j, k) = (cf2 * X(thp, i, j, k) - ULT + RHS1) / cf1;
}

__device__ __forceinline__ void d_rhsw(const Geo &g, const SF &F, const SK &K, int i, int j, int k)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, dt = g.dt;
    const double Ptl = g.Ptl, Ra = g.Ra, Re = g.Re, Gr = g.Gr;
    (void)nr; (void)nz; (void)np; (void)dz; (void)dt; (void)Ra; (void)Re; (void)Gr; (void)Ptl; (void)dp;
    const double *A = F.A, *B = F.B, *C = F.C, *Ap = F.Ap, *Bp = F.Bp, *Cp = F.Cp, *th = F.th, *thp = F.thp;
    const double *uel = F.uel, *vel = F.vel, *wel = F.wel, *An = F.An, *Bn = F.Bn, *Cn = F.Cn;
    (void)A; (void)B; (void)C; (void)Ap; (void)Bp; (void)Cp; (void)th; (void)thp; (void)uel; (void)vel; (void)wel;
    (void)An; (void)Bn; (void)Cn;

    const int km = k - 1;
    const double r = g.r0 + ((double)i - 0.5) * dr;
    const Dr R(r);
    const double Ah = (X(An, i, j, k) - X(An, i, j, km)) / K.dp;
    const double Br = (X(Bn, i, j, k) - X(Bn, i - 1, j, k)) / K.dr;
    double v;
    if (i == 1) v = -X(Bn, 1, j, k) / K.dr - 0.5 * X(Bn, 1, j, k) / R.r;
    else if (i == nr) v = X(Bn, nr - 1, j, k) / K.dr - 0.5 * X(Bn, nr - 1, j, k) / R.r;
    else v = -Br - 0.5 * (X(Bn, i, j, k) + X(Bn, i - 1, j, k)) / R.r;
    if (k == 1) v = v + X(An, i, j, k) / R.r / K.dp;
    else if (k == np) v = v - X(An, i, j, km) / R.r / K.dp;
    else v = v + Ah / R.r;
    XW(F.wrh, i, j, k) = v;
}

__device__ __forceinline__ void d_rhsv(const Geo &g, const SF &F, const SK &K, int i, int j, int k)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, dt = g.dt;
    const double Ptl = g.Ptl, Ra = g.Ra, Re = g.Re, Gr = g.Gr;
    (void)nr; (void)nz; (void)np; (void)dz; (void)dt; (void)Ra; (void)Re; (void)Gr; (void)Ptl; (void)dp;
    const double *A = F.A, *B = F.B, *C = F.C, *Ap = F.Ap, *Bp = F.Bp, *Cp = F.Cp, *th = F.th, *thp = F.thp;
    const double *uel = F.uel, *vel = F.vel, *wel = F.wel, *An = F.An, *Bn = F.Bn, *Cn = F.Cn;
    (void)A; (void)B; (void)C; (void)Ap; (void)Bp; (void)Cp; (void)th; (void)thp; (void)uel; (void)vel; (void)wel;
    (void)An; (void)Bn; (void)Cn;

    const double r = g.r0 + ((double)i - 0.5) * dr;

    const Dr R(r);
    const double Az = (X(An, i, j, k) - X(An, i, j - 1, k)) / K.dz;
    const double Cr = (X(Cn, i, j, k) - X(Cn, i - 1, j, k)) / K.dr;
    double v;
    if (i == 1) v = X(Cn, 1, j, k) / K.dr + X(Cn, 1, j, k) / R.r;
    else if (i == nr) v = -X(Cn, nr - 1, j, k) / K.dr + X(Cn, nr - 1, j, k) / R.r;
    else v = Cr + (X(Cn, i, j, k) + X(Cn, i - 1, j, k)) / R.r;
    if (j == 1) v = v - X(An, i, 1, k) / K.dz;
    else if (j == nz) v = v + X(An, i, nz - 1, k) / K.dz;
    else v = v - Az;
    XW(F.vrh, i, j, k) = v;
}

__device__ __forceinline__ void d_rhsu(const Geo &g, const SF &F, const SK &K, int i, int j, int k)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, dt = g.dt;
    const double Ptl = g.Ptl, Ra = g.Ra, Re = g.Re, Gr = g.Gr;
    (void)nr; (void)nz; (void)np; (void)dz; (void)dt; (void)Ra; (void)Re; (void)Gr; (void)Ptl; (void)dp;
    const double *A = F.A, *B = F.B, *C = F.C, *Ap = F.Ap, *Bp = F.Bp, *Cp = F.Cp, *th = F.th, *thp = F.thp;
    const double *uel = F.uel, *vel = F.vel, *wel = F.wel, *An = F.An, *Bn = F.Bn, *Cn = F.Cn;
    (void)A; (void)B; (void)C; (void)Ap; (void)Bp; (void)Cp; (void)th; (void)thp; (void)uel; (void)vel; (void)wel;
    (void)An; (void)Bn; (void)Cn;

    const int km = k - 1;
    const double r = g.r0 + (double)i * dr;
    const Dr R(r);
    const double Bz = (X(Bn, i, j, k) - X(Bn, i, j - 1, k)) / K.dz;
    const double welz = 0.5 * (X(wel, i, j, k) + X(wel, i + 1, j, k) - X(wel, i, j - 1, k) -
                               X(wel, i + 1, j - 1, k)) / K.dz;
    const double Ch = (X(Cn, i, j, k) - X(Cn, i, j, km)) / K.dp;
    double v;
    if (j == 1) v = X(Bn, i, 1, k) / K.dz - 2.0 * welz / R.r;
    else if (j == nz) v = -X(Bn, i, nz - 1, k) / K.dz - 2.0 * welz / R.r;
    else v = Bz - 2.0 * welz / R.r;
    if (k == 1) v = v - X(Cn, i, j, k) / R.r / K.dp;
    else if (k == np) v = v + X(Cn, i, j, k - 1) / R.r / K.dp;
    else v = v - Ch / R.r;
    XW(F.urh, i, j, k) = v;
}


// x over the (i, j) plane, y over the planes k = 0..np+1.  MASK: 1 A, 2 B,
// 4 C, 8 temperature (all read only the old levels)
template <int MASK>
__global__ void __launch_bounds__(256) ks_explicit(Geo g, SF F, SK K)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ldi * g.ldj) return;
    const int j = t / g.ldi, i = t - j * g.ldi, k = blockIdx.y;
    const int nr = g.nr, nz = g.nz, np = g.np;
    if ((MASK & 1) && i >= 1 && i <= nr && j >= 1 && j <= nz - 1 && k >= 1 && k <= np - 1) d_avort(g, F, K, i, j, k);
    if ((MASK & 2) && i >= 1 && i <= nr - 1 && j >= 1 && j <= nz - 1 && k >= 1 && k <= np) d_bvort(g, F, K, i, j, k);
    if ((MASK & 4) && i >= 1 && i <= nr - 1 && j >= 1 && j <= nz && k >= 1 && k <= np - 1) d_cvort(g, F, K, i, j, k);
    if ((MASK & 8) && i >= 1 && i <= nr) d_temper(g, F, K, i, j, k);
}

// rhsw / rhsv / rhsu on their unknowns
template <int WHICH>
__global__ void __launch_bounds__(256) ks_rhs(Geo g, SF F, SK K)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ldi * g.ldj) return;
    const int j = t / g.ldi, i = t - j * g.ldi, k = blockIdx.y;
    const int i1 = (WHICH == 3) ? g.nr - 1 : g.nr, j1 = (WHICH == 1) ? g.nz - 1 : g.nz;
    const int k1 = (WHICH == 2) ? g.np - 1 : g.np;
    if (i < 1 || i > i1 || j < 1 || j > j1 || k < 1 || k > k1) return;
    if (WHICH == 1) d_rhsw(g, F, K, i, j, k);
    else if (WHICH == 2) d_rhsv(g, F, K, i, j, k);
    else d_rhsu(g, F, K, i, j, k);
}

// heat, sect360v2mg.f:840-881: the flux of each plane k = 0..np+1, a
// workgroup each: its z terms in parallel into LDS, then their sum in the
// reference's order by one thread ...
__global__ void __launch_bounds__(256) ks_heat_planes(Geo g, const double *__restrict__ th, double *__restrict__ tr)
{
    extern __shared__ double term[];   // nz+2 terms
    const int k = blockIdx.x, nz = g.nz;
    const double dr = g.dr, dz = g.dz;
    for (int j = threadIdx.x; j <= nz + 1; j += blockDim.x) {
        const double w = (j == 0 || j == nz + 1) ? 240.0 : (j == 1 || j == nz) ? 80.0 : 60.0;
        term[j] = dz * (-184.0 * X(th, 0, j, k) + 225.0 * X(th, 1, j, k) - 50.0 * X(th, 2, j, k) +
                        9.0 * X(th, 3, j, k)) / (w * dr);
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double trans = term[0];
    trans = trans + term[1];
    for (int j = 2; j <= nz - 1; ++j) trans = trans + term[j];
    trans = trans + term[nz];
    trans = trans + term[nz + 1];
    tr[k] = trans;
}

// ... and their running sum in plane order, trans1 = trans1 + trans dp
__global__ void __launch_bounds__(64) ks_heat_sum(Geo g, double *tr)
{
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int k = 0; k <= g.np + 1; ++k) s = s + tr[k] * g.dp;
    tr[g.np + 2] = s;
}

// Local Nusselt numbers at mid-height, the output of sect360v2mg.f:318-325:
// a thread per plane k = 0..np, in the reference's operation order (L =
// ln(r0) from the host's C library, as the reference's dlog)
__global__ void __launch_bounds__(64) ks_local_nu(Geo g, const double *__restrict__ th, double rmax, double L,
                                                  double *__restrict__ inner, double *__restrict__ outer)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > g.np) return;
    const int nr = g.nr, jm = g.nz / 2;
    const double dr = g.dr;
    inner[k] = -(g.r0 * L * (-184.0 * X(th, 0, jm, k) + 225.0 * X(th, 1, jm, k) - 50.0 * X(th, 2, jm, k) +
                             9.0 * X(th, 3, jm, k)) / (60.0 * dr));
    outer[k] = -(rmax * L * (184.0 * X(th, nr + 1, jm, k) - 225.0 * X(th, nr, jm, k) + 50.0 * X(th, nr - 1, jm, k) -
                             9.0 * X(th, nr - 2, jm, k)) / (60.0 * dr));
}

// update, sect360v2mg.f:936-979: wall vorticity of the end walls, the
// cylinders and the barrier's faces, wall temperatures (no point is written
// twice) ...
__global__ void __launch_bounds__(256) ks_update_walls(Geo g, SF F)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ldi * g.ldj) return;
    const int j = t / g.ldi, i = t - j * g.ldi, k = blockIdx.y;
    const int nr = g.nr, nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp;
    const double *uel = F.uel, *vel = F.vel, *wel = F.wel;
    if (i >= 1 && i <= nr && k >= 1 && k <= np - 1) {
        if (j == 0) X(F.An, i, 0, k) = -(3.0 * X(vel, i, 1, k) - X(vel, i, 2, k) / 3.0) / dz;
        else if (j == nz) X(F.An, i, nz, k) = -(X(vel, i, nz - 1, k) / 3.0 - 3.0 * X(vel, i, nz, k)) / dz;
    }
    if (i >= 1 && i <= nr - 1 && k >= 1 && k <= np) {
        if (j == 0) X(F.Bn, i, 0, k) = 2.0 * X(uel, i, 1, k) / dz;
        else if (j == nz) X(F.Bn, i, nz, k) = -2.0 * X(uel, i, nz, k) / dz;
    }
    if (j >= 1 && j <= nz - 1 && k >= 1 && k <= np) {
        if (i == 0) X(F.Bn, 0, j, k) = -2.0 * X(wel, 1, j, k) / dr;
        else if (i == nr) X(F.Bn, nr, j, k) = 2.0 * X(wel, nr, j, k) / dr;
    }
    if (j >= 1 && j <= nz && k >= 1 && k <= np - 1) {
        if (i == 0) X(F.Cn, 0, j, k) = (3.0 * X(vel, 1, j, k) - X(vel, 2, j, k) / 3.0) / dr;
        else if (i == nr) X(F.Cn, nr, j, k) = (X(vel, nr - 1, j, k) / 3.0 - 3.0 * X(vel, nr, j, k)) / dr;
    }
    if (i >= 1 && i <= nr && j >= 1 && j <= nz) {
        if (k == 0) X(F.An, i, j, 0) = 1.0 / (g.r0 + ((double)i - 0.5) * dr) * (2.0 * X(wel, i, j, 1)) / dp;
        else if (k == np) X(F.An, i, j, np) = 1.0 / (g.r0 + ((double)i - 0.5) * dr) * (-2.0 * X(wel, i, j, np)) / dp;
    }
    if (i >= 1 && i <= nr - 1 && j >= 1 && j <= nz) {
        if (k == 0) X(F.Cn, i, j, 0) = 1.0 / (g.r0 + ((double)i - 0.5) * dr) * (-2.0 * X(uel, i, j, 1)) / dp;
        else if (k == np) X(F.Cn, i, j, np) = 1.0 / (g.r0 + ((double)i - 0.5) * dr) * (2.0 * X(uel, i, j, np)) / dp;
    }
    if (i == 0) X(F.thn, 0, j, k) = 0.0;
    else if (i == nr + 1) X(F.thn, nr + 1, j, k) = 1.0;
}

// ... then the time levels rotate (sect360v2mg.f:981-993): Ap <- A <- An by
// exchanging the buffers, and the old Ap buffer, which becomes An, receives
// a copy of the new A, so every level holds what the reference's copies
// leave (An keeps its values where no routine writes)
struct Copy4 {
    const double *src[4];
    double *dst[4];
};

__global__ void __launch_bounds__(256) ks_copy4(Copy4 C, size_t n)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int q = blockIdx.y;
    C.dst[q][t] = C.src[q][t];
}

// ---------------------------------------------------------------- Poisson

// psip <- psi on the range poisson1/2/3 copy (sect360v2mg.f:1427-1432,
// 1519-1524, 1619-1624)
__global__ void __launch_bounds__(256) ks_psi_pre(Geo g, int which, const double *__restrict__ psi,
                                                  double *__restrict__ psip)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ldi * g.ldj) return;
    const int j = t / g.ldi, i = t - j * g.ldi, k = blockIdx.y;
    const int ie = (which == 3) ? g.nr : g.nr + 1, je = (which == 1) ? g.nz : g.nz + 1;
    const int ke = (which == 2) ? g.np : g.np + 1;
    if (i > ie || j > je || k > ke) return;
    X(psip, i, j, k) = X(psi, i, j, k);
}

// the walls and mirrored ghosts poisson1/2/3 write after the interior
// (sect360v2mg.f:1439-1457, 1530-1547, 1630-1654); reads unknowns only
__global__ void __launch_bounds__(256) ks_psi_post(Geo g, int which, double *psi)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ldi * g.ldj) return;
    const int j = t / g.ldi, i = t - j * g.ldi, k = blockIdx.y;
    const int nr = g.nr, nz = g.nz, np = g.np;
    const int kk = (k == 0) ? 1 : (k == np + 1) ? np : k;
    if (which == 1) {
        if (j == 0 || j == nz) X(psi, i, j, k) = 0.0;
        else if (j <= nz - 1) {
            if (i == 0) X(psi, 0, j, k) = -X(psi, 1, j, kk);
            else if (i == nr + 1) X(psi, nr + 1, j, k) = -X(psi, nr, j, kk);
            else if (k == 0) X(psi, i, j, 0) = -X(psi, i, j, 1);
            else if (k == np + 1) X(psi, i, j, np + 1) = -X(psi, i, j, np);
        }
    } else if (which == 2) {
        if (k > np) return;
        if (k == 0 || k == np || i == 0 || i == nr + 1 || j == 0 || j == nz + 1) X(psi, i, j, k) = 0.0;
    } else {
        if (i > nr) return;
        if (i == 0 || i == nr) X(psi, i, j, k) = 0.0;
        else if (j == 0) X(psi, i, 0, k) = -X(psi, i, 1, kk);
        else if (j == nz + 1) X(psi, i, nz + 1, k) = -X(psi, i, nz, kk);
        else if (k == 0) X(psi, i, j, 0) = -X(psi, i, j, 1);
        else if (k == np + 1) X(psi, i, j, np + 1) = -X(psi, i, j, np);
    }
}

// 1. S[q][u] = sum_k W[q][k] rhs(i, j, k+1), u = (j-1) i1 + (i-1)
struct PhiFwd {
    const double *W;
    int M, N, K;
    const double *rhs;
    double *S;
    int i1, ldi, ldj;
    __device__ double a(int, int m, int k) const { return W[(size_t)m * K + k]; }
    __device__ double b(int, int k, int n) const
    {
        const int j = n / i1, i = n - j * i1;
        return rhs[((size_t)(k + 1) * ldj + (size_t)(j + 1)) * ldi + (size_t)(i + 1)];
    }
    __device__ void c(int, int m, int n, double v) const { S[(size_t)m * N + n] = v; }
};

// 2. T[q][i][l] = sum_j S[q][j][i] Vinv[l][j]
struct ZFwd {
    const double *Vinv;
    int M, N, K;
    const double *S;
    double *T;
    int mt;
    __device__ double a(int q, int m, int k) const { return S[(size_t)q * mt + (size_t)k * M + m]; }
    __device__ double b(int, int k, int n) const { return Vinv[(size_t)n * K + k]; }
    __device__ void c(int q, int m, int n, double v) const { T[(size_t)q * mt + (size_t)m * N + n] = v; }
};

// 4. S[q][j][i] = sum_l V[j][l] T[q][i][l]
struct ZInv {
    const double *V;
    int M, N, K;
    const double *T;
    double *S;
    int mt;
    __device__ double a(int, int m, int k) const { return V[(size_t)m * K + k]; }
    __device__ double b(int q, int k, int n) const { return T[(size_t)q * mt + (size_t)n * K + k]; }
    __device__ void c(int q, int m, int n, double v) const { S[(size_t)q * mt + (size_t)m * N + n] = v; }
};

// 5. psi(i, j, k+1) = sum_q Wi[k][q] S[q][u]
struct PhiInv {
    const double *Wi;
    int M, N, K;
    const double *S;
    double *psi;
    int i1, ldi, ldj;
    __device__ double a(int, int m, int k) const { return Wi[(size_t)m * K + k]; }
    __device__ double b(int, int k, int n) const { return S[(size_t)k * N + n]; }
    __device__ void c(int, int m, int n, double v) const
    {
        const int j = n / i1, i = n - j * i1;
        psi[((size_t)(m + 1) * ldj + (size_t)(j + 1)) * ldi + (size_t)(i + 1)] = v;
    }
};

// 3. per (phi mode q, z eigenvalue l) the r system (R + lphi(q) diag(1/(dp
// r)^2) + lz(l)) x = y, x overwriting y in T (layout [q][i][l]); row 1 has an
// extra coefficient e1 two points inward, removed with row 2, row i1 one
// (en) removed with the eliminated row i1 - 2.  Thomas factors go to cp.
__global__ void __launch_bounds__(64) ks_rsolve(double *__restrict__ T, double *__restrict__ cp,
                                                 const double *__restrict__ rt, const double *__restrict__ lz,
                                                 const double *__restrict__ lphi, int i1, int j1, int k1, double e1,
                                                 double en)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k1 * j1) return;
    const int q = t / j1, l = t - q * j1;
    const size_t mt = (size_t)i1 * j1;
    double *__restrict__ x = T + (size_t)q * mt + l;
    double *__restrict__ c = cp + (size_t)q * mt + l;
    const double *sub = rt, *dia = rt + i1, *sup = rt + 2 * i1, *ir2 = rt + 3 * i1;
    const double mu = lphi[q], lam = lz[l];
    auto diag = [&](int i) { return dia[i] + mu * ir2[i] + lam; };
    double d0 = diag(0), s0 = sup[0], y = x[0];
    if (e1 != 0.0) {
        const double f = e1 / sup[1];
        d0 -= f * sub[1];
        s0 -= f * diag(1);
        y -= f * x[j1];
    }
    double cprev = s0 / d0;
    y /= d0;
    c[0] = cprev;
    x[0] = y;
    // rows 1 .. i1-2 in groups of U with their loads issued ahead
    constexpr int U = 8;
    int i = 1;
    for (; i + U <= i1 - 1; i += U) {
        double rv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) rv[u] = x[(size_t)(i + u) * j1];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t o = (size_t)(i + u) * j1;
            const double a = sub[i + u];
            const double den = diag(i + u) - a * cprev;
            cprev = sup[i + u] / den;
            y = (rv[u] - a * y) / den;
            c[o] = cprev;
            x[o] = y;
        }
    }
    for (; i < i1; ++i) {
        const size_t o = (size_t)i * j1;
        double a = sub[i], rv = x[o];
        if (i == i1 - 1 && en != 0.0) {
            const size_t o2 = (size_t)(i - 2) * j1;
            a -= en * c[o2];
            rv -= en * x[o2];
        }
        const double den = diag(i) - a * cprev;
        cprev = sup[i] / den;
        y = (rv - a * y) / den;
        c[o] = cprev;
        x[o] = y;
    }
    i = i1 - 2;
    for (; i - U + 1 >= 0; i -= U) {
        double fu[U], xu[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t o = (size_t)(i - u) * j1;
            fu[u] = c[o];
            xu[u] = x[o];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            y = xu[u] - fu[u] * y;
            x[(size_t)(i - u) * j1] = y;
        }
    }
    for (; i >= 0; --i) {
        const size_t o = (size_t)i * j1;
        y = x[o] - c[o] * y;
        x[o] = y;
    }
}

template <class T>
hipError_t upload(T **dst, const std::vector<T> &src)
{
    hipError_t e = hipMalloc(dst, src.size() * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
}

}  // namespace

void ranges(const Geo &g, int which, int &i1, int &j1, int &k1)
{
    i1 = (which == 3) ? g.nr - 1 : g.nr;
    j1 = (which == 1) ? g.nz - 1 : g.nz;
    k1 = (which == 2) ? g.np - 1 : g.np;
}

// r rows of the sector's w operator (sect360v2mg.f:1823-1838 without the z
// and phi terms): zero flux through both cylinders
static void r_rows_w(int nr, double r0, double dr, std::vector<double> &sub, std::vector<double> &dia,
                     std::vector<double> &sup)
{
    const double hhr = dr * dr;
    sub.assign(nr, 0.0);
    dia.assign(nr, 0.0);
    sup.assign(nr, 0.0);
    for (int i = 1; i <= nr; ++i) {
        const double rv = r0 + ((double)i - 0.5) * dr;
        double a = 1.0 / hhr - 0.5 / dr / rv, d = -2.0 / hhr, c = 1.0 / hhr + 0.5 / dr / rv;
        if (i == 1) { a = 0.0; d = -1.0 / hhr - 0.5 / dr / rv; }
        if (i == nr) { c = 0.0; d = -1.0 / hhr + 0.5 / dr / rv; }
        sub[i - 1] = a;
        dia[i - 1] = d;
        sup[i - 1] = c;
    }
}

// z rows of the sector's u operator (sect360v2mg.f:2240-2246), scaled by
// dz^2: zero flux through both end walls
static std::vector<double> z_matrix_u(int j1)
{
    std::vector<double> M((size_t)j1 * j1, 0.0);
    for (int j = 0; j < j1; ++j) {
        M[(size_t)j * j1 + j] = (j == 0 || j == j1 - 1) ? -1.0 : -2.0;
        if (j > 0) M[(size_t)j * j1 + j - 1] = 1.0;
        if (j < j1 - 1) M[(size_t)j * j1 + j + 1] = 1.0;
    }
    return M;
}

int spoisson_setup(SPoisson &P, const Geo &g)
{
    using vv::set_error;
    const double pi = std::acos(-1.0);
    const double hhz = g.dz * g.dz, hhp = g.dp * g.dp;
    size_t smax = 0, cmax = 0;
    P.fft = vvt::ytrans_fits(g.np);
    if (P.fft) {
        const int st = vvt::ytrans_setup(P.yt, g.np, g.ldi, g.ldj);
        if (st) return st;
    }
    for (int w = 1; w <= 3; ++w) {
        SOp &O = P.op[w - 1];
        ranges(g, w, O.i1, O.j1, O.k1);
        O.m = O.i1 * O.j1;
        const int k1 = O.k1, np = g.np;
        // phi eigenvectors: w, u cos(q pi (p - 1/2)/np), q = 0..np-1, over the
        // planes p = 1..np; v sin((q+1) pi p/np), q = 0..np-2, p = 1..np-1
        std::vector<double> W((size_t)k1 * k1), Wi((size_t)k1 * k1), lphi(k1);
        for (int q = 0; q < k1; ++q) {
            const int m = (w == 2) ? q + 1 : q;
            lphi[q] = 2.0 * std::cos(pi * m / np) - 2.0;
            const double wgt = (w == 2 || m > 0) ? 2.0 / np : 1.0 / np;
            for (int p = 1; p <= k1; ++p) {
                const double e = (w == 2) ? std::sin(pi * (double)m * p / np) : std::cos(pi * (double)m * (p - 0.5) / np);
                W[(size_t)q * k1 + (p - 1)] = e;
                Wi[(size_t)(p - 1) * k1 + q] = wgt * e;
            }
        }
        std::vector<double> sub, dia, sup;
        O.e1 = O.en = 0.0;
        if (w == 1) r_rows_w(g.nr, g.r0, g.dr, sub, dia, sup);
        else vv3::r_rows(w, g.nr, g.r0, g.dr, sub, dia, sup, O.e1, O.en);
        std::vector<double> rt((size_t)4 * O.i1);
        for (int i = 0; i < O.i1; ++i) {
            const double rv = (w == 3) ? (double)(i + 1) * g.dr + g.r0 : g.r0 + ((double)(i + 1) - 0.5) * g.dr;
            rt[i] = sub[i];
            rt[O.i1 + i] = dia[i];
            rt[2 * O.i1 + i] = sup[i];
            rt[3 * O.i1 + i] = 1.0 / hhp / (rv * rv);
        }
        std::vector<double> lam, V, Vinv;
        const std::vector<double> Z = (w == 3) ? z_matrix_u(O.j1) : vv3::z_matrix(w, O.j1);
        if (!vv3::eigen_banded(Z, O.j1, lam, V, Vinv)) {
            set_error("vvs_create: z eigen-decomposition of a Poisson operator failed");
            return VV_EINVAL;
        }
        for (double &v : lam) v /= hhz;
        if (!P.fft) {
            VV_HIP(upload(&O.W, W));
            VV_HIP(upload(&O.Wi, Wi));
        }
        VV_HIP(upload(&O.lphi, lphi));
        std::vector<double> Fz, Gz;
        O.zsplit = vv3::zsplit_wanted(O.j1) && 2 * k1 <= 65535 && vv3::z_split(O.j1, lam, V, Vinv, Fz, Gz);
        VV_HIP(upload(&O.V, O.zsplit ? Gz : V));
        VV_HIP(upload(&O.Vinv, O.zsplit ? Fz : Vinv));
        VV_HIP(upload(&O.lz, lam));
        VV_HIP(upload(&O.rt, rt));
        smax = std::max(smax, (size_t)k1 * O.m);
        cmax = std::max(cmax, (size_t)k1 * O.m);
    }
    VV_HIP(hipMalloc(&P.S, smax * sizeof(double)));
    VV_HIP(hipMalloc(&P.T, smax * sizeof(double)));
    VV_HIP(hipMalloc(&P.cp, cmax * sizeof(double)));
    // the second set is optional: without it the solves run in turn
    if (hipMalloc(&P.S2, smax * sizeof(double)) != hipSuccess || hipMalloc(&P.T2, smax * sizeof(double)) != hipSuccess ||
        hipMalloc(&P.cp2, cmax * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        for (double **q : {&P.S2, &P.T2, &P.cp2})
            if (*q) { (void)hipFree(*q); *q = nullptr; }
    }
    return VV_OK;
}

void spoisson_free(SPoisson &P)
{
    for (double *p : {P.S, P.T, P.cp, P.S2, P.T2, P.cp2})
        if (p) (void)hipFree(p);
    P.S = P.T = P.cp = P.S2 = P.T2 = P.cp2 = nullptr;
    vvt::ytrans_free(P.yt);
    for (auto &O : P.op) {
        for (double *p : {O.W, O.Wi, O.lphi, O.V, O.Vinv, O.lz, O.rt})
            if (p) (void)hipFree(p);
        O = SOp();
    }
}

int spoisson_launch(const SPoisson &P, const Geo &g, int which, double *psi, double *psip, const double *rhs,
                    hipStream_t s, bool second)
{
    const SOp &O = P.op[which - 1];
    double *S = second ? P.S2 : P.S, *T = second ? P.T2 : P.T, *cp = second ? P.cp2 : P.cp;
    const int i1 = O.i1, j1 = O.j1, k1 = O.k1, mt = O.m;
    const dim3 pg((unsigned)((g.ldi * g.ldj + 255) / 256), (unsigned)(g.np + 2));
    hipLaunchKernelGGL(ks_psi_pre, pg, dim3(256), 0, s, g, which, (const double *)psi, psip);
    if (P.fft) vvt::ytrans_fwd(P.yt, which == 2, rhs, S, i1, mt, s);
    else vv::gemm_f64<true, false>(PhiFwd{O.W, k1, mt, k1, rhs, S, i1, g.ldi, g.ldj}, 1, s);
    if (O.zsplit) vv::zsplit_fwd(O.Vinv, S, T, i1, j1, k1, s);
    else vv::gemm_f64<false, true>(ZFwd{O.Vinv, i1, j1, j1, S, T, mt}, k1, s);
    hipLaunchKernelGGL(ks_rsolve, dim3((unsigned)((k1 * j1 + 63) / 64)), dim3(64), 0, s, T, cp,
                       (const double *)O.rt, (const double *)O.lz, (const double *)O.lphi, i1, j1, k1, O.e1, O.en);
    if (O.zsplit) vv::zsplit_inv(O.V, T, S, i1, j1, k1, s);
    else vv::gemm_f64<true, true>(ZInv{O.V, j1, i1, j1, T, S, mt}, k1, s);
    if (P.fft) vvt::ytrans_inv(P.yt, which == 2, S, psi, i1, mt, s);
    else vv::gemm_f64<true, false>(PhiInv{O.Wi, k1, mt, k1, S, psi, i1, g.ldi, g.ldj}, 1, s);
    hipLaunchKernelGGL(ks_psi_post, pg, dim3(256), 0, s, g, which, psi);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

}  // namespace vvs

// ---------------------------------------------------------------- C ABI

struct vvs_run {
    // every C entry point on a run holds this lock
    std::recursive_mutex mu;
    vv::PinnedStage stage;   // host copies (vvs_put / vvs_get)
    vv3_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    vvs::Geo g;
    size_t nf = 0;                       // doubles per field: ldi ldj (np+2)
    double *f[VV3_NFIELD] = {nullptr};
    vvs::SPoisson pois;
    vvs::SK K;                           // the explicit routines' constant divisors
    hipStream_t stream2 = nullptr;       // heat and v's right-hand side and solve, beside w's and u's
    hipEvent_t fork = nullptr, join = nullptr;
    double *heat = nullptr;              // plane fluxes k = 0..np+1, then trans1
    double *nu = nullptr;                // local Nusselt numbers: inner, outer (np+1 each)
};

using namespace vvs;

static SF sfields(vvs_run *r)
{
    SF F;
    F.A = r->f[VV3_A]; F.B = r->f[VV3_B]; F.C = r->f[VV3_C];
    F.Ap = r->f[VV3_AP]; F.Bp = r->f[VV3_BP]; F.Cp = r->f[VV3_CP];
    F.th = r->f[VV3_TH]; F.thp = r->f[VV3_THP];
    F.uel = r->f[VV3_UEL]; F.vel = r->f[VV3_VEL]; F.wel = r->f[VV3_WEL];
    F.An = r->f[VV3_AN]; F.Bn = r->f[VV3_BN]; F.Cn = r->f[VV3_CN]; F.thn = r->f[VV3_THN];
    F.urh = r->f[VV3_URH]; F.vrh = r->f[VV3_VRH]; F.wrh = r->f[VV3_WRH];
    return F;
}

static dim3 pgrid(const vvs_run *r) { return dim3((unsigned)((r->g.ldi * r->g.ldj + 255) / 256), (unsigned)(r->g.np + 2)); }

static int launch_explicit(vvs_run *r, int mask)
{
    const SF F = sfields(r);
    const dim3 gr = pgrid(r);
    switch (mask) {
    case 1: hipLaunchKernelGGL(ks_explicit<1>, gr, dim3(256), 0, r->stream, r->g, F, r->K); break;
    case 2: hipLaunchKernelGGL(ks_explicit<2>, gr, dim3(256), 0, r->stream, r->g, F, r->K); break;
    case 4: hipLaunchKernelGGL(ks_explicit<4>, gr, dim3(256), 0, r->stream, r->g, F, r->K); break;
    case 7: hipLaunchKernelGGL(ks_explicit<7>, gr, dim3(256), 0, r->stream, r->g, F, r->K); break;
    case 8: hipLaunchKernelGGL(ks_explicit<8>, gr, dim3(256), 0, r->stream, r->g, F, r->K); break;
    default: hipLaunchKernelGGL(ks_explicit<15>, gr, dim3(256), 0, r->stream, r->g, F, r->K); break;
    }
    VV_HIP(hipGetLastError());
    return VV_OK;
}

static int launch_heat(vvs_run *r, hipStream_t s)
{
    hipLaunchKernelGGL(ks_heat_planes, dim3((unsigned)(r->g.np + 2)), dim3(256), (size_t)(r->g.nz + 2) * sizeof(double),
                       s, r->g, (const double *)r->f[VV3_TH], r->heat);
    hipLaunchKernelGGL(ks_heat_sum, dim3(1), dim3(64), 0, s, r->g, r->heat);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

static int launch_rhs(vvs_run *r, int which, hipStream_t s)
{
    const SF F = sfields(r);
    const dim3 gr = pgrid(r);
    if (which == 1) hipLaunchKernelGGL(ks_rhs<1>, gr, dim3(256), 0, s, r->g, F, r->K);
    else if (which == 2) hipLaunchKernelGGL(ks_rhs<2>, gr, dim3(256), 0, s, r->g, F, r->K);
    else hipLaunchKernelGGL(ks_rhs<3>, gr, dim3(256), 0, s, r->g, F, r->K);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

static int launch_poisson(vvs_run *r, int which, hipStream_t s, bool second = false)
{
    static const int PSI[3][3] = {{VV3_WEL, VV3_WELP, VV3_WRH}, {VV3_VEL, VV3_VELP, VV3_VRH},
                                  {VV3_UEL, VV3_UELP, VV3_URH}};
    const int *p = PSI[which - 1];
    return spoisson_launch(r->pois, r->g, which, r->f[p[0]], r->f[p[1]], r->f[p[2]], s, second);
}

static int launch_update(vvs_run *r)
{
    const SF F = sfields(r);
    hipLaunchKernelGGL(ks_update_walls, pgrid(r), dim3(256), 0, r->stream, r->g, F);
    static const int LV[4][3] = {{VV3_AP, VV3_A, VV3_AN}, {VV3_BP, VV3_B, VV3_BN}, {VV3_CP, VV3_C, VV3_CN},
                                 {VV3_THP, VV3_TH, VV3_THN}};
    Copy4 C;
    for (int q = 0; q < 4; ++q) {
        double *old_p = r->f[LV[q][0]];
        r->f[LV[q][0]] = r->f[LV[q][1]];
        r->f[LV[q][1]] = r->f[LV[q][2]];
        r->f[LV[q][2]] = old_p;
        C.src[q] = r->f[LV[q][1]];
        C.dst[q] = r->f[LV[q][2]];
    }
    hipLaunchKernelGGL(ks_copy4, dim3((unsigned)((r->nf + 255) / 256), 4), dim3(256), 0, r->stream, C, r->nf);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

// host (ldi x ldj x (np+2)) <-> device
static int copys(vvs_run *r, int field, double *host, int ldi, int ldj, bool to_dev)
{
    const int di = r->g.ldi, dj = r->g.ldj, nk = r->g.np + 2;
    if (ldi == 0) ldi = di;
    if (ldj == 0) ldj = dj;
    if (ldi < di || ldj < dj) {
        vv::set_error("vvs_put/vvs_get: ldi must be >= nr+2 and ldj >= nz+2");
        return VV_EINVAL;
    }
    VV_HIP(hipSetDevice(r->device));
    double *dev = r->f[field];
    const size_t words = (size_t)di * dj * nk;
    double *st = r->stage.get(words);
    if (!st) { vv::set_error("vvs_put/vvs_get: pinned staging buffer"); return VV_ENOMEM; }
    auto rows = [&](bool pack) {
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < dj; ++j) {
                double *h = host + ((size_t)k * ldj + j) * ldi;
                double *s = st + ((size_t)k * dj + j) * di;
                if (pack) std::memcpy(s, h, (size_t)di * sizeof(double));
                else std::memcpy(h, s, (size_t)di * sizeof(double));
            }
    };
    if (to_dev) {
        rows(true);
        VV_HIP(hipMemcpyAsync(dev, st, words * sizeof(double), hipMemcpyHostToDevice, r->stream));
        VV_HIP(hipStreamSynchronize(r->stream));
    } else {
        VV_HIP(hipMemcpyAsync(st, dev, words * sizeof(double), hipMemcpyDeviceToHost, r->stream));
        VV_HIP(hipStreamSynchronize(r->stream));
        rows(false);
    }
    return VV_OK;
}

extern "C" {

int vvs_create(const vv3_config *cfg, int device, vvs_run **out)
{
    using vv::set_error;
    if (!cfg || !out) { set_error("vvs_create: null argument"); return VV_EINVAL; }
    *out = nullptr;
    if (cfg->nr < 4 || cfg->nz < 4 || cfg->nz > VVS_NZ_MAX || cfg->np < 4 || cfg->np > VVS_NP_MAX) {
        set_error("vvs_create: need nr >= 4, 4 <= nz <= VVS_NZ_MAX and 4 <= np <= VVS_NP_MAX");
        return VV_EINVAL;
    }
    if (!(cfg->dt > 0) || !(cfg->rmax > cfg->r0) || !(cfg->r0 > 0) || !(cfg->zmax > 0) || !(cfg->Ptl > 0) ||
        cfg->Re == 0.0) {
        set_error("vvs_create: bad physical parameters");
        return VV_EINVAL;
    }
    VV_HIP(hipSetDevice(device));
    vvs_run *r = new vvs_run();
    r->cfg = *cfg;
    r->device = device;
    Geo &g = r->g;
    g.nr = cfg->nr; g.nz = cfg->nz; g.np = cfg->np; g.ldi = cfg->nr + 2; g.ldj = cfg->nz + 2;
    // sect360v2mg.f:84-86, unless the caller passes its own values
    g.dr = (cfg->dr > 0.0) ? cfg->dr : (cfg->rmax - cfg->r0) / cfg->nr;
    g.dz = (cfg->dz > 0.0) ? cfg->dz : cfg->zmax / cfg->nz;
    g.dp = (cfg->dp > 0.0) ? cfg->dp : 2.0 * std::acos(-1.0) / cfg->np;
    g.dt = cfg->dt; g.Ra = cfg->Ra; g.Re = cfg->Re; g.Ptl = cfg->Ptl; g.Gr = cfg->Gr; g.r0 = cfg->r0;
    g.model = 0;
    {   // the divisors as the reference's expressions form them, with RN(1/b)
        auto dv = [](double v) { return Dv{v, 1.0 / v}; };
        SK &K = r->K;
        K.dr = dv(g.dr); K.dz = dv(g.dz); K.dp = dv(g.dp);
        K.dr2 = dv(g.dr * g.dr); K.dz2 = dv(g.dz * g.dz); K.dp2 = dv(g.dp * g.dp);
        K.three = dv(3.0); K.ptl = dv(g.Ptl); K.ptl2 = dv(2.0 * g.Ptl); K.ptldt = dv(g.Ptl * g.dt);
        K.dt = dv(g.dt); K.dp2x = dv(2.0 * g.dp);
        K.re = Dvs{g.Re, 1.0 / g.Re};
    }
    r->nf = (size_t)g.ldi * g.ldj * (cfg->np + 2);
    auto fail = [&](int s) { vvs_destroy(r); return s; };
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) return fail(VV_EHIP);
    for (int q = 0; q < VV3_NFIELD; ++q) {
        if (hipMalloc(&r->f[q], r->nf * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
        if (hipMemsetAsync(r->f[q], 0, r->nf * sizeof(double), r->stream) != hipSuccess) return fail(VV_EHIP);
    }
    if (hipMalloc(&r->heat, (size_t)(cfg->np + 3) * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
    if (hipMalloc(&r->nu, 2 * (size_t)(cfg->np + 1) * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
    if (hipMemsetAsync(r->heat, 0, (size_t)(cfg->np + 3) * sizeof(double), r->stream) != hipSuccess)
        return fail(VV_EHIP);
    {
        const int st = spoisson_setup(r->pois, g);
        if (st) return fail(st);
    }
    if (r->pois.S2 && (hipStreamCreateWithFlags(&r->stream2, hipStreamNonBlocking) != hipSuccess ||
                       hipEventCreateWithFlags(&r->fork, hipEventDisableTiming) != hipSuccess ||
                       hipEventCreateWithFlags(&r->join, hipEventDisableTiming) != hipSuccess))
        return fail(VV_EHIP);
    if (hipStreamSynchronize(r->stream) != hipSuccess) return fail(VV_EHIP);
    *out = r;
    return VV_OK;
}

int vvs_destroy(vvs_run *r)
{
    if (!r) return VV_OK;
    {
        std::lock_guard<std::recursive_mutex> lk(r->mu);
        (void)hipSetDevice(r->device);
        if (r->stream) (void)hipStreamSynchronize(r->stream);
        for (int q = 0; q < VV3_NFIELD; ++q)
            if (r->f[q]) (void)hipFree(r->f[q]);
        if (r->stream2) (void)hipStreamSynchronize(r->stream2);
        spoisson_free(r->pois);
        if (r->heat) (void)hipFree(r->heat);
        for (hipEvent_t e : {r->fork, r->join})
            if (e) (void)hipEventDestroy(e);
        if (r->stream2) (void)hipStreamDestroy(r->stream2);
        if (r->nu) (void)hipFree(r->nu);
        if (r->stream) (void)hipStreamDestroy(r->stream);
    }
    delete r;
    return VV_OK;
}

int vvs_put(vvs_run *r, int field, const double *host, int ldi, int ldj)
{
    if (!r || !host || field < 0 || field >= VV3_NFIELD) { vv::set_error("vvs_put: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    return copys(r, field, const_cast<double *>(host), ldi, ldj, true);
}

int vvs_get(vvs_run *r, int field, double *host, int ldi, int ldj)
{
    if (!r || !host || field < 0 || field >= VV3_NFIELD) { vv::set_error("vvs_get: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    return copys(r, field, host, ldi, ldj, false);
}

int vvs_vorticity(vvs_run *r, int which)
{
    if (!r || which < 0 || which > 3) { vv::set_error("vvs_vorticity: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    VVS_TRY(launch_explicit(r, which == 0 ? 7 : 1 << (which - 1)));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vvs_temperature(vvs_run *r)
{
    if (!r) { vv::set_error("vvs_temperature: null run"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    VVS_TRY(launch_explicit(r, 8));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vvs_heat(vvs_run *r, double *trans1)
{
    if (!r || !trans1) { vv::set_error("vvs_heat: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    VVS_TRY(launch_heat(r, r->stream));
    VV_HIP(hipMemcpyAsync(trans1, r->heat + r->g.np + 2, sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vvs_local_nu(vvs_run *r, double *inner, double *outer)
{
    if (!r || !inner || !outer) { vv::set_error("vvs_local_nu: null argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    const int np = r->g.np;
    hipLaunchKernelGGL(ks_local_nu, dim3((unsigned)((np + 1 + 63) / 64)), dim3(64), 0, r->stream, r->g,
                       (const double *)r->f[VV3_TH], r->cfg.rmax, std::log(r->cfg.r0), r->nu, r->nu + np + 1);
    VV_HIP(hipGetLastError());
    VV_HIP(hipMemcpyAsync(inner, r->nu, (np + 1) * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipMemcpyAsync(outer, r->nu + np + 1, (np + 1) * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vvs_rhs(vvs_run *r, int which)
{
    if (!r || which < 1 || which > 3) { vv::set_error("vvs_rhs: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    VVS_TRY(launch_rhs(r, which, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vvs_poisson(vvs_run *r, int which)
{
    if (!r || which < 1 || which > 3) { vv::set_error("vvs_poisson: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    VVS_TRY(launch_poisson(r, which, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vvs_update(vvs_run *r)
{
    if (!r) { vv::set_error("vvs_update: null run"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    VVS_TRY(launch_update(r));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vvs_step(vvs_run *r, int n, double *trans1)
{
    if (!r || n < 0) { vv::set_error("vvs_step: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    for (int s = 0; s < n; ++s) {
        VVS_TRY(launch_explicit(r, 15));
        if (r->stream2) {
            // the heat flux (it reads th) and v's right-hand side (An and Cn
            // only, sect360v2mg.f:1034-1063) and solve on the second stream
            // with the second work arrays, w's and u's (u's right-hand side
            // reads the new w) on the first; the update waits for both
            VV_HIP(hipEventRecord(r->fork, r->stream));
            VV_HIP(hipStreamWaitEvent(r->stream2, r->fork, 0));
            VVS_TRY(launch_heat(r, r->stream2));
            VVS_TRY(launch_rhs(r, 2, r->stream2));
            VVS_TRY(launch_poisson(r, 2, r->stream2, true));
            VV_HIP(hipEventRecord(r->join, r->stream2));
            VVS_TRY(launch_rhs(r, 1, r->stream));
            VVS_TRY(launch_poisson(r, 1, r->stream));
            VVS_TRY(launch_rhs(r, 3, r->stream));
            VVS_TRY(launch_poisson(r, 3, r->stream));
            VV_HIP(hipStreamWaitEvent(r->stream, r->join, 0));
        } else {
            VVS_TRY(launch_heat(r, r->stream));
            for (int w = 1; w <= 3; ++w) {
                VVS_TRY(launch_rhs(r, w, r->stream));
                VVS_TRY(launch_poisson(r, w, r->stream));
            }
        }
        VVS_TRY(launch_update(r));
    }
    if (trans1)
        VV_HIP(hipMemcpyAsync(trans1, r->heat + r->g.np + 2, sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

}  // extern "C"
