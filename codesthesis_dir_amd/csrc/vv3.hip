// vv3.hip -- the 3D annulus time step of rot3dmg.f on gfx950, and of its
// stationary horizontal-annulus variant rot3dmghor.f (model VV3_MODEL_HOR:
// gravity in the r-theta plane instead of the rotation terms, swapped wall
// temperatures, heat flux through both cylinders near mid-axis).
//
// Explicit part: one thread per point for Avorticity / Bvorticity /
// Cvorticity / temperature, the Poisson right-hand sides and the wall
// update, evaluating the reference's expressions in its operation order
// (the library is compiled with -ffp-contract=off), so the results equal
// the reference bit for bit.
//
// The three Poisson solves are exact separable solves (vv3_poisson.hip).

#include "vv3_internal.h"
#include "vv_reduce.h"

#include <cmath>
#include <cstring>
#include <vector>

struct vv3_run {
    // every C entry point on a run holds this lock, so calls on one run from
    // several threads are serialised
    std::recursive_mutex mu;
    vv::PinnedStage stage;   // host copies (vv3_put / vv3_get)
    vv3_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    int nr = 0, nz = 0, np = 0, ldi = 0, ldj = 0;
    double dr = 0, dz = 0, dp = 0;
    size_t nf = 0;                       // doubles per field
    double *f[VV3_NFIELD] = {nullptr};   // device fields (rotated by pointer swap)
    vv3::Poisson3 pois;                  // the three Poisson operators
    hipStream_t stream2 = nullptr;       // heat and v's right-hand side and solve, beside w's and u's
    hipEvent_t fork = nullptr, join = nullptr;
    double *heat = nullptr;              // flux terms (nz+2) x np, then the totals trans1, trans2
    double *trig = nullptr;              // VV3_MODEL_HOR: sin, cos of each plane's azimuth (2 x (np+1))
    double *kepart = nullptr;            // kinetic energy: partials [KE3_MAXB], then the total
    double *dgbuf = nullptr;             // diagnostics, allocated at first use: plane sums (np),
                                         // cos / sin table (2 x MODES_MAX x np), the 2 x MODES_MAX
                                         // mode integrals, local Nusselt numbers (2 x (np+1))
};

constexpr int MODES_MAX = 16;

constexpr int KE3_MAXB = 1024;

namespace vv3 {

namespace {

constexpr int WG = 256;

// field element (i, j, k); fields stay below 2^32 elements, so offsets are
// 32-bit (uniform base + per-lane offset addressing)
#define X(a, i, j, k) (a)[(unsigned)(((k) * g.ldj + (j)) * g.ldi + (i))]

struct Fields {
    const double *A, *B, *C, *Ap, *Bp, *Cp, *th, *thp, *uel, *vel, *wel;
    double *An, *Bn, *Cn, *thn;
};

// A positive divisor with its correctly rounded reciprocal.  x / b is
// evaluated as q0 = x rb, t = fma(b, q0, -x) = b q0 - x (exact),
// q = fma(-t, rb, q0): for rb = RN(1/b) this is the correctly rounded
// quotient (Markstein's correction), so the results equal those of IEEE
// division.  Forming -t rather than x - b q0 gives a zero quotient the sign
// IEEE division gives it when b > 0 (t = +0, -t = -0 for x = +-0).
struct Rc {
    double v, r;
};
__device__ __forceinline__ double dv(double x, Rc b)
{
    const double q0 = x * b.r;
    const double t = __builtin_fma(b.v, q0, -x);
    return __builtin_fma(-t, b.r, q0);
}
// the same for a divisor of either sign (the zero case kept apart)
__device__ __forceinline__ double dvs(double x, Rc b)
{
    const double q0 = x * b.r;
    const double e = __builtin_fma(-b.v, q0, x);
    return (e == 0.0) ? q0 : __builtin_fma(e, b.r, q0);
}

// constants of the explicit routines, each evaluated on the host exactly as
// the reference's expression writes it
struct Kc {
    Rc dr, dz, dp, dr2, dz2, dp2, ptl, ptl2, ptldt, three, dt;   // positive
    Rc re;                                                       // either sign
    double cnr_int, cnz_int;   // -1/(dr dr), -1/(dz dz)
    double cnr_w, cnz_w;       // -2.5/(dr dr), -2.5/(dz dz)
    double cnz_t;              // -4/(dz dz)
    double hdr;                // 0.5/dr
    double m1dp2;              // -1/(dp dp)
    double ptldt_inv, dt_inv;  // 1/(Ptl dt), 1/dt
    double grptl;              // Gr Ptl
    const double *trig;        // VV3_MODEL_HOR: sin(theta_k) at [k], cos at [np+1+k]
};

// per-point radius with 1/r and 1/r^2
struct Rad {
    Rc r, r2, r2h;   // r, r*r, 2*(r*r)
    __device__ explicit Rad(double rv)
    {
        const double q = rv * rv, iq = 1.0 / q;
        r = Rc{rv, 1.0 / rv};
        r2 = Rc{q, iq};
        r2h = Rc{2.0 * q, 0.5 * iq};
    }
};

// The routines read the fields through a view: fv.x<field, di, dj, dk>() is
// the value at (i + di, j + dj, k + dk), k + dk taken periodically (km, ka).
// GView reads it from memory; MView (below) carries the values a marching
// column reuses from plane to plane.  Both give the same numbers, so the
// routines' results do not depend on the view.
constexpr int FA = 0, FB = 1, FC = 2, FTH = 3, FU = 4, FV = 5, FW = 6, FAP = 7, FBP = 8, FCP = 9, FTHP = 10;

template <int FLD>
__device__ __forceinline__ const double *fptr(const Fields &F)
{
    if constexpr (FLD == FA) return F.A;
    else if constexpr (FLD == FB) return F.B;
    else if constexpr (FLD == FC) return F.C;
    else if constexpr (FLD == FTH) return F.th;
    else if constexpr (FLD == FU) return F.uel;
    else if constexpr (FLD == FV) return F.vel;
    else if constexpr (FLD == FW) return F.wel;
    else if constexpr (FLD == FAP) return F.Ap;
    else if constexpr (FLD == FBP) return F.Bp;
    else if constexpr (FLD == FCP) return F.Cp;
    else return F.thp;
}

struct GView {
    const Fields &F;
    unsigned b[3];   // offset of (i, j) in the planes km, k, ka
    int li;          // row pitch
    template <int FLD, int DI, int DJ, int DK>
    __device__ double x() const
    {
        return fptr<FLD>(F)[b[DK + 1] + (unsigned)(DJ * li + DI)];
    }
};

#define XV(f, di, dj, dk) fv.template x<f, di, dj, dk>()

// Avorticity at (i, j, k), rot3dmg.f:359-465; M = 1: rot3dmghor.f:373-483
template <int M, class V>
__device__ double avort(const Geo &g, const Kc &K, const V &fv, int i, int j, int k)
{
    const int nr = g.nr, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, Ra = g.Ra, Re = g.Re;
    const Rad R(g.r0 + ((double)i - 0.5) * dr);
    const double u = (XV(FU, 0, 0, 0) + XV(FU, 0, 1, 0) + XV(FU, 0, 0, 1) + XV(FU, 0, 1, 1) + XV(FU, -1, 0, 0) +
                      XV(FU, -1, 1, 0) + XV(FU, -1, 0, 1) + XV(FU, -1, 1, 1)) * 0.125;
    const double v = (XV(FV, 0, 1, 0) + XV(FV, 0, 0, 0)) * 0.5;
    const double w = (XV(FW, 0, 0, 1) + XV(FW, 0, 0, 0)) * 0.5;
    const double tem = (XV(FTH, 0, 0, 0) + XV(FTH, 0, 0, 1) + XV(FTH, 0, 1, 0) + XV(FTH, 0, 1, 1)) * 0.25;
    const double BB = (XV(FB, 0, 0, 0) + XV(FB, -1, 0, 0) + XV(FB, 0, 0, 1) + XV(FB, -1, 0, 1)) * 0.25;
    const double CC = (XV(FC, 0, 0, 0) + XV(FC, 0, 1, 0) + XV(FC, -1, 0, 0) + XV(FC, -1, 1, 0)) * 0.25;
    const double Ac = XV(FA, 0, 0, 0);
    double Arra, Ara, Ar, cnr;
    if (i == 1) {
        const double Am = XV(FA, -1, 0, 0), A1 = XV(FA, 1, 0, 0);
        Arra = dv(3.2 * Am + 2.0 * A1 - 0.2 * XV(FA, 2, 0, 0), K.dr2);
        Ara = dv(dv(-4.0 * Am, K.three) + dv(A1, K.three), K.dr);
        Ar = dv(dv(-4.0 * Am, K.three) + Ac + dv(A1, K.three), K.dr);
        cnr = K.cnr_w + dv(K.hdr, R.r);
    } else if (i == nr) {
        const double Am = XV(FA, -1, 0, 0), A1 = XV(FA, 1, 0, 0);
        Arra = dv(3.2 * A1 + 2.0 * Am - 0.2 * XV(FA, -2, 0, 0), K.dr2);
        Ara = dv(dv(-Am, K.three) + dv(4.0 * A1, K.three), K.dr);
        Ar = dv(dv(-4.0 * Am, K.three) + Ac + dv(A1, K.three), K.dr);
        cnr = K.cnr_w - dv(K.hdr, R.r);
    } else {
        const double Am = XV(FA, -1, 0, 0), A1 = XV(FA, 1, 0, 0);
        Arra = dv(Am + A1, K.dr2);
        Ara = dv((A1 - Am) * 0.5, K.dr);
        Ar = Ara;
        cnr = K.cnr_int;
        if (fabs(dv(u * dr, K.ptl)) > 2) {          // hybrid upwind, rot3dmg.f:397-406
            if (u > 0) Ar = dv(Ac - Am, K.dr);
            else if (u < 0) Ar = dv(A1 - Ac, K.dr);
        }
    }
    const double uelr = dv((XV(FU, 0, 0, 0) + XV(FU, 0, 1, 0) + XV(FU, 0, 0, 1) + XV(FU, 0, 1, 1)) * 0.25 -
                           (XV(FU, -1, 0, 0) + XV(FU, -1, 1, 0) + XV(FU, -1, 0, 1) + XV(FU, -1, 1, 1)) * 0.25, K.dr);
    const double Azza = dv(XV(FA, 0, -1, 0) + XV(FA, 0, 1, 0), K.dz2);
    double Az = dv((XV(FA, 0, 1, 0) - XV(FA, 0, -1, 0)) * 0.5, K.dz);
    if (fabs(dv(w * dz, K.ptl)) > 2) {
        if (w > 0) Az = dv(Ac - XV(FA, 0, -1, 0), K.dz);
        else if (w < 0) Az = dv(XV(FA, 0, 1, 0) - Ac, K.dz);
    }
    const double uelz = dv((XV(FU, 0, 1, 0) + XV(FU, 0, 1, 1) + XV(FU, -1, 1, 0) + XV(FU, -1, 1, 1)) * 0.25 -
                           (XV(FU, 0, 0, 0) + XV(FU, 0, 0, 1) + XV(FU, -1, 0, 0) + XV(FU, -1, 0, 1)) * 0.25, K.dz);
    const double thz = dv((XV(FTH, 0, 1, 0) + XV(FTH, 0, 1, 1) - XV(FTH, 0, 0, 0) - XV(FTH, 0, 0, 1)) * 0.5, K.dz);
    const double thh = dv((XV(FTH, 0, 0, 1) + XV(FTH, 0, 1, 1) - XV(FTH, 0, 0, 0) - XV(FTH, 0, 1, 0)) * 0.5, K.dp);
    const double Ahha = dv(XV(FA, 0, 0, 1) + XV(FA, 0, 0, -1), K.dp2);
    double Ah = dv((XV(FA, 0, 0, 1) - XV(FA, 0, 0, -1)) * 0.5, K.dp);
    if (fabs(dv(v * dp * R.r.v, K.ptl)) > 2) {
        if (v > 0) Ah = dv(Ac - XV(FA, 0, 0, -1), K.dp);
        else if (v < 0) Ah = dv(XV(FA, 0, 0, 1) - Ac, K.dp);
    }
    const double Bh = dv((XV(FB, 0, 0, 1) + XV(FB, -1, 0, 1) - XV(FB, 0, 0, 0) - XV(FB, -1, 0, 0)) * 0.5, K.dp);
    const double uelh = dv((XV(FU, 0, 0, 1) + XV(FU, 0, 1, 1) + XV(FU, -1, 0, 1) + XV(FU, -1, 1, 1)) * 0.25 -
                           (XV(FU, 0, 0, 0) + XV(FU, 0, 1, 0) + XV(FU, -1, 0, 0) + XV(FU, -1, 1, 0)) * 0.25, K.dp);
    const double cnz = K.cnz_int;
    const double cnh = dv(K.m1dp2, R.r2);
    const double RHS1 = Arra + dv(Ara, R.r) + dv(Ahha, R.r2) + Azza - dv(2.0 * Bh, R.r2) + dv(Ac, K.ptldt);
    const double ULT = dv(u * Ar + dv(v * Ah, R.r) + w * Az - dv(uelh * BB, R.r) - uelz * CC, K.ptl);
    double RHS2;
    if constexpr (M == 1) RHS2 = K.trig[k] * g.Gr * g.Ptl * thz;   // sintheta*Gr*Ptl*thz
    else RHS2 = -2.0 * Re * uelz + dv(dvs(2.0 * Ra * (tem * uelz + thz * u), K.re), K.ptl) - dv(K.grptl, R.r) * thh;
    const double cf1 = K.ptldt_inv - dv(uelr, K.ptl2) + dv(1.0, R.r2h) - cnr - cnz - cnh;
    const double cf2 = dv(uelr, K.ptl2) - dv(1.0, R.r2h) + cnr + cnz + cnh;
    return (cf2 * XV(FAP, 0, 0, 0) - ULT - RHS2 + RHS1) / cf1;
}

// Bvorticity at (i, j, k), rot3dmg.f:495-590; M = 1: rot3dmghor.f:514-613
template <int M, class V>
__device__ double bvort(const Geo &g, const Kc &K, const V &fv, int i, int j, int k)
{
    const int np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, Ra = g.Ra, Re = g.Re;
    const Rad R(g.r0 + (double)i * dr);
    const double r = R.r.v;
    const double u = (XV(FU, 0, 0, 0) + XV(FU, 0, 1, 0)) * 0.5;
    const double v = (XV(FV, 0, 0, -1) + XV(FV, 0, 1, -1) + XV(FV, 0, 0, 0) + XV(FV, 0, 1, 0) + XV(FV, 1, 0, -1) +
                      XV(FV, 1, 1, -1) + XV(FV, 1, 0, 0) + XV(FV, 1, 1, 0)) * 0.125;
    const double w = (XV(FW, 0, 0, 0) + XV(FW, 1, 0, 0)) * 0.5;
    const double tem = (XV(FTH, 0, 1, 0) + XV(FTH, 1, 1, 0) + XV(FTH, 0, 0, 0) + XV(FTH, 1, 0, 0)) * 0.25;
    const double AA = (XV(FA, 0, 0, 0) + XV(FA, 1, 0, 0) + XV(FA, 0, 0, -1) + XV(FA, 1, 0, -1)) * 0.25;
    const double CC = (XV(FC, 0, 0, 0) + XV(FC, 0, 1, 0) + XV(FC, 0, 0, -1) + XV(FC, 0, 1, -1)) * 0.25;
    const double Bc = XV(FB, 0, 0, 0);
    const double Brra = dv(XV(FB, -1, 0, 0) + XV(FB, 1, 0, 0), K.dr2);
    const double Bra = dv((XV(FB, 1, 0, 0) - XV(FB, -1, 0, 0)) * 0.5, K.dr);
    double Br = Bra;
    if (fabs(dv(u * dr, K.ptl)) > 2) {
        if (u > 0) Br = dv(Bc - XV(FB, -1, 0, 0), K.dr);
        else if (u < 0) Br = dv(XV(FB, 1, 0, 0) - Bc, K.dr);
    }
    const double velr = dv((XV(FV, 1, 0, -1) + XV(FV, 1, 1, -1) + XV(FV, 1, 0, 0) + XV(FV, 1, 1, 0)) * 0.25 -
                           (XV(FV, 0, 0, -1) + XV(FV, 0, 1, -1) + XV(FV, 0, 0, 0) + XV(FV, 0, 1, 0)) * 0.25, K.dr);
    const double Bzza = dv(XV(FB, 0, -1, 0) + XV(FB, 0, 1, 0), K.dz2);
    double Bz = dv((XV(FB, 0, 1, 0) - XV(FB, 0, -1, 0)) * 0.5, K.dz);
    if (fabs(dv(w * dz, K.ptl)) > 2) {
        if (w > 0) Bz = dv(Bc - XV(FB, 0, -1, 0), K.dz);
        else if (w < 0) Bz = dv(XV(FB, 0, 1, 0) - Bc, K.dz);
    }
    const double velz = dv((XV(FV, 0, 1, -1) + XV(FV, 1, 1, -1) + XV(FV, 0, 1, 0) + XV(FV, 1, 1, 0)) * 0.25 -
                           (XV(FV, 0, 0, -1) + XV(FV, 1, 0, -1) + XV(FV, 0, 0, 0) + XV(FV, 1, 0, 0)) * 0.25, K.dz);
    const double thz = dv((XV(FTH, 0, 1, 0) + XV(FTH, 1, 1, 0) - XV(FTH, 0, 0, 0) - XV(FTH, 1, 0, 0)) * 0.5, K.dz);
    const double thr = dv((XV(FTH, 1, 1, 0) + XV(FTH, 1, 0, 0) - XV(FTH, 0, 1, 0) - XV(FTH, 0, 0, 0)) * 0.5, K.dr);
    const double Bhha = dv(XV(FB, 0, 0, 1) + XV(FB, 0, 0, -1), K.dp2);
    double Bh = dv((XV(FB, 0, 0, 1) - XV(FB, 0, 0, -1)) * 0.5, K.dp);
    if (fabs(dv(v * dp * r, K.ptl)) > 2) {
        if (v > 0) Bh = dv(Bc - XV(FB, 0, 0, -1), K.dp);
        else if (v < 0) Bh = dv(XV(FB, 0, 0, 1) - Bc, K.dp);
    }
    const double Ah = dv((XV(FA, 0, 0, 0) + XV(FA, 1, 0, 0) - XV(FA, 0, 0, -1) - XV(FA, 1, 0, -1)) * 0.5, K.dp);
    const double velh = dv((XV(FV, 0, 0, 0) + XV(FV, 1, 0, 0) + XV(FV, 0, 1, 0) + XV(FV, 1, 1, 0)) * 0.25 -
                           (XV(FV, 0, 0, -1) + XV(FV, 1, 0, -1) + XV(FV, 0, 1, -1) + XV(FV, 1, 1, -1)) * 0.25, K.dp);
    const double cnr = K.cnr_int;
    const double cnz = K.cnz_int;
    const double cnh = dv(K.m1dp2, R.r2);
    const double RHS1 = Brra + dv(Bra, R.r) + dv(Bhha, R.r2) + Bzza + dv(2.0 * Ah, R.r2) + dv(Bc, K.ptldt);
    const double ULT = dv(u * Br + dv(v * Bh, R.r) + w * Bz - (velr - dv(v, R.r)) * AA - velz * CC, K.ptl);
    double RHS2;
    if constexpr (M == 1) RHS2 = K.trig[np + 1 + k] * g.Gr * g.Ptl * thz;   // costheta*Gr*Ptl*thz
    else RHS2 = -2.0 * Re * velz + r * Ra * thz + dv(dvs(2.0 * Ra * (tem * velz + thz * v), K.re), K.ptl) + K.grptl * thr;
    const double uvh = dv(u, R.r) + dv(velh, R.r);
    const double cf1 = K.ptldt_inv - dv(uvh, K.ptl2) + dv(1.0, R.r2h) - cnr - cnz - cnh;
    const double cf2 = dv(uvh, K.ptl2) - dv(1.0, R.r2h) + cnr + cnz + cnh;
    return (cf2 * XV(FBP, 0, 0, 0) - ULT - RHS2 + RHS1) / cf1;
}

// Cvorticity at (i, j, k), rot3dmg.f:619-725; M = 1: rot3dmghor.f:643-754
template <int M, class V>
__device__ double cvort(const Geo &g, const Kc &K, const V &fv, int i, int j, int k)
{
    const int nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp, Ra = g.Ra, Re = g.Re;
    const Rad R(g.r0 + (double)i * dr);
    const double r = R.r.v;
    const double u = (XV(FU, 0, 0, 1) + XV(FU, 0, 0, 0)) * 0.5;
    const double v = (XV(FV, 0, 0, 0) + XV(FV, 1, 0, 0)) * 0.5;
    const double w = (XV(FW, 0, 0, 0) + XV(FW, 1, 0, 0) + XV(FW, 0, 0, 1) + XV(FW, 1, 0, 1) + XV(FW, 0, -1, 0) +
                      XV(FW, 1, -1, 0) + XV(FW, 0, -1, 1) + XV(FW, 1, -1, 1)) * 0.125;
    const double AA = (XV(FA, 0, 0, 0) + XV(FA, 0, -1, 0) + XV(FA, 1, 0, 0) + XV(FA, 1, -1, 0)) * 0.25;
    const double BB = (XV(FB, 0, 0, 0) + XV(FB, 0, -1, 0) + XV(FB, 0, 0, 1) + XV(FB, 0, -1, 1)) * 0.25;
    const double tem = (XV(FTH, 0, 0, 0) + XV(FTH, 1, 0, 0) + XV(FTH, 1, 0, 1) + XV(FTH, 0, 0, 1)) * 0.25;
    const double Cc = XV(FC, 0, 0, 0);
    const double Crra = dv(XV(FC, -1, 0, 0) + XV(FC, 1, 0, 0), K.dr2);
    const double Cra = dv((XV(FC, 1, 0, 0) - XV(FC, -1, 0, 0)) * 0.5, K.dr);
    double Cr = Cra;
    if (fabs(dv(u * dr, K.ptl)) > 2) {
        if (u > 0) Cr = dv(Cc - XV(FC, -1, 0, 0), K.dr);
        else if (u < 0) Cr = dv(XV(FC, 1, 0, 0) - Cc, K.dr);
    }
    const double welr = dv((XV(FW, 1, 0, 0) + XV(FW, 1, 0, 1) + XV(FW, 1, -1, 0) + XV(FW, 1, -1, 1)) * 0.25 -
                           (XV(FW, 0, 0, 0) + XV(FW, 0, 0, 1) + XV(FW, 0, -1, 0) + XV(FW, 0, -1, 1)) * 0.25, K.dr);
    const double thr = dv((XV(FTH, 1, 0, 0) + XV(FTH, 1, 0, 1) - XV(FTH, 0, 0, 0) - XV(FTH, 0, 0, 1)) * 0.5, K.dr);
    double Czza, Cz, cnz;
    if (j == 1) {
        const double Cm = XV(FC, 0, -1, 0), C1 = XV(FC, 0, 1, 0);
        Czza = dv(3.2 * Cm + 2.0 * C1 - 0.2 * XV(FC, 0, 2, 0), K.dz2);
        Cz = dv(dv(-4.0 * Cm, K.three) + Cc + dv(C1, K.three), K.dz);
        cnz = K.cnz_w;
    } else if (j == nz) {
        const double Cm = XV(FC, 0, -1, 0), C1 = XV(FC, 0, 1, 0);
        Czza = dv(3.2 * C1 + 2.0 * Cm - 0.2 * XV(FC, 0, -2, 0), K.dz2);
        Cz = dv(dv(-Cm, K.three) - Cc + dv(4.0 * C1, K.three), K.dz);
        cnz = K.cnz_w;
    } else {
        const double Cm = XV(FC, 0, -1, 0), C1 = XV(FC, 0, 1, 0);
        Czza = dv(Cm + C1, K.dz2);
        Cz = dv((C1 - Cm) * 0.5, K.dz);
        if (fabs(dv(w * dz, K.ptl)) > 2) {
            if (w > 0) Cz = dv(Cc - Cm, K.dz);
            else if (w < 0) Cz = dv(C1 - Cc, K.dz);
        }
        cnz = K.cnz_int;
    }
    const double welz = dv((XV(FW, 0, 0, 0) + XV(FW, 1, 0, 0) + XV(FW, 0, 0, 1) + XV(FW, 1, 0, 1)) * 0.25 -
                           (XV(FW, 0, -1, 0) + XV(FW, 1, -1, 0) + XV(FW, 0, -1, 1) + XV(FW, 1, -1, 1)) * 0.25, K.dz);
    const double Chha = dv(XV(FC, 0, 0, 1) + XV(FC, 0, 0, -1), K.dp2);
    double Ch = dv((XV(FC, 0, 0, 1) - XV(FC, 0, 0, -1)) * 0.5, K.dp);
    if (fabs(dv(v * dp * r, K.ptl)) > 2) {
        if (v > 0) Ch = dv(Cc - XV(FC, 0, 0, -1), K.dp);
        else if (v < 0) Ch = dv(XV(FC, 0, 0, 1) - Cc, K.dp);
    }
    const double welh = dv((XV(FW, 0, 0, 1) + XV(FW, 1, 0, 1) + XV(FW, 0, -1, 1) + XV(FW, 1, -1, 1)) * 0.25 -
                           (XV(FW, 0, 0, 0) + XV(FW, 1, 0, 0) + XV(FW, 0, -1, 0) + XV(FW, 1, -1, 0)) * 0.25, K.dp);
    const double thh = dv((XV(FTH, 0, 0, 1) + XV(FTH, 1, 0, 1) - XV(FTH, 0, 0, 0) - XV(FTH, 1, 0, 0)) * 0.5, K.dp);
    const double cnr = K.cnr_int;
    const double cnh = dv(K.m1dp2, R.r2);
    const double RHS1 = Crra + dv(Cra, R.r) + dv(Chha, R.r2) + Czza + dv(Cc, K.ptldt);
    const double ULT = dv(u * Cr + dv(v * Ch, R.r) + w * Cz - welr * AA - dv(welh * BB, R.r), K.ptl);
    double RHS2;
    if constexpr (M == 1)   // -1.d0/r*Gr*Ptl*costheta*thh-Gr*Ptl*sintheta*thr
        RHS2 = -(dv(1.0, R.r) * g.Gr * g.Ptl * K.trig[np + 1 + k] * thh) - K.grptl * K.trig[k] * thr;
    else
        RHS2 = -2.0 * Re * welz - Ra * thh + dv(dvs(2.0 * Ra * (tem * welz - thr * u - dv(thh * v, R.r)), K.re), K.ptl);
    const double cf1 = K.ptldt_inv - dv(welz, K.ptl2) - cnr - cnz - cnh;
    const double cf2 = dv(welz, K.ptl2) + cnr + cnz + cnh;
    return (cf2 * XV(FCP, 0, 0, 0) - ULT - RHS2 + RHS1) / cf1;
}

// temperature at (i, j, k), rot3dmg.f:754-871
template <class V>
__device__ double temper(const Geo &g, const Kc &K, const V &fv, int i, int j, int k)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz, dp = g.dp;
    const Rad R(g.r0 + ((double)i - 0.5) * dr);
    const double r = R.r.v;
    double u, v, w;
    if (j == 0 || j == nz + 1) {
        u = 0.0;
        v = 0.0;
        w = 0.0;
    } else {
        u = (XV(FU, 0, 0, 0) + XV(FU, -1, 0, 0)) * 0.5;
        v = (XV(FV, 0, 0, 0) + XV(FV, 0, 0, -1)) * 0.5;
        w = (XV(FW, 0, 0, 0) + XV(FW, 0, -1, 0)) * 0.5;
    }
    const double Tc = XV(FTH, 0, 0, 0);
    double thrra, thra, thr, cnr;
    if (i == 1) {
        const double Tm = XV(FTH, -1, 0, 0), T1 = XV(FTH, 1, 0, 0);
        thrra = dv(3.2 * Tm + 2.0 * T1 - 0.2 * XV(FTH, 2, 0, 0), K.dr2);
        thra = dv(dv(-4.0 * Tm, K.three) + dv(T1, K.three), K.dr);
        thr = dv(dv(-4.0 * Tm, K.three) + Tc + dv(T1, K.three), K.dr);
        cnr = K.cnr_w + dv(K.hdr, R.r);
    } else if (i == nr) {
        const double Tm = XV(FTH, -1, 0, 0), T1 = XV(FTH, 1, 0, 0);
        thrra = dv(3.2 * T1 + 2.0 * Tm - 0.2 * XV(FTH, -2, 0, 0), K.dr2);
        thra = dv(dv(-Tm, K.three) + dv(4.0 * T1, K.three), K.dr);
        thr = dv(dv(-Tm, K.three) - Tc + dv(4.0 * T1, K.three), K.dr);
        cnr = K.cnr_w - dv(K.hdr, R.r);
    } else {
        const double Tm = XV(FTH, -1, 0, 0), T1 = XV(FTH, 1, 0, 0);
        thrra = dv(Tm + T1, K.dr2);
        thra = dv((T1 - Tm) * 0.5, K.dr);
        thr = thra;
        if (fabs(dv(u * dr, K.ptl)) > 2) {
            if (u > 0) thr = dv(Tc - Tm, K.dr);
            else if (u < 0) thr = dv(T1 - Tc, K.dr);
        }
        cnr = K.cnr_int;
    }
    double thzza, thz, cnz;
    if (j == 1) {
        const double Tm = XV(FTH, 0, -1, 0), T1 = XV(FTH, 0, 1, 0);
        thzza = dv(3.2 * Tm + 2.0 * T1 - 0.2 * XV(FTH, 0, 2, 0), K.dz2);
        thz = dv(dv(-4.0 * Tm, K.three) + Tc + dv(T1, K.three), K.dz);
        cnz = K.cnz_w;
    } else if (j == nz) {
        const double Tm = XV(FTH, 0, -1, 0), T1 = XV(FTH, 0, 1, 0);
        thzza = dv(3.2 * T1 + 2.0 * Tm - 0.2 * XV(FTH, 0, -2, 0), K.dz2);
        thz = dv(dv(-Tm, K.three) - Tc + dv(4.0 * T1, K.three), K.dz);
        cnz = K.cnz_w;
    } else if (j == 0) {
        thzza = dv(8.0 * XV(FTH, 0, 1, 0), K.dz2);
        thz = 0.0;
        cnz = K.cnz_t;
    } else if (j == nz + 1) {
        thzza = dv(8.0 * XV(FTH, 0, -1, 0), K.dz2);
        thz = 0.0;
        cnz = K.cnz_t;
    } else {
        const double Tm = XV(FTH, 0, -1, 0), T1 = XV(FTH, 0, 1, 0);
        thzza = dv(Tm + T1, K.dz2);
        thz = dv((T1 - Tm) * 0.5, K.dz);
        if (fabs(dv(w * dz, K.ptl)) > 2) {
            if (w > 0) thz = dv(Tc - Tm, K.dz);
            else if (w < 0) thz = dv(T1 - Tc, K.dz);
        }
        cnz = K.cnz_int;
    }
    const double thhha = dv(XV(FTH, 0, 0, 1) + XV(FTH, 0, 0, -1), K.dp2);
    double thh = dv((XV(FTH, 0, 0, 1) - XV(FTH, 0, 0, -1)) * 0.5, K.dp);
    if (fabs(dv(v * dp * r, K.ptl)) > 2) {
        if (v > 0) thh = dv(Tc - XV(FTH, 0, 0, -1), K.dp);
        else if (v < 0) thh = dv(XV(FTH, 0, 0, 1) - Tc, K.dp);
    }
    const double cnp = dv(K.m1dp2, R.r2);
    const double RHS1 = thrra + dv(thra, R.r) + dv(thhha, R.r2) + thzza + dv(Tc, K.dt);
    const double ULT = u * thr + dv(v * thh, R.r) + w * thz;
    const double cf1 = K.dt_inv - cnr - cnz - cnp;
    const double cf2 = cnr + cnz + cnp;
    return (cf2 * XV(FTHP, 0, 0, 0) - ULT + RHS1) / cf1;
}

// The vorticity routines' reads while a column marches in k (MView): the
// values a column needs at several planes -- its node and the neighbours the
// staggered averages take across k -- stay in registers (K: plane k, M: its
// km, N: its ka), each loaded once per plane; the others are read at plane
// k.  A member's name gives the field and the offsets in i and j (m1 = -1).
// The set is that of the reads of avort, bvort and cvort.
struct MView {
    const Fields &F;
    unsigned o, li, bk = 0;   // (i, j) within a plane; row pitch; plane k
    double a00M, a00K, a00N, a10M, a10K, bm10K, bm10N, b0m1K;
    double b0m1N, b00M, b00K, b00N, c00M, c00K, c00N, c01M;
    double c01K, th00K, th00N, th01K, th01N, th10K, th10N, um10K;
    double um10N, um11K, um11N, u00K, u00N, u01K, u01N, v00M;
    double v00K, v01M, v01K, v10M, v10K, v11M, v11K, w0m1K;
    double w0m1N, w00K, w00N, w1m1K, w1m1N, w10K, w10N;
    __device__ MView(const Fields &F_, unsigned o_, unsigned li_) : F(F_), o(o_), li(li_) {}
    template <int FLD, int DI, int DJ, int DK>
    __device__ double x() const
    {
        if constexpr (FLD == FA && DI == 0 && DJ == 0)
            return DK == -1 ? a00M : DK == 0 ? a00K : a00N;
        else if constexpr (FLD == FA && DI == 1 && DJ == 0)
            return DK == -1 ? a10M : a10K;
        else if constexpr (FLD == FB && DI == -1 && DJ == 0)
            return DK == 0 ? bm10K : bm10N;
        else if constexpr (FLD == FB && DI == 0 && DJ == -1)
            return DK == 0 ? b0m1K : b0m1N;
        else if constexpr (FLD == FB && DI == 0 && DJ == 0)
            return DK == -1 ? b00M : DK == 0 ? b00K : b00N;
        else if constexpr (FLD == FC && DI == 0 && DJ == 0)
            return DK == -1 ? c00M : DK == 0 ? c00K : c00N;
        else if constexpr (FLD == FC && DI == 0 && DJ == 1)
            return DK == -1 ? c01M : c01K;
        else if constexpr (FLD == FTH && DI == 0 && DJ == 0)
            return DK == 0 ? th00K : th00N;
        else if constexpr (FLD == FTH && DI == 0 && DJ == 1)
            return DK == 0 ? th01K : th01N;
        else if constexpr (FLD == FTH && DI == 1 && DJ == 0)
            return DK == 0 ? th10K : th10N;
        else if constexpr (FLD == FU && DI == -1 && DJ == 0)
            return DK == 0 ? um10K : um10N;
        else if constexpr (FLD == FU && DI == -1 && DJ == 1)
            return DK == 0 ? um11K : um11N;
        else if constexpr (FLD == FU && DI == 0 && DJ == 0)
            return DK == 0 ? u00K : u00N;
        else if constexpr (FLD == FU && DI == 0 && DJ == 1)
            return DK == 0 ? u01K : u01N;
        else if constexpr (FLD == FV && DI == 0 && DJ == 0)
            return DK == -1 ? v00M : v00K;
        else if constexpr (FLD == FV && DI == 0 && DJ == 1)
            return DK == -1 ? v01M : v01K;
        else if constexpr (FLD == FV && DI == 1 && DJ == 0)
            return DK == -1 ? v10M : v10K;
        else if constexpr (FLD == FV && DI == 1 && DJ == 1)
            return DK == -1 ? v11M : v11K;
        else if constexpr (FLD == FW && DI == 0 && DJ == -1)
            return DK == 0 ? w0m1K : w0m1N;
        else if constexpr (FLD == FW && DI == 0 && DJ == 0)
            return DK == 0 ? w00K : w00N;
        else if constexpr (FLD == FW && DI == 1 && DJ == -1)
            return DK == 0 ? w1m1K : w1m1N;
        else if constexpr (FLD == FW && DI == 1 && DJ == 0)
            return DK == 0 ? w10K : w10N;
        else {
            static_assert(DK == 0, "a read across k without a register");
            return fptr<FLD>(F)[bk + o + (unsigned)(DJ * (int)li + DI)];
        }
    }
    // the march starts at plane k0 (offset b0) after km(k0) (offset bm)
    __device__ void start(unsigned bm, unsigned b0)
    {
        a00M = F.A[bm + o];
        a00K = F.A[b0 + o];
        a10M = F.A[bm + o + (unsigned)(0 * (int)li + 1)];
        bm10K = F.B[b0 + o + (unsigned)(0 * (int)li + -1)];
        b0m1K = F.B[b0 + o + (unsigned)(-1 * (int)li + 0)];
        b00M = F.B[bm + o];
        b00K = F.B[b0 + o];
        c00M = F.C[bm + o];
        c00K = F.C[b0 + o];
        c01M = F.C[bm + o + (unsigned)(1 * (int)li + 0)];
        th00K = F.th[b0 + o];
        th01K = F.th[b0 + o + (unsigned)(1 * (int)li + 0)];
        th10K = F.th[b0 + o + (unsigned)(0 * (int)li + 1)];
        um10K = F.uel[b0 + o + (unsigned)(0 * (int)li + -1)];
        um11K = F.uel[b0 + o + (unsigned)(1 * (int)li + -1)];
        u00K = F.uel[b0 + o];
        u01K = F.uel[b0 + o + (unsigned)(1 * (int)li + 0)];
        v00M = F.vel[bm + o];
        v01M = F.vel[bm + o + (unsigned)(1 * (int)li + 0)];
        v10M = F.vel[bm + o + (unsigned)(0 * (int)li + 1)];
        v11M = F.vel[bm + o + (unsigned)(1 * (int)li + 1)];
        w0m1K = F.wel[b0 + o + (unsigned)(-1 * (int)li + 0)];
        w00K = F.wel[b0 + o];
        w1m1K = F.wel[b0 + o + (unsigned)(-1 * (int)li + 1)];
        w10K = F.wel[b0 + o + (unsigned)(0 * (int)li + 1)];
    }
    // plane k (offset bk_) with its ka (offset bn)
    __device__ void advance(unsigned bk_, unsigned bn)
    {
        bk = bk_;
        a00N = F.A[bn + o];
        a10K = F.A[bk + o + (unsigned)(0 * (int)li + 1)];
        bm10N = F.B[bn + o + (unsigned)(0 * (int)li + -1)];
        b0m1N = F.B[bn + o + (unsigned)(-1 * (int)li + 0)];
        b00N = F.B[bn + o];
        c00N = F.C[bn + o];
        c01K = F.C[bk + o + (unsigned)(1 * (int)li + 0)];
        th00N = F.th[bn + o];
        th01N = F.th[bn + o + (unsigned)(1 * (int)li + 0)];
        th10N = F.th[bn + o + (unsigned)(0 * (int)li + 1)];
        um10N = F.uel[bn + o + (unsigned)(0 * (int)li + -1)];
        um11N = F.uel[bn + o + (unsigned)(1 * (int)li + -1)];
        u00N = F.uel[bn + o];
        u01N = F.uel[bn + o + (unsigned)(1 * (int)li + 0)];
        v00K = F.vel[bk + o];
        v01K = F.vel[bk + o + (unsigned)(1 * (int)li + 0)];
        v10K = F.vel[bk + o + (unsigned)(0 * (int)li + 1)];
        v11K = F.vel[bk + o + (unsigned)(1 * (int)li + 1)];
        w0m1N = F.wel[bn + o + (unsigned)(-1 * (int)li + 0)];
        w00N = F.wel[bn + o];
        w1m1N = F.wel[bn + o + (unsigned)(-1 * (int)li + 1)];
        w10N = F.wel[bn + o + (unsigned)(0 * (int)li + 1)];
    }
    __device__ void rotate()
    {
        a00M = a00K;
        a00K = a00N;
        a10M = a10K;
        bm10K = bm10N;
        b0m1K = b0m1N;
        b00M = b00K;
        b00K = b00N;
        c00M = c00K;
        c00K = c00N;
        c01M = c01K;
        th00K = th00N;
        th01K = th01N;
        th10K = th10N;
        um10K = um10N;
        um11K = um11N;
        u00K = u00N;
        u01K = u01N;
        v00M = v00K;
        v01M = v01K;
        v10M = v10K;
        v11M = v11K;
        w0m1K = w0m1N;
        w00K = w00N;
        w1m1K = w1m1N;
        w10K = w10N;
    }
};

// A, B and C in one pass, marching: one thread per (i, j) column over KC3
// consecutive planes, the values reused across planes carried in registers
// (MView); tiles dealt to the XCDs in contiguous ranges (the hardware deals
// linear workgroup ids to the 8 XCDs in turn), so the rows j-1 and j+1 a
// block reads were just fetched into that XCD's L2 by its neighbours.
constexpr int KC3 = 8;
constexpr int MARCH_MIN = 16384;   // columns (nr nz) from which the march pays

template <int M>
__global__ void __launch_bounds__(WG) k3_explicit_abc(Geo g, Kc K, Fields F)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const unsigned T = gridDim.x * gridDim.y, L = blockIdx.y * gridDim.x + blockIdx.x;
    const unsigned x = L & 7u, q = L >> 3, base = T >> 3, rem = T & 7u;
    const unsigned Wt = x * base + (x < rem ? x : rem) + q;
    const int kc = (int)(Wt / gridDim.x), bx = (int)(Wt - (unsigned)kc * gridDim.x);
    const int t = bx * WG + threadIdx.x;
    if (t >= nr * nz) return;
    const int jj = t / nr;
    const int i = 1 + (t - jj * nr), j = 1 + jj;
    const int k0 = kc * KC3, k1 = min(np, k0 + KC3 - 1);
    const unsigned pl = (unsigned)(g.ldi * g.ldj);
    MView fv(F, (unsigned)(j * g.ldi + i), (unsigned)g.ldi);
    fv.start((unsigned)(k0 == 0 ? np - 1 : k0 - 1) * pl, (unsigned)k0 * pl);
    for (int k = k0; k <= k1; ++k) {
        fv.advance((unsigned)k * pl, (unsigned)(k == np ? 1 : k + 1) * pl);
        const unsigned e = (unsigned)k * pl + fv.o;
        if (j <= nz - 1) F.An[e] = avort<M>(g, K, fv, i, j, k);
        if (i <= nr - 1 && j <= nz - 1) F.Bn[e] = bvort<M>(g, K, fv, i, j, k);
        if (i <= nr - 1) F.Cn[e] = cvort<M>(g, K, fv, i, j, k);
        fv.rotate();
    }
}

// One routine (W = 0: A, 1: B, 2: C, 3: T) per launch, or (W = 4) A, B and
// C together; blockIdx.y = k, a thread per (i, j) point of the routine's
// loop range (for W = 4 the union i = 1..nr, j = 1..nz, each routine
// evaluated where its own range holds).  A kernel per routine set keeps its
// constants and field pointers within the scalar registers.  M is the model
// (enum vv3_model).  W = 4 serves planes of fewer than MARCH_MIN columns,
// where marching leaves too few threads to fill the chip.
template <int W, int M>
__global__ void __launch_bounds__(WG) k3_explicit(Geo g, Kc K, Fields F)
{
    const int nr = g.nr, nz = g.nz;
    constexpr int which = W;
    int i0 = 1, ni, j0 = 1, nj;
    if (which == 4) { ni = nr; nj = nz; }
    else if (which == 0) { ni = nr; nj = nz - 1; }
    else if (which == 1) { ni = nr - 1; nj = nz - 1; }
    else if (which == 2) { ni = nr - 1; nj = nz; }
    else { ni = nr; j0 = 0; nj = nz + 2; }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * nj) return;
    const int jj = t / ni;
    const int i = i0 + (t - jj * ni), j = j0 + jj, k = blockIdx.y;
    const int km = (k == 0) ? g.np - 1 : k - 1, ka = (k == g.np) ? 1 : k + 1;
    const unsigned o = (unsigned)(j * g.ldi + i), pl = (unsigned)(g.ldi * g.ldj);
    const GView fv{F, {(unsigned)km * pl + o, (unsigned)k * pl + o, (unsigned)ka * pl + o}, g.ldi};
    if constexpr (W == 4) {
        if (j <= nz - 1) X(F.An, i, j, k) = avort<M>(g, K, fv, i, j, k);
        if (i <= nr - 1 && j <= nz - 1) X(F.Bn, i, j, k) = bvort<M>(g, K, fv, i, j, k);
        if (i <= nr - 1) X(F.Cn, i, j, k) = cvort<M>(g, K, fv, i, j, k);
        return;
    }
    if constexpr (W == 0) X(F.An, i, j, k) = avort<M>(g, K, fv, i, j, k);
    else if constexpr (W == 1) X(F.Bn, i, j, k) = bvort<M>(g, K, fv, i, j, k);
    else if constexpr (W == 2) X(F.Cn, i, j, k) = cvort<M>(g, K, fv, i, j, k);
    else X(F.thn, i, j, k) = temper(g, K, fv, i, j, k);
}

// heat, rot3dmg.f:882-905.  The flux terms of every (j, k) in parallel,
// stored in the reference's summation order over j (0, 1, nz, nz+1, 2, ...,
// nz-1): hs[p * np + k-1].
__global__ void __launch_bounds__(WG) k3_heat_terms(Geo g, const double *th, double *hs)
{
    const int nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (nz + 2) * np) return;
    const int k = 1 + t % np, j = t / np;
    const double w = (j == 0 || j == nz + 1) ? 240.0 : (j == 1 || j == nz) ? 80.0 : 60.0;
    const int p = (j == 0) ? 0 : (j == 1) ? 1 : (j == nz) ? 2 : (j == nz + 1) ? 3 : j + 2;
    hs[(size_t)p * np + (k - 1)] =
        dz * (-184.0 * X(th, 0, j, k) + 225.0 * X(th, 1, j, k) - 50.0 * X(th, 2, j, k) + 9.0 * X(th, 3, j, k)) /
        (w * dr);
}

// the sums in the reference's order: trans per k over j, then trans1 over k
// (the per-k sums kept in LDS, np doubles of dynamic shared memory)
__global__ void __launch_bounds__(WG) k3_heat_sum(Geo g, const double *hs, double *out)
{
    extern __shared__ double part[];
    const int nz = g.nz, np = g.np;
    for (int k = threadIdx.x; k < np; k += blockDim.x) {
        double trans = hs[k];
        int p = 1;
        for (; p + 8 <= nz + 2; p += 8) {   // loads issued ahead of the in-order sum
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = hs[(size_t)(p + u) * np + k];
#pragma unroll
            for (int u = 0; u < 8; ++u) trans = trans + v[u];
        }
        for (; p < nz + 2; ++p) trans = trans + hs[(size_t)p * np + k];
        part[k] = trans;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double trans1 = 0.0;
        int k = 0;
        for (; k + 8 <= np; k += 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = part[k + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) trans1 = trans1 + v[u] * g.dp;
        }
        for (; k < np; ++k) trans1 = trans1 + part[k] * g.dp;
        *out = trans1;
    }
}

// heat of rot3dmghor.f:913-946.  Per plane k (a thread each) the
// inner-cylinder sum transa and the outer-cylinder sum transb over
// j = nz/2-3 .. nz/2+3 in the reference's order (the two end rows with
// weight 1/120 first, then the rows between with 1/60): hs[k-1] and
// hs[np + k-1].
__global__ void __launch_bounds__(WG) k3_heat_hor_terms(Geo g, const double *th, double *hs)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz;
    const int k = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k > np) return;
    const int nza = nz / 2 - 3, nzb = nz / 2 + 3;
    auto ta = [&](int j, double w) {
        return dz * (-184.0 * X(th, 0, j, k) + 225.0 * X(th, 1, j, k) - 50.0 * X(th, 2, j, k) + 9.0 * X(th, 3, j, k)) /
               (w * dr);
    };
    auto tb = [&](int j, double w) {
        return dz * (184.0 * X(th, nr + 1, j, k) - 225.0 * X(th, nr, j, k) + 50.0 * X(th, nr - 1, j, k) -
                     9.0 * X(th, nr - 2, j, k)) / (w * dr);
    };
    double a = ta(nza, 120.0), b = tb(nza, 120.0);
    a = a + ta(nzb, 120.0);
    b = b + tb(nzb, 120.0);
    for (int j = nza + 1; j <= nzb - 1; ++j) {
        a = a + ta(j, 60.0);
        b = b + tb(j, 60.0);
    }
    hs[k - 1] = a;
    hs[np + k - 1] = b;
}

// trans1 = sum over k of transa*dp and trans2 of transb*dp, in k order
// (one lane of wave 0 and one of wave 1; 128 threads)
__global__ void __launch_bounds__(128) k3_heat_hor_sum(Geo g, const double *hs, double *out)
{
    const int np = g.np, w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) != 0 || w > 1) return;
    const double *h = hs + (size_t)w * np;
    double t = 0.0;
    int k = 0;
    for (; k + 8 <= np; k += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = h[k + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) t = t + v[u] * g.dp;
    }
    for (; k < np; ++k) t = t + h[k] * g.dp;
    out[w] = t;
}

// Kinetic energy (no counterpart in the reference; vvmi3.h gives the
// definition).  Workgroup b sums the planes k = b, b + gridDim.x, ... over
// their (i, j) points; partials are reduced by wavefront shuffles and
// k3_energy_sum adds them in index order, so the result repeats bit for bit.
__global__ void __launch_bounds__(WG) k3_energy(Geo g, const double *uel, const double *vel, const double *wel,
                                                double *part)
{
    __shared__ double sh[WG / 64];
    const int nr = g.nr, nz = g.nz, np = g.np;
    const int plane = g.ldi * g.ldj;
    double s = 0.0;
    for (int k = blockIdx.x; k < np; k += gridDim.x) {
        for (int t = threadIdx.x; t < plane; t += WG) {
            const int j = t / g.ldi, i = t - j * g.ldi;
            if (i < 1 || i > nr || j < 1 || j > nz) continue;
            const double rc = g.r0 + ((double)i - 0.5) * g.dr;
            const double v = X(vel, i, j, k);
            s += rc * v * v;
            if (j <= nz - 1) {
                const double w = X(wel, i, j, k);
                s += rc * w * w;
            }
            if (i <= nr - 1) {
                const double u = X(uel, i, j, k);
                s += (g.r0 + (double)i * g.dr) * u * u;
            }
        }
    }
    s = vv::block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(WG) k3_energy_sum(const double *part, int nb, double scale, double *out)
{
    __shared__ double sh[WG / 64];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += WG) s += part[b];
    s = vv::block_sum(s, sh);
    if (threadIdx.x == 0) *out = scale * s;
}

// Local Nusselt numbers at mid-height, sect360v2mg.f:318-325: a thread per
// plane k = 0..np, in the reference's operation order (L = ln(r0) from the
// host's C library, as the reference's dlog).
__global__ void __launch_bounds__(WG) k3_local_nu(Geo g, const double *th, double rmax, double L, double *inner,
                                                  double *outer)
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

// Fourier-mode integrals after rot3dmghor.f:954-991.  The plane sums
// P_k = sum_{i,j} f_i f_j (r0 + i dr) T(i,j,k) do not depend on the mode,
// so workgroup k-1 forms P_k (k = 1..np) and k3_modes_dft takes the cos and
// sin sums over k of each mode, one reduction per (mode, kind).
__global__ void __launch_bounds__(WG) k3_plane_sums(Geo g, const double *th, double *psum)
{
    __shared__ double sh[WG / 64];
    const int k = blockIdx.x + 1;
    const int nr = g.nr, nz = g.nz, plane = g.ldi * g.ldj;
    double sm = 0.0;
    for (int t = threadIdx.x; t < plane; t += WG) {
        const int j = t / g.ldi, i = t - j * g.ldi;
        if (i > nr + 1 || j > nz + 1) continue;
        double f = 1.0;
        if (i == 0 || i == nr + 1) f *= 0.5;
        if (j == 0 || j == nz + 1) f *= 0.5;
        sm += (g.r0 + (double)i * g.dr) * X(th, i, j, k) * f;
    }
    sm = vv::block_sum(sm, sh);
    if (threadIdx.x == 0) psum[blockIdx.x] = sm;
}

__global__ void __launch_bounds__(WG) k3_modes_dft(const double *psum, const double *tw, int np, int nmax,
                                                   double scale, double *out)
{
    __shared__ double sh[WG / 64];
    for (int q = 0; q < 2 * nmax; ++q) {   // q < nmax: cos of mode q+1; else sin of mode q-nmax+1
        const double *w = tw + (size_t)((q < nmax) ? q : MODES_MAX + q - nmax) * np;
        double sm = 0.0;
        for (int k = threadIdx.x; k < np; k += WG) sm += psum[k] * w[k];
        sm = vv::block_sum(sm, sh);
        if (threadIdx.x == 0) out[q] = scale * sm;
        __syncthreads();
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
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * nj) return;
    const int jj = t / ni;
    const int i = 1 + (t - jj * ni), j = 1 + jj, k = blockIdx.y;
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
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ldi * g.ldj) return;
    const int j = t / g.ldi, i = t - j * g.ldi, k = blockIdx.y;
    const unsigned p = (unsigned)(k * g.ldi * g.ldj + t);
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
    // th: 0 on the inner, 1 on the outer cylinder (rot3dmg.f:987-990);
    // 1 inside, 0 outside in rot3dmghor.f:1042-1043
    double th = F.thn[p];
    if (i == 0 || i == nr + 1) {
        th = ((i == 0) == (g.model == VV3_MODEL_HOR)) ? 1.0 : 0.0;
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
    g.model = r->cfg.model;
    return g;
}

extern "C" {

int vv3_create(const vv3_config *cfg, int device, vv3_run **out)
{
    using vv::set_error;
    if (!cfg || !out) { set_error("vv3_create: null argument"); return VV_EINVAL; }
    *out = nullptr;
    if (cfg->model != VV3_MODEL_ROT3DMG && cfg->model != VV3_MODEL_HOR) {
        set_error("vv3_create: unknown model");
        return VV_EINVAL;
    }
    const bool hor = cfg->model == VV3_MODEL_HOR;
    if (cfg->nr < 4 || cfg->nz < (hor ? 6 : 4) || cfg->np < 4 || (cfg->np & 1) || cfg->np > 8192) {
        set_error("vv3_create: need nr >= 4, nz >= 4 (6 for the horizontal annulus) and an even np, "
                  "4 <= np <= 8192");
        return VV_EINVAL;
    }
    // rot3dmg.f divides by Re; rot3dmghor.f has no rotation terms
    if (!(cfg->dt > 0) || !(cfg->rmax > cfg->r0) || !(cfg->r0 > 0) || !(cfg->zmax > 0) || !(cfg->Ptl > 0) ||
        (!hor && cfg->Re == 0.0)) {
        set_error("vv3_create: bad physical parameters");
        return VV_EINVAL;
    }
    VV_HIP(hipSetDevice(device));
    vv3_run *r = new vv3_run();
    r->cfg = *cfg;
    r->device = device;
    r->nr = cfg->nr; r->nz = cfg->nz; r->np = cfg->np;
    r->ldi = cfg->nr + 2; r->ldj = cfg->nz + 2;
    // rot3dmg.f:91-93, unless the caller passes its own values
    r->dr = (cfg->dr > 0.0) ? cfg->dr : (cfg->rmax - cfg->r0) / cfg->nr;
    r->dz = (cfg->dz > 0.0) ? cfg->dz : cfg->zmax / cfg->nz;
    r->dp = (cfg->dp > 0.0) ? cfg->dp : 2.0 * std::acos(-1.0) / cfg->np;
    r->nf = (size_t)r->ldi * r->ldj * (cfg->np + 1);
    auto fail = [&](int s) { vv3_destroy(r); return s; };
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) return fail(VV_EHIP);
    for (int q = 0; q < VV3_NFIELD; ++q) {
        if (hipMalloc(&r->f[q], r->nf * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
        if (hipMemsetAsync(r->f[q], 0, r->nf * sizeof(double), r->stream) != hipSuccess) return fail(VV_EHIP);
    }
    {
        const size_t nh = (size_t)(cfg->nz + 3) * cfg->np + 2;
        if (hipMalloc(&r->heat, nh * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
        if (hipMemsetAsync(r->heat, 0, nh * sizeof(double), r->stream) != hipSuccess) return fail(VV_EHIP);
    }
    if (hipMalloc(&r->kepart, (KE3_MAXB + 1) * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
    if (hor) {
        // the azimuth of plane k as rot3dmghor.f:374 forms it,
        // theta=6.283185307d0/np*k, and its sine and cosine from the host's
        // libm (the library the reference's dsin/dcos call)
        const int n1 = cfg->np + 1;
        std::vector<double> t(2 * (size_t)n1);
        for (int k = 0; k < n1; ++k) {
            const double theta = 6.283185307 / cfg->np * k;
            t[k] = std::sin(theta);
            t[n1 + k] = std::cos(theta);
        }
        if (hipMalloc(&r->trig, t.size() * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
        if (hipMemcpyAsync(r->trig, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, r->stream) !=
            hipSuccess)
            return fail(VV_EHIP);
        if (hipStreamSynchronize(r->stream) != hipSuccess) return fail(VV_EHIP);
    }
    {
        const int st = poisson3_setup(r->pois, geo(r));
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

int vv3_destroy(vv3_run *r)
{
    if (!r) return VV_OK;
    {   // the lock is released before the run is deleted
        std::lock_guard<std::recursive_mutex> lk(r->mu);
        (void)hipSetDevice(r->device);
        if (r->stream) (void)hipStreamSynchronize(r->stream);
        if (r->stream2) (void)hipStreamSynchronize(r->stream2);
        for (int q = 0; q < VV3_NFIELD; ++q)
            if (r->f[q]) (void)hipFree(r->f[q]);
        poisson3_free(r->pois);
        for (hipEvent_t e : {r->fork, r->join})
            if (e) (void)hipEventDestroy(e);
        if (r->stream2) (void)hipStreamDestroy(r->stream2);
        if (r->heat) (void)hipFree(r->heat);
        if (r->trig) (void)hipFree(r->trig);
        if (r->kepart) (void)hipFree(r->kepart);
        if (r->dgbuf) (void)hipFree(r->dgbuf);
        if (r->stream) (void)hipStreamDestroy(r->stream);
    }
    delete r;
    return VV_OK;
}

// host (ldi x ldj x (np+1)) <-> device ((nr+2) x (nz+2) x (np+1))
static int copy3(vv3_run *r, int field, double *host, int ldi, int ldj, bool to_dev)
{
    const int di = r->ldi, dj = r->ldj, nk = r->np + 1;
    if (ldi == 0) ldi = di;
    if (ldj == 0) ldj = dj;
    if (ldi < di || ldj < dj) {
        vv::set_error("vv3_put/vv3_get: ldi must be >= nr+2 and ldj >= nz+2");
        return VV_EINVAL;
    }
    VV_HIP(hipSetDevice(r->device));
    double *dev = r->f[field];
    // planes k = 0..np through the pinned stage in the device layout (rows
    // of nr+2 doubles, di x dj per plane): one DMA copy
    const size_t words = (size_t)di * dj * nk;
    double *st = r->stage.get(words);
    if (!st) { vv::set_error("vv3_put/vv3_get: pinned staging buffer"); return VV_ENOMEM; }
    auto rows = [&](bool pack) {
        if (ldi == di && ldj == dj) {
            if (pack) std::memcpy(st, host, words * sizeof(double));
            else std::memcpy(host, st, words * sizeof(double));
            return;
        }
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

int vv3_put(vv3_run *r, int field, const double *host, int ldi, int ldj)
{
    if (!r || !host || field < 0 || field >= VV3_NFIELD) { vv::set_error("vv3_put: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    return copy3(r, field, const_cast<double *>(host), ldi, ldj, true);
}

int vv3_get(vv3_run *r, int field, double *host, int ldi, int ldj)
{
    if (!r || !host || field < 0 || field >= VV3_NFIELD) { vv::set_error("vv3_get: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    return copy3(r, field, host, ldi, ldj, false);
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

// heat's terms and sums; the total lands at heat_total(r)
static double *heat_total(vv3_run *r) { return r->heat + (size_t)(r->nz + 3) * r->np; }

static int launch_heat(vv3_run *r, hipStream_t s)
{
    const Geo g = geo(r);
    double *hs = r->heat;
    if (g.model == VV3_MODEL_HOR) {
        hipLaunchKernelGGL(k3_heat_hor_terms, dim3(nblk(r->np, WG)), dim3(WG), 0, s, g,
                           (const double *)r->f[VV3_TH], hs);
        hipLaunchKernelGGL(k3_heat_hor_sum, dim3(1), dim3(128), 0, s, g, (const double *)hs, heat_total(r));
        VV_HIP(hipGetLastError());
        return VV_OK;
    }
    hipLaunchKernelGGL(k3_heat_terms, dim3(nblk((long long)(r->nz + 2) * r->np, WG)), dim3(WG), 0, s, g,
                       (const double *)r->f[VV3_TH], hs);
    hipLaunchKernelGGL(k3_heat_sum, dim3(1), dim3(WG), r->np * sizeof(double), s, g, (const double *)hs,
                       heat_total(r));
    VV_HIP(hipGetLastError());
    return VV_OK;
}

// the explicit routines' constants, evaluated as the reference writes them
static Kc constants(const vv3_run *r)
{
    const double dr = r->dr, dz = r->dz, dp = r->dp, dt = r->cfg.dt, Ptl = r->cfg.Ptl;
    auto rc = [](double v) { return Rc{v, 1.0 / v}; };
    Kc K;
    K.dr = rc(dr); K.dz = rc(dz); K.dp = rc(dp);
    K.dr2 = rc(dr * dr); K.dz2 = rc(dz * dz); K.dp2 = rc(dp * dp);
    K.ptl = rc(Ptl); K.ptl2 = rc(2.0 * Ptl); K.ptldt = rc(Ptl * dt);
    K.re = rc(r->cfg.Re); K.three = rc(3.0); K.dt = rc(dt);
    K.cnr_int = -1.0 / (dr * dr);
    K.cnz_int = -1.0 / (dz * dz);
    K.cnr_w = -2.5 / (dr * dr);
    K.cnz_w = -2.5 / (dz * dz);
    K.cnz_t = -4.0 / (dz * dz);
    K.hdr = 0.5 / dr;
    K.m1dp2 = -1.0 / (dp * dp);
    K.ptldt_inv = 1.0 / (Ptl * dt);
    K.dt_inv = 1.0 / dt;
    K.grptl = r->cfg.Gr * Ptl;
    K.trig = r->trig;
    return K;
}

extern "C++" {
template <int M>
static int launch_explicit_m(vv3_run *r, int mask)
{
    const Geo g = geo(r);
    const Kc K = constants(r);
    const Fields F = fields(r);
    const int np1 = r->np + 1;
    const long long pa = (long long)r->nr * (r->nz - 1), pb = (long long)(r->nr - 1) * (r->nz - 1);
    const long long pc = (long long)(r->nr - 1) * r->nz, pt = (long long)r->nr * (r->nz + 2);
    if ((mask & 7) == 7) {
        const long long pu = (long long)r->nr * r->nz;
        if (pu >= MARCH_MIN)
            hipLaunchKernelGGL((k3_explicit_abc<M>), dim3(nblk(pu, WG), (np1 + KC3 - 1) / KC3), dim3(WG), 0,
                               r->stream, g, K, F);
        else
            hipLaunchKernelGGL((k3_explicit<4, M>), dim3(nblk(pu, WG), np1), dim3(WG), 0, r->stream, g, K, F);
    } else {
        if (mask & 1)
            hipLaunchKernelGGL((k3_explicit<0, M>), dim3(nblk(pa, WG), np1), dim3(WG), 0, r->stream, g, K, F);
        if (mask & 2)
            hipLaunchKernelGGL((k3_explicit<1, M>), dim3(nblk(pb, WG), np1), dim3(WG), 0, r->stream, g, K, F);
        if (mask & 4)
            hipLaunchKernelGGL((k3_explicit<2, M>), dim3(nblk(pc, WG), np1), dim3(WG), 0, r->stream, g, K, F);
    }
    // temperature is the same in both programs
    if (mask & 8) hipLaunchKernelGGL((k3_explicit<3, 0>), dim3(nblk(pt, WG), np1), dim3(WG), 0, r->stream, g, K, F);
    VV_HIP(hipGetLastError());
    return VV_OK;
}
}  // extern "C++"

static int launch_explicit(vv3_run *r, int mask)
{
    return (r->cfg.model == VV3_MODEL_HOR) ? launch_explicit_m<1>(r, mask) : launch_explicit_m<0>(r, mask);
}

static int launch_rhs(vv3_run *r, int which, hipStream_t s)
{
    RhsF F;
    F.An = r->f[VV3_AN]; F.Bn = r->f[VV3_BN]; F.Cn = r->f[VV3_CN]; F.wel = r->f[VV3_WEL];
    F.wrh = r->f[VV3_WRH]; F.vrh = r->f[VV3_VRH]; F.urh = r->f[VV3_URH];
    const long long cnt = (long long)r->nr * r->nz;
    hipLaunchKernelGGL(k3_rhs, dim3(nblk(cnt, WG), r->np + 1), dim3(WG), 0, s, geo(r), F, which);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

static int launch_poisson(vv3_run *r, int which, hipStream_t s, bool second = false)
{
    const int rhs = (which == 1) ? VV3_WRH : (which == 2) ? VV3_VRH : VV3_URH;
    const int psi = (which == 1) ? VV3_WEL : (which == 2) ? VV3_VEL : VV3_UEL;
    const int psip = (which == 1) ? VV3_WELP : (which == 2) ? VV3_VELP : VV3_UELP;
    return poisson3_launch(r->pois, geo(r), which, r->f[psi], r->f[psip], r->f[rhs], s, second);
}

static int launch_update(vv3_run *r)
{
    UpdF F;
    F.uel = r->f[VV3_UEL]; F.vel = r->f[VV3_VEL]; F.wel = r->f[VV3_WEL];
    F.An = r->f[VV3_AN]; F.Bn = r->f[VV3_BN]; F.Cn = r->f[VV3_CN]; F.thn = r->f[VV3_THN];
    F.Ap = r->f[VV3_AP]; F.Bp = r->f[VV3_BP]; F.Cp = r->f[VV3_CP]; F.thp = r->f[VV3_THP];
    hipLaunchKernelGGL(k3_update, dim3(nblk((long long)r->ldi * r->ldj, WG), r->np + 1), dim3(WG), 0, r->stream,
                       geo(r), F);
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
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    int st = launch_explicit(r, which == 0 ? 7 : 1 << (which - 1));
    if (!st) VV_HIP(hipStreamSynchronize(r->stream));
    return st;
}

int vv3_temperature(vv3_run *r)
{
    if (!r) { vv::set_error("vv3: null run"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    int st = launch_explicit(r, 8);
    if (!st) VV_HIP(hipStreamSynchronize(r->stream));
    return st;
}

int vv3_heat(vv3_run *r, double *trans1)
{
    if (!r) { vv::set_error("vv3: null run"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    const int st = launch_heat(r, r->stream);
    if (st) return st;
    double t = 0.0;
    VV_HIP(hipMemcpyAsync(&t, heat_total(r), sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    if (trans1) *trans1 = t;
    return VV_OK;
}

int vv3_energy(vv3_run *r, double *ke)
{
    if (!r || !ke) { vv::set_error("vv3_energy: null argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    const int nb = r->np < KE3_MAXB ? r->np : KE3_MAXB;
    const Geo g = geo(r);
    hipLaunchKernelGGL(k3_energy, dim3(nb), dim3(WG), 0, r->stream, g, (const double *)r->f[VV3_UEL],
                       (const double *)r->f[VV3_VEL], (const double *)r->f[VV3_WEL], r->kepart);
    hipLaunchKernelGGL(k3_energy_sum, dim3(1), dim3(WG), 0, r->stream, (const double *)r->kepart, nb,
                       0.5 * r->dr * r->dz * r->dp, r->kepart + KE3_MAXB);
    VV_HIP(hipGetLastError());
    VV_HIP(hipMemcpyAsync(ke, r->kepart + KE3_MAXB, sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

// the diagnostics buffer, allocated and tabulated at first use
static int diag_buffer(vv3_run *r)
{
    if (r->dgbuf) return VV_OK;
    const int np = r->np;
    const size_t nw = (size_t)np + 2 * MODES_MAX * (size_t)np + 2 * MODES_MAX + 2 * ((size_t)np + 1);
    std::vector<double> t(2 * MODES_MAX * (size_t)np);
    for (int n = 1; n <= MODES_MAX; ++n)
        for (int k = 1; k <= np; ++k) {
            const double ang = (double)k * r->dp;   // ang=dble(k)*dp, rot3dmghor.f:970
            t[(size_t)(n - 1) * np + (k - 1)] = std::cos((double)n * ang);
            t[(size_t)(MODES_MAX + n - 1) * np + (k - 1)] = std::sin((double)n * ang);
        }
    if (hipMalloc(&r->dgbuf, nw * sizeof(double)) != hipSuccess) {
        r->dgbuf = nullptr;
        vv::set_error("vv3 diagnostics: device allocation failed");
        return VV_ENOMEM;
    }
    VV_HIP(hipMemcpyAsync(r->dgbuf + np, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, r->stream));
    return VV_OK;
}

int vv3_local_nu(vv3_run *r, double *inner, double *outer)
{
    if (!r || !inner || !outer) { vv::set_error("vv3_local_nu: null argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    int st = diag_buffer(r);
    if (st) return st;
    const int np = r->np;
    double *nu = r->dgbuf + (size_t)np + 2 * MODES_MAX * (size_t)np + 2 * MODES_MAX;
    hipLaunchKernelGGL(k3_local_nu, dim3(nblk(np + 1, WG)), dim3(WG), 0, r->stream, geo(r),
                       (const double *)r->f[VV3_TH], r->cfg.rmax, std::log(r->cfg.r0), nu, nu + np + 1);
    VV_HIP(hipGetLastError());
    VV_HIP(hipMemcpyAsync(inner, nu, (np + 1) * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipMemcpyAsync(outer, nu + np + 1, (np + 1) * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vv3_modes(vv3_run *r, int nmax, double *c, double *s)
{
    if (!r || !c || !s || nmax < 1 || nmax > MODES_MAX) {
        vv::set_error("vv3_modes: need 1 <= nmax <= 16 and output arrays");
        return VV_EINVAL;
    }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    int st = diag_buffer(r);
    if (st) return st;
    const int np = r->np;
    double *psum = r->dgbuf, *tw = r->dgbuf + np, *out = r->dgbuf + np + 2 * MODES_MAX * (size_t)np;
    hipLaunchKernelGGL(k3_plane_sums, dim3(np), dim3(WG), 0, r->stream, geo(r), (const double *)r->f[VV3_TH], psum);
    hipLaunchKernelGGL(k3_modes_dft, dim3(1), dim3(WG), 0, r->stream, (const double *)psum, (const double *)tw, np,
                       nmax, r->dr * r->dz * r->dp, out);
    VV_HIP(hipGetLastError());
    VV_HIP(hipMemcpyAsync(c, out, nmax * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipMemcpyAsync(s, out + nmax, nmax * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vv3_fluxes(vv3_run *r, double *trans1, double *trans2)
{
    if (!r) { vv::set_error("vv3: null run"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    double t[2] = {0.0, 0.0};
    VV_HIP(hipMemcpyAsync(t, heat_total(r), sizeof(t), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    if (trans1) *trans1 = t[0];
    if (trans2) *trans2 = t[1];
    return VV_OK;
}

int vv3_rhs(vv3_run *r, int which)
{
    if (!r || which < 1 || which > 3) { vv::set_error("vv3_rhs: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    int st = launch_rhs(r, which, r->stream);
    if (!st) VV_HIP(hipStreamSynchronize(r->stream));
    return st;
}

int vv3_poisson(vv3_run *r, int which)
{
    if (!r || which < 1 || which > 3) { vv::set_error("vv3_poisson: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    int st = launch_poisson(r, which, r->stream);
    if (!st) VV_HIP(hipStreamSynchronize(r->stream));
    return st;
}

int vv3_update(vv3_run *r)
{
    if (!r) { vv::set_error("vv3: null run"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    int st = launch_update(r);
    if (!st) VV_HIP(hipStreamSynchronize(r->stream));
    return st;
}

int vv3_step(vv3_run *r, int n, double *trans1)
{
    if (!r || n < 0) { vv::set_error("vv3_step: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    for (int s = 0; s < n; ++s) {
        int st = launch_explicit(r, 15);
        if (!st && r->stream2) {
            // the heat flux (it reads th) and v's right-hand side (An and Cn
            // only, rot3dmg.f:1066-1093) and solve on the second stream with
            // the second work arrays, w's and u's (u's right-hand side reads
            // the new w) on the first; the update waits for both
            VV_HIP(hipEventRecord(r->fork, r->stream));
            VV_HIP(hipStreamWaitEvent(r->stream2, r->fork, 0));
            st = launch_heat(r, r->stream2);
            if (!st) st = launch_rhs(r, 2, r->stream2);
            if (!st) st = launch_poisson(r, 2, r->stream2, true);
            if (!st) VV_HIP(hipEventRecord(r->join, r->stream2));
            if (!st) st = launch_rhs(r, 1, r->stream);
            if (!st) st = launch_poisson(r, 1, r->stream);
            if (!st) st = launch_rhs(r, 3, r->stream);
            if (!st) st = launch_poisson(r, 3, r->stream);
            if (!st) VV_HIP(hipStreamWaitEvent(r->stream, r->join, 0));
        } else {
            if (!st) st = launch_heat(r, r->stream);
            for (int w = 1; w <= 3 && !st; ++w) {
                st = launch_rhs(r, w, r->stream);
                if (!st) st = launch_poisson(r, w, r->stream);
            }
        }
        if (!st) st = launch_update(r);
        if (st) return st;
    }
    double t = 0.0;
    VV_HIP(hipMemcpyAsync(&t, heat_total(r), sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    if (trans1) *trans1 = t;
    return VV_OK;
}

}  // extern "C"
