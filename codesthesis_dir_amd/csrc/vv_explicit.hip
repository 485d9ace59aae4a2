// vv_explicit.hip -- explicit part of the velocity-vorticity time step on
// gfx950: compact-scheme derivatives (batched constant-coefficient Thomas
// and cyclic solves), DuFort-Frankel/leapfrog stencils with the Robert
// filter, Poisson right-hand sides, wall update, wall heat flux.
//
// Arithmetic follows the reference expressions operation by operation
// (rb.f = velvor2dperiodicrbconv.f) and the library is compiled with
// -ffp-contract=off, so these kernels reproduce the reference bit for bit.
// Field layout: F[k*ld + i], i = 0..nr+1 across the walls, k = 0..np.

#include "vv_internal.h"

#include <cmath>

namespace vv {

namespace {

constexpr int WG = 256;

__device__ __forceinline__ size_t ix(int i, int k, int ld) { return (size_t)k * ld + i; }

// single-precision literals of rb.f:660-668 promoted to double
__device__ __forceinline__ double f23() { return (double)(2.f / 3.f); }
__device__ __forceinline__ double f16() { return (double)(1.f / 6.f); }

struct CompactArgs {
    const double *C, *th;
    double *dCr, *dCh, *dthrC, *dthhC, *dthrT, *dthhT;
    const double *tab_r1, *tab_rT, *tab_p;
    int nr, np, ld, mask;
    double dri, dpi;
};

// ---- batched Thomas along r (one thread per column k), rb.f:3046-3064
// with factors from TdmaTab: fact | bp | c, each of length n.
template <int FAM>
__device__ void tdma_column(const CompactArgs &a, int k)
{
    const int nr = a.nr, ld = a.ld, np = a.np;
    const int n = (FAM == 2) ? nr : nr - 1;
    const double *tab = (FAM == 2) ? a.tab_rT : a.tab_r1;
    const double *fact = tab, *bp = tab + n, *cc = tab + 2 * n;
    const double dri = a.dri;
    double *out = (FAM == 0) ? a.dCr : (FAM == 1) ? a.dthrC : a.dthrT;
    const double *C = a.C, *th = a.th;
    const int ka = (k == np) ? 1 : k + 1;
    double rprev = 0.0;
    for (int j = 0; j < n; ++j) {
        const int i = j + 1;
        double r;
        if (FAM == 0) {            // Cr, rb.f:810-824
            if (i == 1)
                r = 3.0 * (C[ix(2, k, ld)] - C[ix(0, k, ld)]) * dri
                  - (-1.5 * C[ix(0, k, ld)] + 2.0 * C[ix(1, k, ld)] - 0.5 * C[ix(2, k, ld)]) * dri;
            else if (i == nr - 1)
                r = 3.0 * (C[ix(nr, k, ld)] - C[ix(nr - 2, k, ld)]) * dri
                  - (0.5 * C[ix(nr - 2, k, ld)] - 2.0 * C[ix(nr - 1, k, ld)] + 1.5 * C[ix(nr, k, ld)]) * dri;
            else
                r = 3.0 * (C[ix(i + 1, k, ld)] - C[ix(i - 1, k, ld)]) * dri;
        } else if (FAM == 1) {     // thr at vorticity points, rb.f:854-875
            if (i == 1)
                r = 1.5 * (th[ix(2, k, ld)] + th[ix(2, ka, ld)] - th[ix(1, k, ld)] - th[ix(1, ka, ld)]) * dri
                  - (-f23() * (th[ix(0, k, ld)] + th[ix(0, ka, ld)])
                     + 0.5 * (th[ix(1, k, ld)] + th[ix(1, ka, ld)])
                     + f16() * (th[ix(2, k, ld)] + th[ix(2, ka, ld)])) * dri;
            else if (i == nr - 1)
                r = 1.5 * (th[ix(nr, k, ld)] + th[ix(nr, ka, ld)] - th[ix(nr - 1, k, ld)] - th[ix(nr - 1, ka, ld)]) * dri
                  - (-f16() * (th[ix(nr - 1, k, ld)] + th[ix(nr - 1, ka, ld)])
                     - 0.5 * (th[ix(nr, k, ld)] + th[ix(nr, ka, ld)])
                     + f23() * (th[ix(nr + 1, k, ld)] + th[ix(nr + 1, ka, ld)])) * dri;
            else
                r = 1.5 * (th[ix(i + 1, k, ld)] + th[ix(i + 1, ka, ld)] - th[ix(i, k, ld)] - th[ix(i, ka, ld)]) * dri;
        } else {                   // thr at temperature points, rb.f:1244-1262
            if (i == 1)
                r = dri * (0.533333333 * th[ix(0, k, ld)] - 3.6 * th[ix(1, k, ld)]
                           + 2.966666667 * th[ix(2, k, ld)] + 0.1 * th[ix(4, k, ld)]);
            else if (i == nr)
                r = dri * (-0.1 * th[ix(nr - 3, k, ld)] - 2.966666667 * th[ix(nr - 1, k, ld)]
                           + 3.6 * th[ix(nr, k, ld)] - 0.533333333 * th[ix(nr + 1, k, ld)]);
            else
                r = 3.0 * (th[ix(i + 1, k, ld)] - th[ix(i - 1, k, ld)]) * dri;
        }
        if (j > 0) r = r - fact[j] * rprev;
        out[ix(i, k, ld)] = r;
        rprev = r;
    }
    double y = rprev / bp[n - 1];
    out[ix(n, k, ld)] = y;
    for (int j = n - 2; j >= 0; --j) {
        y = (out[ix(j + 1, k, ld)] - cc[j] * y) / bp[j];
        out[ix(j + 1, k, ld)] = y;
    }
}

// ---- batched cyclic solve along phi (one thread per row i), the
// reference ptdma (rb.f:3066-3116) on n = np+1 points.  Table layout
// (1-based, stride n+1): f1 | f2 | a | b | c | fA fB fC.
template <int FAM>
__device__ void ptdma_row(const CompactArgs &a, int i)
{
    const int np = a.np, ld = a.ld;
    const int n = np + 1;
    const int s = n + 1;
    const double *f1 = a.tab_p, *f2 = a.tab_p + s, *aa = a.tab_p + 2 * s;
    const double *bb = a.tab_p + 3 * s, *cc = a.tab_p + 4 * s;
    const double fA = a.tab_p[5 * s], fB = a.tab_p[5 * s + 1], fC = a.tab_p[5 * s + 2];
    const double dpi = a.dpi;
    const double *X = (FAM == 1) ? a.th : (FAM == 0 ? a.C : a.th);
    double *out = (FAM == 0) ? a.dCh : (FAM == 1) ? a.dthhC : a.dthhT;
    // rhs of entry j (1-based), column k = j-1
    auto rhs = [&](int j) -> double {
        const int k = j - 1;
        if (FAM == 1) {            // thh at vorticity points, rb.f:884-891
            const int ka = (k == np) ? 1 : k + 1;
            return 1.5 * (X[ix(i + 1, ka, ld)] + X[ix(i, ka, ld)] - X[ix(i + 1, k, ld)] - X[ix(i, k, ld)]) * dpi;
        }
        // Ch (rb.f:836-842), thh at temperature points (rb.f:1276-1282)
        if (k == np || k == 0) return 3.0 * (X[ix(i, 1, ld)] - X[ix(i, np - 1, ld)]) * dpi;
        return 3.0 * (X[ix(i, k + 1, ld)] - X[ix(i, k - 1, ld)]) * dpi;
    };
#define R_(j) out[ix(i, (j) - 1, ld)]
    double rn = rhs(n);
    double rm = rhs(1);
    R_(1) = rm;
    for (int j = 2; j <= n - 2; ++j) {
        double r = rhs(j) - f1[j] * rm;
        rn = rn - f2[j] * rm;
        R_(j) = r;
        rm = r;
    }
    double rn1 = rhs(n - 1) - fA * R_(n - 2);
    rn = rn - fB * R_(n - 2);
    rn = rn - fC * rn1;
    const double yn = rn / bb[n];
    double y = (rn1 - cc[n - 1] * yn) / bb[n - 1];
    R_(n) = yn;
    R_(n - 1) = y;
    for (int j = n - 2; j >= 1; --j) {
        y = (R_(j) - cc[j] * y - aa[j] * yn) / bb[j];
        R_(j) = y;
    }
#undef R_
}

__global__ void __launch_bounds__(WG) k_compact(CompactArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int fam = blockIdx.y;
    if (!(a.mask & (1 << fam))) return;
    switch (fam) {
    case 0: if (t <= a.np) tdma_column<0>(a, t); break;
    case 1: if (t <= a.np) tdma_column<1>(a, t); break;
    case 2: if (t <= a.np) tdma_column<2>(a, t); break;
    case 3: if (t + 1 <= a.nr - 1) ptdma_row<0>(a, t + 1); break;
    case 4: if (t + 1 <= a.nr - 1) ptdma_row<1>(a, t + 1); break;
    case 5: if (t + 1 <= a.nr) ptdma_row<2>(a, t + 1); break;
    }
}

struct DFArgs {
    const double *C, *Cp, *th, *thp, *uel, *vel;
    const double *dCr, *dCh, *dthrC, *dthhC, *dthrT, *dthhT;
    double *Cn, *Cf, *thn, *thf;
    const double *rC, *rT;
    int nr, np, ld, first, do_c, do_t, geom;
    double dri, dpi, dt, Pr, Ra, robert, Re;
};

// DuFort-Frankel vorticity (rb.f:900-982, first step rb.f:699-775) and
// temperature (rb.f:1292-1377, first step rb.f:1124-1204).  The Robert
// filtered current level goes to Cf/thf (the reference filters in place,
// rb.f:1041-1045 and 1435-1440, after all neighbours were read).
// Annulus (geom 1, velvor2dnewsch3.4.f): r = r0+i*dr at vorticity rows
// (ann.f:918) and r0+(i-1/2)*dr at temperature rows (ann.f:1331); RHS1
// gains Cra/r and thra/r, the leapfrog term uses Cp, and RHS2 gains the
// centrifugal buoyancy 2Ra(-thr u - thh v/r)/Re/Pr (ann.f:981-987, 1383).
__global__ void __launch_bounds__(WG) k_df(DFArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nr = a.nr, np = a.np, ld = a.ld;
    const int rows = nr + 2;
    if (t >= rows * (np + 1)) return;
    const int k = t / rows, i = t - k * rows;
    const int km = (k == 0) ? np - 1 : k - 1;
    const int ka = (k == np) ? 1 : k + 1;
    const double dri = a.dri, dpi = a.dpi, dt = a.dt, Ptl = a.Pr;
    const size_t o = ix(i, k, ld);
    const bool ann = (a.geom == 1);
    if (a.do_c) {
        const double Cc = a.C[o];
        if (i >= 1 && i <= nr - 1) {
            const double r = ann ? a.rC[i] : 1.0;
            const double ri = 1.0 / r;
            const double u = (a.uel[ix(i, ka, ld)] + a.uel[o]) * 0.5;
            const double v = (a.vel[o] + a.vel[ix(i + 1, k, ld)]) * 0.5;
            const double Chha = (a.C[ix(i, ka, ld)] + a.C[ix(i, km, ld)]) * (dpi * dpi);
            const double Ch = a.dCh[o];
            const double Crra = (a.C[ix(i - 1, k, ld)] + a.C[ix(i + 1, k, ld)]) * (dri * dri);
            const double cnr = -1.0 * (dri * dri);
            const double Cr = a.dCr[o];
            const double thh = a.dthhC[o];
            const double cnh = (-1.0 * (dpi * dpi)) * (ri * ri);
            double RHS1, cf1, RHS2;
            if (ann) {
                const double Cra = (a.C[ix(i + 1, k, ld)] - a.C[ix(i - 1, k, ld)]) * 0.5 * dri;
                const double thr = a.dthrC[o];
                if (a.first) {
                    RHS1 = Crra + Cra * ri + Chha * (ri * ri) + Cc / (Ptl * dt);
                    cf1 = 1.0 / (Ptl * dt) - cnr - cnh;
                } else {
                    RHS1 = Crra + Cra * ri + Chha * (ri * ri) + 0.5 * a.Cp[o] / (Ptl * dt);
                    cf1 = 0.5 / (Ptl * dt) - cnr - cnh;
                }
                RHS2 = -(a.Ra * thh) + 2.0 * a.Ra * (-(thr * u) - thh * v * ri) / a.Re / Ptl;
            } else {
                if (a.first) {
                    RHS1 = Crra + Chha * (ri * ri) + Cc / (Ptl * dt);
                    cf1 = 1.0 / (Ptl * dt) - cnr - cnh;
                } else {
                    RHS1 = Crra + Chha * (ri * ri) + 0.5 * Cc / (Ptl * dt);
                    cf1 = 0.5 / (Ptl * dt) - cnr - cnh;
                }
                RHS2 = -(a.Ra * thh);
            }
            const double ULT = (u * Cr + v * Ch * ri) / Ptl;
            const double cf2 = cnr + cnh;
            const double cn = (cf2 * a.Cp[o] - ULT - RHS2 + RHS1) / cf1;
            a.Cn[o] = cn;
            a.Cf[o] = a.first ? Cc : Cc + a.robert * (cn - 2.0 * Cc + a.Cp[o]);
        } else {
            a.Cn[o] = 0.0;
            a.Cf[o] = Cc;
        }
    }
    if (a.do_t) {
        const double T = a.th[o];
        if (i >= 1 && i <= nr) {
            const double r = ann ? a.rT[i] : 1.0;
            const double ri = 1.0 / r;
            const double u = (a.uel[o] + a.uel[ix(i - 1, k, ld)]) * 0.5;
            const double v = (a.vel[o] + a.vel[ix(i, km, ld)]) * 0.5;
            double thrra, cnr, thra;
            if (i == 1) {
                thrra = (3.2 * a.th[ix(0, k, ld)] + 2.0 * a.th[ix(2, k, ld)] - 0.2 * a.th[ix(3, k, ld)]) * (dri * dri);
                thra = (-4.0 * a.th[ix(0, k, ld)] / 3.0 + a.th[ix(2, k, ld)] / 3.0) * dri;
                cnr = -(2.5 * (dri * dri)) + 0.5 * dri * ri;
            } else if (i == nr) {
                thrra = (3.2 * a.th[ix(nr + 1, k, ld)] + 2.0 * a.th[ix(nr - 1, k, ld)] - 0.2 * a.th[ix(nr - 2, k, ld)]) * (dri * dri);
                thra = (-a.th[ix(nr - 1, k, ld)] / 3.0 + 4.0 * a.th[ix(nr + 1, k, ld)] / 3.0) * dri;
                cnr = -(2.5 * (dri * dri)) - 0.5 * dri * ri;
            } else {
                thrra = (a.th[ix(i - 1, k, ld)] + a.th[ix(i + 1, k, ld)]) * (dri * dri);
                thra = (a.th[ix(i + 1, k, ld)] - a.th[ix(i - 1, k, ld)]) * 0.5 * dri;
                cnr = -1.0 * (dri * dri);
            }
            const double thr = a.dthrT[o];
            const double thhha = (a.th[ix(i, ka, ld)] + a.th[ix(i, km, ld)]) * (dpi * dpi);
            const double thh = a.dthhT[o];
            const double cnp = -(1.0 * (dpi * dpi)) * (ri * ri);
            double RHS1, cf1;
            if (a.first) {
                RHS1 = ann ? thrra + thra * ri + thhha * (ri * ri) + T / dt
                           : thrra + thhha * (ri * ri) + T / dt;
                cf1 = 1.0 / dt - cnr - cnp;
            } else {
                RHS1 = ann ? thrra + thra * ri + thhha * (ri * ri) + 0.5 * a.thp[o] / dt
                           : thrra + thhha * (ri * ri) + 0.5 * a.thp[o] / dt;
                cf1 = 0.5 / dt - cnr - cnp;
            }
            const double ULT = u * thr + v * thh * ri;
            const double cf2 = cnr + cnp;
            const double tn = (cf2 * a.thp[o] - ULT + RHS1) / cf1;
            a.thn[o] = tn;
            a.thf[o] = a.first ? T : T + a.robert * (tn - 2.0 * T + a.thp[o]);
        } else {
            a.thf[o] = T;
        }
    }
}

// Wall heat flux, rb.f:1452-1500: the per-column terms are computed in
// parallel, then accumulated in the reference's order by one thread.
__global__ void __launch_bounds__(WG) k_heat(const double *th, int nr, int np, int ld,
                                             double dri, double dp, double *dst)
{
    __shared__ double xb[WG + 1], xt[WG + 1];
    double t1 = 0.0, t2 = 0.0;
    double first_b = 0.0, first_t = 0.0;
    for (int base = 1; base <= np; base += WG) {
        const int k = base + (int)threadIdx.x;
        if (k <= np) {
            xb[threadIdx.x] = 0.5 * dp * (-3.35238095 * th[ix(0, k, ld)] + 4.375 * th[ix(1, k, ld)]
                                          - 1.45833333 * th[ix(2, k, ld)] + 0.525 * th[ix(3, k, ld)]
                                          - 0.08928571 * th[ix(4, k, ld)]) * dri;
            xt[threadIdx.x] = 0.5 * dp * (-3.35238095 * th[ix(nr + 1, k, ld)] + 4.375 * th[ix(nr, k, ld)]
                                          - 1.45833333 * th[ix(nr - 1, k, ld)] + 0.525 * th[ix(nr - 2, k, ld)]
                                          - 0.08928571 * th[ix(nr - 3, k, ld)]) * dri;
        }
        if (threadIdx.x == 0) {   // term of column k+1 for the last k of this chunk
            const int kn = base + WG;
            if (kn <= np) {
                xb[WG] = 0.5 * dp * (-3.35238095 * th[ix(0, kn, ld)] + 4.375 * th[ix(1, kn, ld)]
                                     - 1.45833333 * th[ix(2, kn, ld)] + 0.525 * th[ix(3, kn, ld)]
                                     - 0.08928571 * th[ix(4, kn, ld)]) * dri;
                xt[WG] = 0.5 * dp * (-3.35238095 * th[ix(nr + 1, kn, ld)] + 4.375 * th[ix(nr, kn, ld)]
                                     - 1.45833333 * th[ix(nr - 1, kn, ld)] + 0.525 * th[ix(nr - 2, kn, ld)]
                                     - 0.08928571 * th[ix(nr - 3, kn, ld)]) * dri;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (base == 1) { first_b = xb[0]; first_t = xt[0]; }
            const int cnt = (np - base + 1 < WG) ? np - base + 1 : WG;
            for (int j = 0; j < cnt; ++j) {
                const int k = base + j;
                const double nb = (k == np) ? first_b : xb[j + 1];
                const double nt = (k == np) ? first_t : xt[j + 1];
                t1 = t1 + xb[j] + nb;
                t2 = t2 + xt[j] + nt;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { dst[0] = t1; dst[1] = t2; }
}

struct RhsArgs {
    const double *Cn, *vel, *velp, *uel, *uelp;
    double *vrh, *urh;
    double *fv, *uv, *fu, *uu;     // multigrid level 0 (coloured layout)
    unsigned long long *fnorm;     // [2]
    unsigned long long *cfl_bits;  // [2] reset here
    int *nclamp;                   // reset here
    const double *rC, *rT;
    int nr, np, ld, which, from_fields, geom;
    double dr, dp;
};

// multigrid level 0 is column-major: row j of column q at q*n + j
__device__ __forceinline__ size_t mgix(int j, int q, int n) { return (size_t)q * n + j; }

// rhsv (rb.f:1629-1676), rhsu (rb.f:1679-1721), the extrapolated initial
// guess of poisson2/3 (rb.f:1827-1835) scattered into the multigrid level 0,
// and the max-norms of the right-hand sides.
__global__ void __launch_bounds__(WG) k_rhs_guess(RhsArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nr = a.nr, np = a.np, ld = a.ld;
    const int rows = nr + 2;
    if (t == 0) {
        a.cfl_bits[0] = 0ull;
        a.cfl_bits[1] = 0ull;
        *a.nclamp = 0;
    }
    if (t >= rows * (np + 1)) return;
    const int k = t / rows, i = t - k * rows;
    const size_t o = ix(i, k, ld);
    if ((a.which & 1) && i >= 1 && i <= nr) {
        double v;
        if (a.from_fields) {
            v = a.vrh[o];
        } else {
            if (a.geom == 1) {          // ann.f:1663-1670
                const double r = a.rT[i];
                const double Cr = (a.Cn[o] - a.Cn[ix(i - 1, k, ld)]) / a.dr;
                if (i == 1) v = a.Cn[ix(1, k, ld)] / a.dr + a.Cn[ix(1, k, ld)] / r;
                else if (i == nr) v = -a.Cn[ix(nr - 1, k, ld)] / a.dr + a.Cn[ix(nr - 1, k, ld)] / r;
                else v = Cr + (a.Cn[o] + a.Cn[ix(i - 1, k, ld)]) / r;
            } else {
                if (i == 1) v = a.Cn[ix(1, k, ld)] / a.dr;
                else if (i == nr) v = -a.Cn[ix(nr - 1, k, ld)] / a.dr;
                else v = (a.Cn[o] - a.Cn[ix(i - 1, k, ld)]) / a.dr;
            }
            a.vrh[o] = v;
        }
        if (k >= 1) {
            const size_t m = mgix(i - 1, k - 1, nr);
            a.fv[m] = v;
            a.uv[m] = 2.0 * a.vel[o] - a.velp[o];
            atomicMax(&a.fnorm[0], (unsigned long long)__double_as_longlong(fabs(v)));
        }
    }
    if ((a.which & 2) && i >= 1 && i <= nr - 1) {
        double u;
        if (a.from_fields) {
            u = a.urh[o];
        } else {
            const int km = (k == 0) ? np - 1 : k - 1;
            const double Ch = (a.Cn[o] - a.Cn[ix(i, km, ld)]) / a.dp;
            u = -Ch / (a.geom == 1 ? a.rC[i] : 1.0);   // ann.f:1715-1719
            a.urh[o] = u;
        }
        if (k >= 1) {
            const size_t m = mgix(i - 1, k - 1, nr - 1);
            a.fu[m] = u;
            a.uu[m] = 2.0 * a.uel[o] - a.uelp[o];
            atomicMax(&a.fnorm[1], (unsigned long long)__double_as_longlong(fabs(u)));
        }
    }
}

struct UpdArgs {
    double *Cn, *thn;
    const double *vel, *uel;
    const double *rT;
    int geom;
    double *mT2, *mT, *mU2, *mV2;
    unsigned long long *cfl_bits;
    int *nclamp;
    int nr, np, ld, kt, stats;
    double dr, dp, dt;
};

// update (rb.f:1557-1626) on the new level; the rotation Cp<-C<-Cn,
// thp<-th<-thn is a pointer swap on the host.  Running means
// (rb.f:409-417) and CFL maxima (rb.f:421-458) are fused here.
__global__ void __launch_bounds__(WG) k_update(UpdArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nr = a.nr, np = a.np, ld = a.ld;
    const int rows = nr + 2;
    if (t >= rows * (np + 1)) return;
    const int k = t / rows, i = t - k * rows;
    const size_t o = ix(i, k, ld);
    if (i == 0) a.Cn[o] = (3.0 * a.vel[ix(1, k, ld)] - a.vel[ix(2, k, ld)] / 3.0) / a.dr;
    if (i == nr) a.Cn[o] = (a.vel[ix(nr - 1, k, ld)] / 3.0 - 3.0 * a.vel[ix(nr, k, ld)]) / a.dr;
    double T = a.thn[o];
    if (i == 0) T = 0.0;
    if (i == nr + 1) T = 1.0;
    int clamped = 0;
    if (T > 1.0) { T = 0.99; ++clamped; }
    if (T < 0.0) { T = 0.01; ++clamped; }
    a.thn[o] = T;
    if (clamped) atomicAdd(a.nclamp, clamped);
    if (a.stats) {
        const double U = a.uel[o], V = a.vel[o];
        if (a.kt >= 2) {
            const double tkti = 1.0 / (double)(a.kt - 1);
            const double km2 = (double)(a.kt - 2);
            a.mT2[o] = tkti * a.mT2[o] * km2 + tkti * (T * T);
            a.mT[o] = tkti * a.mT[o] * km2 + tkti * T;
            a.mU2[o] = tkti * a.mU2[o] * km2 + tkti * (U * U);
            a.mV2[o] = tkti * a.mV2[o] * km2 + tkti * (V * V);
        }
        if (i >= 1 && i <= nr) {
            const double c = fabs(U * a.dt / a.dr);
            atomicMax(&a.cfl_bits[0], (unsigned long long)__double_as_longlong(c));
        }
        if (i >= 1 && i <= nr - 1) {
            const double c = fabs(V * a.dt / ((a.geom == 1 ? a.rT[i] : 1.0) * a.dp));
            atomicMax(&a.cfl_bits[1], (unsigned long long)__double_as_longlong(c));
        }
    }
}

inline int nblk(long long n, int wg) { return (int)((n + wg - 1) / wg); }

}  // namespace

int launch_compact(vv_run *r)
{
    CompactArgs a;
    a.C = r->C; a.th = r->th;
    a.dCr = r->dCr; a.dCh = r->dCh; a.dthrC = r->dthrC; a.dthhC = r->dthhC;
    a.dthrT = r->dthrT; a.dthhT = r->dthhT;
    a.tab_r1 = r->tab_r1; a.tab_rT = r->tab_rT; a.tab_p = r->tab_p;
    a.nr = r->nr; a.np = r->np; a.ld = r->ld; a.mask = 0x3f;
    a.dri = 1.0 / r->dr; a.dpi = 1.0 / r->dp;
    const int len = (r->np + 1 > r->nr) ? r->np + 1 : r->nr;
    dim3 grid(nblk(len, 64), 6);
    hipEvent_t pe_;
    prof_begin(r, &pe_);
    hipLaunchKernelGGL(k_compact, grid, dim3(64), 0, r->stream, a);
    // read C and th, write six derivative fields
    prof_end(r, VV_K_COMPACT, pe_, 8.0 * 8.0 * (double)(r->nr + 2) * (r->np + 1));
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_compact_mask(vv_run *r, int mask)
{
    CompactArgs a;
    a.C = r->C; a.th = r->th;
    a.dCr = r->dCr; a.dCh = r->dCh; a.dthrC = r->dthrC; a.dthhC = r->dthhC;
    a.dthrT = r->dthrT; a.dthhT = r->dthhT;
    a.tab_r1 = r->tab_r1; a.tab_rT = r->tab_rT; a.tab_p = r->tab_p;
    a.nr = r->nr; a.np = r->np; a.ld = r->ld; a.mask = mask;
    a.dri = 1.0 / r->dr; a.dpi = 1.0 / r->dp;
    const int len = (r->np + 1 > r->nr) ? r->np + 1 : r->nr;
    dim3 grid(nblk(len, 64), 6);
    hipLaunchKernelGGL(k_compact, grid, dim3(64), 0, r->stream, a);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_df(vv_run *r, int first, int do_c, int do_t)
{
    DFArgs a;
    a.C = r->C; a.Cp = r->Cp; a.th = r->th; a.thp = r->thp; a.uel = r->uel; a.vel = r->vel;
    a.dCr = r->dCr; a.dCh = r->dCh; a.dthrC = r->dthrC; a.dthhC = r->dthhC; a.dthrT = r->dthrT;
    a.dthhT = r->dthhT;
    a.Cn = r->Cn; a.Cf = r->Cf; a.thn = r->thn; a.thf = r->thf;
    a.rC = r->rC; a.rT = r->rT;
    a.nr = r->nr; a.np = r->np; a.ld = r->ld; a.first = first; a.do_c = do_c; a.do_t = do_t;
    a.geom = r->cfg.geom;
    a.dri = 1.0 / r->dr; a.dpi = 1.0 / r->dp; a.dt = r->cfg.dt; a.Pr = r->cfg.Pr; a.Ra = r->cfg.Ra;
    a.robert = r->cfg.robert; a.Re = r->cfg.Re;
    const long long n = (long long)(r->nr + 2) * (r->np + 1);
    hipEvent_t pe_;
    prof_begin(r, &pe_);
    hipLaunchKernelGGL(k_df, dim3(nblk(n, WG)), dim3(WG), 0, r->stream, a);
    prof_end(r, VV_K_DF, pe_, 8.0 * 15.0 * (double)n);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_heat(vv_run *r, double *dst2)
{
    hipEvent_t pe_;
    prof_begin(r, &pe_);
    hipLaunchKernelGGL(k_heat, dim3(1), dim3(WG), 0, r->stream, (const double *)r->thf, r->nr,
                       r->np, r->ld, 1.0 / r->dr, r->dp, dst2);
    prof_end(r, VV_K_HEAT, pe_, 8.0 * 10.0 * (r->np + 1));
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_heat_from(vv_run *r, const double *th, double *dst2)
{
    hipLaunchKernelGGL(k_heat, dim3(1), dim3(WG), 0, r->stream, th, r->nr, r->np, r->ld,
                       1.0 / r->dr, r->dp, dst2);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

static int launch_rhs_common(vv_run *r, int which, int from_fields)
{
    RhsArgs a;
    a.from_fields = from_fields;
    a.Cn = r->Cn; a.vel = r->vel; a.velp = r->velp; a.uel = r->uel; a.uelp = r->uelp;
    a.vrh = r->vrh; a.urh = r->urh;
    a.fv = r->mg.prob[0].lev[0].f; a.uv = r->mg.prob[0].lev[0].u;
    a.fu = r->mg.prob[1].lev[0].f; a.uu = r->mg.prob[1].lev[0].u;
    a.fnorm = r->mg.fnorm;
    a.cfl_bits = r->cfl_bits;
    a.nclamp = r->diag_i;
    a.nr = r->nr; a.np = r->np; a.ld = r->ld; a.which = which;
    a.dr = r->dr; a.dp = r->dp;
    a.rC = r->rC; a.rT = r->rT; a.geom = r->cfg.geom;
    const long long n = (long long)(r->nr + 2) * (r->np + 1);
    hipEvent_t pe_;
    prof_begin(r, &pe_);
    hipLaunchKernelGGL(k_rhs_guess, dim3(nblk(n, WG)), dim3(WG), 0, r->stream, a);
    prof_end(r, VV_K_RHS, pe_, 8.0 * 11.0 * (double)n);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_rhs_guess(vv_run *r, int which) { return launch_rhs_common(r, which, 0); }
int launch_rhs_fields(vv_run *r, int which) { return launch_rhs_common(r, which, 1); }

int launch_update(vv_run *r, int kt)
{
    UpdArgs a;
    a.Cn = r->Cn; a.thn = r->thn; a.vel = r->vel; a.uel = r->uel;
    a.rT = r->rT; a.geom = r->cfg.geom;
    a.mT2 = r->mT2; a.mT = r->mT; a.mU2 = r->mU2; a.mV2 = r->mV2;
    a.cfl_bits = r->cfl_bits; a.nclamp = r->diag_i;
    a.nr = r->nr; a.np = r->np; a.ld = r->ld; a.kt = kt; a.stats = r->cfg.stats;
    a.dr = r->dr; a.dp = r->dp; a.dt = r->cfg.dt;
    const long long n = (long long)(r->nr + 2) * (r->np + 1);
    hipEvent_t pe_;
    prof_begin(r, &pe_);
    hipLaunchKernelGGL(k_update, dim3(nblk(n, WG)), dim3(WG), 0, r->stream, a);
    prof_end(r, VV_K_UPDATE, pe_, 8.0 * (r->cfg.stats ? 13.0 : 5.0) * (double)n);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

}  // namespace vv
