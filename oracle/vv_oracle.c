/*
 * vv_oracle.c -- CPU oracle (test infrastructure only) for the
 * velocity-vorticity time step of velvor2dperiodicrbconv.f ("rb.f").
 *
 * Each function cites the reference lines it restates.  Expressions keep the
 * reference's left-to-right evaluation order and its literal constants,
 * including the single-precision literals 2./3. and 1./6. (rb.f:660-668)
 * and the truncated decimals (rb.f:1088-1094, 1483-1495).
 * Build with -O2 -ffp-contract=off (oracle/Makefile).
 */
#include "vv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IX(i, k) ((size_t)(k) * (size_t)ld + (size_t)(i))

/* ---------------------------------------------------------------- banded */

/* rb.f:3046-3064 */
void vo_tdma(double *a, double *b, double *c, double *y, double *r, int n)
{
    for (int j = 1; j < n; ++j) {
        double fact = a[j] / b[j - 1];
        b[j] = b[j] - fact * c[j - 1];
        r[j] = r[j] - fact * r[j - 1];
    }
    y[n - 1] = r[n - 1] / b[n - 1];
    for (int j = n - 2; j >= 0; --j)
        y[j] = (r[j] - c[j] * y[j + 1]) / b[j];
}

/* rb.f:3066-3116, cyclic tridiagonal; a[0] couples row 1 to row n and
 * c[n-1] couples row n to row 1.  Indices below are 1-based as in the
 * reference: A(j) == a[j-1]. */
void vo_ptdma(double *a, double *b, double *c, double *y, double *r, int n)
{
#define A_(j) a[(j) - 1]
#define B_(j) b[(j) - 1]
#define C_(j) c[(j) - 1]
#define R_(j) r[(j) - 1]
#define Y_(j) y[(j) - 1]
    double fact;
    for (int j = 2; j <= n - 2; ++j) {
        fact = A_(j) / B_(j - 1);
        B_(j) = B_(j) - fact * C_(j - 1);
        A_(j) = -fact * A_(j - 1);
        R_(j) = R_(j) - fact * R_(j - 1);
        fact = C_(n) / B_(j - 1);
        C_(n) = -fact * C_(j - 1);
        B_(n) = B_(n) - fact * A_(j - 1);
        R_(n) = R_(n) - fact * R_(j - 1);
    }
    fact = A_(n - 1) / B_(n - 2);
    B_(n - 1) = B_(n - 1) - fact * C_(n - 2);
    C_(n - 1) = C_(n - 1) - fact * A_(n - 2);
    R_(n - 1) = R_(n - 1) - fact * R_(n - 2);
    fact = C_(n) / B_(n - 2);
    A_(n) = A_(n) - fact * C_(n - 2);
    B_(n) = B_(n) - fact * A_(n - 2);
    R_(n) = R_(n) - fact * R_(n - 2);
    fact = A_(n) / B_(n - 1);
    B_(n) = B_(n) - fact * C_(n - 1);
    R_(n) = R_(n) - fact * R_(n - 1);
    Y_(n) = R_(n) / B_(n);
    Y_(n - 1) = (R_(n - 1) - C_(n - 1) * Y_(n)) / B_(n - 1);
    for (int j = n - 2; j >= 1; --j)
        Y_(j) = (R_(j) - C_(j) * Y_(j + 1) - A_(j) * Y_(n)) / B_(j);
#undef A_
#undef B_
#undef C_
#undef R_
#undef Y_
}

/* rb.f:3118-3164, pentadiagonal without pivoting */
void vo_pdma(double *a, double *b, double *c, double *d, double *e,
             double *r, double *x, int n)
{
#define A_(j) a[(j) - 1]
#define B_(j) b[(j) - 1]
#define C_(j) c[(j) - 1]
#define D_(j) d[(j) - 1]
#define E_(j) e[(j) - 1]
#define R_(j) r[(j) - 1]
#define X_(j) x[(j) - 1]
    for (int k = 1; k <= n - 1; ++k) {
        D_(k) = D_(k) / C_(k);
        E_(k) = E_(k) / C_(k);
        R_(k) = R_(k) / C_(k);
        C_(k) = 1.0;
        C_(k + 1) = C_(k + 1) - B_(k + 1) * D_(k);
        if (k != n - 1) D_(k + 1) = D_(k + 1) - B_(k + 1) * E_(k);
        R_(k + 1) = R_(k + 1) - B_(k + 1) * R_(k);
        B_(k + 1) = 0.0;
        if (k != n - 1) {
            B_(k + 2) = B_(k + 2) - A_(k + 2) * D_(k);
            C_(k + 2) = C_(k + 2) - A_(k + 2) * E_(k);
            R_(k + 2) = R_(k + 2) - A_(k + 2) * R_(k);
        }
    }
    X_(n) = R_(n) / C_(n);
    X_(n - 1) = R_(n - 1) - D_(n - 1) * X_(n);
    for (int j = n - 2; j >= 1; --j)
        X_(j) = R_(j) - D_(j) * X_(j + 1) - E_(j) * X_(j + 2);
#undef A_
#undef B_
#undef C_
#undef D_
#undef E_
#undef R_
#undef X_
}

/* ------------------------------------------------- compact derivatives */

/* single-precision literals of the reference (rb.f:660-668) */
static const float F23 = 2.f / 3.f;
static const float F16 = 1.f / 6.f;

/* Cr: rb.f:810-830 (same as Cvorticity1 rb.f:607-629) */
static void compact_Cr(const vo_params *p, const double *C, double *out,
                       double *wa, double *wb, double *wc, double *wr, double *ws)
{
    const int nr = p->nr, np = p->np, ld = p->ld;
    const double dri = 1.0 / p->dr;
    for (int k = 0; k <= np; ++k) {
        for (int i = 1; i <= nr - 1; ++i) {
            wa[i - 1] = 1.0; wb[i - 1] = 4.0; wc[i - 1] = 1.0;
            if (i == 1)
                wr[0] = 3.0 * (C[IX(2, k)] - C[IX(0, k)]) * dri
                      - (-1.5 * C[IX(0, k)] + 2.0 * C[IX(1, k)] - 0.5 * C[IX(2, k)]) * dri;
            else if (i == nr - 1)
                wr[i - 1] = 3.0 * (C[IX(nr, k)] - C[IX(nr - 2, k)]) * dri
                          - (0.5 * C[IX(nr - 2, k)] - 2.0 * C[IX(nr - 1, k)] + 1.5 * C[IX(nr, k)]) * dri;
            else
                wr[i - 1] = 3.0 * (C[IX(i + 1, k)] - C[IX(i - 1, k)]) * dri;
        }
        vo_tdma(wa, wb, wc, ws, wr, nr - 1);
        for (int i = 1; i <= nr - 1; ++i) out[IX(i, k)] = ws[i - 1];
    }
}

/* Ch (cyclic on np+1 points): rb.f:833-851 */
static void compact_Ch(const vo_params *p, const double *C, int i0, int i1, double *out,
                       double *wa, double *wb, double *wc, double *wr, double *ws)
{
    const int np = p->np, ld = p->ld;
    const double dpi = 1.0 / p->dp;
    for (int i = i0; i <= i1; ++i) {
        for (int k = 0; k <= np; ++k) {
            wa[k] = 1.0; wb[k] = 4.0; wc[k] = 1.0;
            if (k == np)
                wr[np] = 3.0 * (C[IX(i, 1)] - C[IX(i, np - 1)]) * dpi;
            else if (k == 0)
                wr[0] = 3.0 * (C[IX(i, 1)] - C[IX(i, np - 1)]) * dpi;
            else
                wr[k] = 3.0 * (C[IX(i, k + 1)] - C[IX(i, k - 1)]) * dpi;
        }
        vo_ptdma(wa, wb, wc, ws, wr, np + 1);
        for (int k = 0; k <= np; ++k) out[IX(i, k)] = ws[k];
    }
}

/* thr at vorticity points, half derivative (rb.f:854-880) */
static void compact_thrC(const vo_params *p, const double *th, double *out,
                         double *wa, double *wb, double *wc, double *wr, double *ws)
{
    const int nr = p->nr, np = p->np, ld = p->ld;
    const double dri = 1.0 / p->dr;
    for (int k = 0; k <= np; ++k) {
        int ka = k + 1;
        if (k == np) ka = 1;
        for (int i = 1; i <= nr - 1; ++i) {
            wa[i - 1] = 1.0; wb[i - 1] = 4.0; wc[i - 1] = 1.0;
            if (i == 1)
                wr[0] = 1.5 * (th[IX(2, k)] + th[IX(2, ka)] - th[IX(1, k)] - th[IX(1, ka)]) * dri
                      - (-(double)F23 * (th[IX(0, k)] + th[IX(0, ka)])
                         + 0.5 * (th[IX(1, k)] + th[IX(1, ka)])
                         + (double)F16 * (th[IX(2, k)] + th[IX(2, ka)])) * dri;
            else if (i == nr - 1)
                wr[i - 1] = 1.5 * (th[IX(nr, k)] + th[IX(nr, ka)] - th[IX(nr - 1, k)] - th[IX(nr - 1, ka)]) * dri
                          - (-(double)F16 * (th[IX(nr - 1, k)] + th[IX(nr - 1, ka)])
                             - 0.5 * (th[IX(nr, k)] + th[IX(nr, ka)])
                             + (double)F23 * (th[IX(nr + 1, k)] + th[IX(nr + 1, ka)])) * dri;
            else
                wr[i - 1] = 1.5 * (th[IX(i + 1, k)] + th[IX(i + 1, ka)] - th[IX(i, k)] - th[IX(i, ka)]) * dri;
        }
        vo_tdma(wa, wb, wc, ws, wr, nr - 1);
        for (int i = 1; i <= nr - 1; ++i) out[IX(i, k)] = ws[i - 1];
    }
}

/* thh at vorticity points, half derivative, cyclic np+1 (rb.f:882-897) */
static void compact_thhC(const vo_params *p, const double *th, double *out,
                         double *wa, double *wb, double *wc, double *wr, double *ws)
{
    const int nr = p->nr, np = p->np, ld = p->ld;
    const double dpi = 1.0 / p->dp;
    for (int i = 1; i <= nr - 1; ++i) {
        for (int k = 0; k <= np; ++k) {
            wa[k] = 1.0; wb[k] = 4.0; wc[k] = 1.0;
            int ka = k + 1;
            if (k == np) ka = 1;
            wr[k] = 1.5 * (th[IX(i + 1, ka)] + th[IX(i, ka)] - th[IX(i + 1, k)] - th[IX(i, k)]) * dpi;
        }
        vo_ptdma(wa, wb, wc, ws, wr, np + 1);
        for (int k = 0; k <= np; ++k) out[IX(i, k)] = ws[k];
    }
}

/* thr at temperature points with one-sided closures (rb.f:1244-1268) */
static void compact_thrT(const vo_params *p, const double *th, double *out,
                         double *wa, double *wb, double *wc, double *wr, double *ws)
{
    const int nr = p->nr, np = p->np, ld = p->ld;
    const double dri = 1.0 / p->dr;
    for (int k = 0; k <= np; ++k) {
        for (int i = 1; i <= nr; ++i) {
            wa[i - 1] = 1.0; wb[i - 1] = 4.0; wc[i - 1] = 1.0;
            if (i == 1) {
                wb[0] = 1.0; wc[0] = 2.0;
                wr[0] = dri * (0.533333333 * th[IX(0, k)] - 3.6 * th[IX(1, k)]
                               + 2.966666667 * th[IX(2, k)] + 0.1 * th[IX(4, k)]);
            } else if (i == nr) {
                wa[i - 1] = 2.0; wb[i - 1] = 1.0;
                wr[i - 1] = dri * (-0.1 * th[IX(nr - 3, k)] - 2.966666667 * th[IX(nr - 1, k)]
                                   + 3.6 * th[IX(nr, k)] - 0.533333333 * th[IX(nr + 1, k)]);
            } else {
                wr[i - 1] = 3.0 * (th[IX(i + 1, k)] - th[IX(i - 1, k)]) * dri;
            }
        }
        vo_tdma(wa, wb, wc, ws, wr, nr);
        for (int i = 1; i <= nr; ++i) out[IX(i, k)] = ws[i - 1];
    }
}

/* -------------------------------------------------------- explicit step */

/* rb.f:584-778 (first step) and rb.f:786-1048 (DuFort-Frankel/leapfrog) */
void vo_cvorticity(const vo_params *p, vo_fields *f, int first)
{
    const int nr = p->nr, np = p->np, ld = p->ld;
    const double dri = 1.0 / p->dr, dpi = 1.0 / p->dp;
    const double Ptl = p->Pr, dt = p->dt, Ra = p->Ra;
    int nmax = (nr > np ? nr : np) + 2;
    double *w = (double *)malloc(sizeof(double) * 5 * nmax);
    double *wa = w, *wb = w + nmax, *wc = w + 2 * nmax, *wr = w + 3 * nmax, *ws = w + 4 * nmax;
    double *C = f->C, *Cp = f->Cp, *Cn = f->Cn, *th = f->th;
    double *uel = f->uel, *vel = f->vel;

    compact_Cr(p, C, f->A, wa, wb, wc, wr, ws);
    compact_Ch(p, C, 1, nr - 1, f->B, wa, wb, wc, wr, ws);
    compact_thrC(p, th, f->An, wa, wb, wc, wr, ws);
    compact_thhC(p, th, f->Bn, wa, wb, wc, wr, ws);

    const int ann = (p->geom == 1);
    for (int i = 1; i <= nr - 1; ++i) {
        /* r = 1 in the box (rb.f:922); r = r0 + i*dr in the annulus (ann.f:918) */
        const double r = ann ? p->r0 + (double)i * p->dr : 1.0;
        const double ri = 1.0 / r;
        for (int k = 0; k <= np; ++k) {
            int ip = i + 1;
            int km = k - 1, ka = k + 1;
            if (k == 0) km = np - 1;
            if (k == np) ka = 1;
            double u = (uel[IX(i, ka)] + uel[IX(i, k)]) * 0.5;
            double v = (vel[IX(i, k)] + vel[IX(ip, k)]) * 0.5;
            double thr = f->An[IX(i, k)];
            double Chha = (C[IX(i, ka)] + C[IX(i, km)]) * (dpi * dpi);
            double Ch = f->B[IX(i, k)];
            double Crra = (C[IX(i - 1, k)] + C[IX(i + 1, k)]) * (dri * dri);
            double cnr = -1.0 * (dri * dri);
            double Cra = (C[IX(i + 1, k)] - C[IX(i - 1, k)]) * 0.5 * dri;
            double Cr = f->A[IX(i, k)];
            double thh = f->Bn[IX(i, k)];
            double cnh = (-1.0 * (dpi * dpi)) * (ri * ri);
            double RHS1, cf1, RHS2;
            if (ann) {
                /* ann.f:981-987 (first step ann.f:734-740): Cra/r term, the
                 * level Cp in the leapfrog term, centrifugal buoyancy */
                if (first) {
                    RHS1 = Crra + Cra * ri + Chha * (ri * ri) + C[IX(i, k)] / (Ptl * dt);
                    cf1 = 1.0 / (Ptl * dt) - cnr - cnh;
                } else {
                    RHS1 = Crra + Cra * ri + Chha * (ri * ri) + 0.5 * Cp[IX(i, k)] / (Ptl * dt);
                    cf1 = 0.5 / (Ptl * dt) - cnr - cnh;
                }
                RHS2 = -(Ra * thh) + 2.0 * Ra * (-(thr * u) - thh * v * ri) / p->Re / Ptl;
            } else {
                if (first) {
                    RHS1 = Crra + Chha * (ri * ri) + C[IX(i, k)] / (Ptl * dt);
                    cf1 = 1.0 / (Ptl * dt) - cnr - cnh;
                } else {
                    RHS1 = Crra + Chha * (ri * ri) + 0.5 * C[IX(i, k)] / (Ptl * dt);
                    cf1 = 0.5 / (Ptl * dt) - cnr - cnh;
                }
                RHS2 = -(Ra * thh);
            }
            double ULT = (u * Cr + v * Ch * ri) / Ptl;
            double cf2 = cnr + cnh;
            Cn[IX(i, k)] = (cf2 * Cp[IX(i, k)] - ULT - RHS2 + RHS1) / cf1;
            if (!first) f->Ap[IX(i, k)] = Cn[IX(i, k)];
        }
    }
    if (!first) {
        /* Robert filter, rb.f:1041-1045 (0.02; 0.05 in ann.f:1055) */
        for (int i = 1; i <= nr - 1; ++i)
            for (int k = 0; k <= np; ++k)
                C[IX(i, k)] = C[IX(i, k)] + p->robert * (Cn[IX(i, k)] - 2.0 * C[IX(i, k)] + Cp[IX(i, k)]);
    }
    free(w);
}

/* rb.f:1056-1212 (first step) and rb.f:1220-1443 */
void vo_temperature(const vo_params *p, vo_fields *f, int first)
{
    const int nr = p->nr, np = p->np, ld = p->ld;
    const double dri = 1.0 / p->dr, dpi = 1.0 / p->dp;
    const double dt = p->dt;
    int nmax = (nr > np ? nr : np) + 2;
    double *w = (double *)malloc(sizeof(double) * 5 * nmax);
    double *wa = w, *wb = w + nmax, *wc = w + 2 * nmax, *wr = w + 3 * nmax, *ws = w + 4 * nmax;
    double *th = f->th, *thp = f->thp, *thn = f->thn;
    double *uel = f->uel, *vel = f->vel;

    compact_thrT(p, th, f->A, wa, wb, wc, wr, ws);
    compact_Ch(p, th, 1, nr, f->B, wa, wb, wc, wr, ws);   /* thh, rb.f:1271-1289 */

    const int ann = (p->geom == 1);
    for (int i = 1; i <= nr; ++i) {
        /* r = 1 in the box; r = r0 + (i-1/2)*dr in the annulus (ann.f:1331) */
        const double r = ann ? p->r0 + ((double)i - 0.5) * p->dr : 1.0;
        const double ri = 1.0 / r;
        for (int k = 0; k <= np; ++k) {
            int im = i - 1;
            int km = k - 1, ka = k + 1;
            if (k == 0) km = np - 1;
            if (k == np) ka = 1;
            double u = (uel[IX(i, k)] + uel[IX(im, k)]) * 0.5;
            double v = (vel[IX(i, k)] + vel[IX(i, km)]) * 0.5;
            double thrra, cnr, thra;
            if (i == 1) {
                thrra = (3.2 * th[IX(i - 1, k)] + 2.0 * th[IX(i + 1, k)] - 0.2 * th[IX(i + 2, k)]) * (dri * dri);
                thra = (-4.0 * th[IX(i - 1, k)] / 3.0 + th[IX(i + 1, k)] / 3.0) * dri;
                cnr = -(2.5 * (dri * dri)) + 0.5 * dri * ri;
            } else if (i == nr) {
                thrra = (3.2 * th[IX(i + 1, k)] + 2.0 * th[IX(i - 1, k)] - 0.2 * th[IX(i - 2, k)]) * (dri * dri);
                thra = (-th[IX(i - 1, k)] / 3.0 + 4.0 * th[IX(i + 1, k)] / 3.0) * dri;
                cnr = -(2.5 * (dri * dri)) - 0.5 * dri * ri;
            } else {
                thrra = (th[IX(i - 1, k)] + th[IX(i + 1, k)]) * (dri * dri);
                thra = (th[IX(i + 1, k)] - th[IX(i - 1, k)]) * 0.5 * dri;
                cnr = -1.0 * (dri * dri);
            }
            double thr = f->A[IX(i, k)];
            double thhha = (th[IX(i, ka)] + th[IX(i, km)]) * (dpi * dpi);
            double thh = f->B[IX(i, k)];
            double cnp = -(1.0 * (dpi * dpi)) * (ri * ri);
            double RHS1, cf1;
            if (first) {
                if (ann) RHS1 = thrra + thra * ri + thhha * (ri * ri) + th[IX(i, k)] / dt;
                else RHS1 = thrra + thhha * (ri * ri) + th[IX(i, k)] / dt;
                cf1 = 1.0 / dt - cnr - cnp;
            } else {
                if (ann) RHS1 = thrra + thra * ri + thhha * (ri * ri) + 0.5 * thp[IX(i, k)] / dt;
                else RHS1 = thrra + thhha * (ri * ri) + 0.5 * thp[IX(i, k)] / dt;
                cf1 = 0.5 / dt - cnr - cnp;
            }
            double ULT = u * thr + v * thh * ri;
            double cf2 = cnr + cnp;
            thn[IX(i, k)] = (cf2 * thp[IX(i, k)] - ULT + RHS1) / cf1;
            if (!first) f->Ap[IX(i, k)] = thn[IX(i, k)];
        }
    }
    if (!first) {
        /* Robert filter, rb.f:1435-1440 */
        for (int i = 1; i <= nr; ++i)
            for (int k = 0; k <= np; ++k)
                th[IX(i, k)] = th[IX(i, k)] + p->robert * (thn[IX(i, k)] - 2.0 * th[IX(i, k)] + thp[IX(i, k)]);
    }
    free(w);
}

/* rb.f:1452-1500: wall heat flux, trapezoid over k with a 5-point one-sided
 * derivative (truncated Fornberg weights) */
void vo_heat(const vo_params *p, const double *th, double *trans1, double *trans2)
{
    const int nr = p->nr, np = p->np, ld = p->ld;
    const double dri = 1.0 / p->dr, dp = p->dp;
    double t1 = 0.0, t2 = 0.0;
    for (int k = 1; k <= np; ++k) {
        int ka = k + 1;
        if (k == np) ka = 1;
        t1 = t1 + 0.5 * dp * (-3.35238095 * th[IX(0, k)] + 4.375 * th[IX(1, k)]
                              - 1.45833333 * th[IX(2, k)] + 0.525 * th[IX(3, k)]
                              - 0.08928571 * th[IX(4, k)]) * dri
                + 0.5 * dp * (-3.35238095 * th[IX(0, ka)] + 4.375 * th[IX(1, ka)]
                              - 1.45833333 * th[IX(2, ka)] + 0.525 * th[IX(3, ka)]
                              - 0.08928571 * th[IX(4, ka)]) * dri;
        t2 = t2 + 0.5 * dp * (-3.35238095 * th[IX(nr + 1, k)] + 4.375 * th[IX(nr, k)]
                              - 1.45833333 * th[IX(nr - 1, k)] + 0.525 * th[IX(nr - 2, k)]
                              - 0.08928571 * th[IX(nr - 3, k)]) * dri
                + 0.5 * dp * (-3.35238095 * th[IX(nr + 1, ka)] + 4.375 * th[IX(nr, ka)]
                              - 1.45833333 * th[IX(nr - 1, ka)] + 0.525 * th[IX(nr - 2, ka)]
                              - 0.08928571 * th[IX(nr - 3, ka)]) * dri;
    }
    *trans1 = t1;
    *trans2 = t2;
}

/* kinetic energy, the build's definition (vv_oracle.h) */
double vo_energy(const vo_params *p, const double *uel, const double *vel)
{
    const int nr = p->nr, np = p->np, ld = p->ld;
    double s = 0.0;
    for (int k = 0; k < np; ++k) {
        for (int i = 1; i <= nr; ++i) {
            const double w = p->geom ? p->r0 + ((double)i - 0.5) * p->dr : 1.0;
            s += w * vel[IX(i, k)] * vel[IX(i, k)];
        }
        for (int i = 1; i <= nr - 1; ++i) {
            const double w = p->geom ? p->r0 + (double)i * p->dr : 1.0;
            s += w * uel[IX(i, k)] * uel[IX(i, k)];
        }
    }
    return 0.5 * p->dr * p->dp * s;
}

/* rb.f:1629-1676 */
void vo_rhsv(const vo_params *p, const double *Cn, double *vrh)
{
    const int nr = p->nr, np = p->np, ld = p->ld;
    const double dr = p->dr;
    const int ann = (p->geom == 1);
    for (int i = 1; i <= nr; ++i) {
        const double r = ann ? p->r0 + ((double)i - 0.5) * dr : 1.0;
        for (int k = 0; k <= np; ++k) {
            double Cr = (Cn[IX(i, k)] - Cn[IX(i - 1, k)]) / dr;
            if (ann) {   /* ann.f:1663-1670 */
                if (i == 1) vrh[IX(i, k)] = Cn[IX(1, k)] / dr + Cn[IX(1, k)] / r;
                else if (i == nr) vrh[IX(i, k)] = -Cn[IX(nr - 1, k)] / dr + Cn[IX(nr - 1, k)] / r;
                else vrh[IX(i, k)] = Cr + (Cn[IX(i, k)] + Cn[IX(i - 1, k)]) / r;
            } else {
                if (i == 1) vrh[IX(i, k)] = Cn[IX(1, k)] / dr;
                else if (i == nr) vrh[IX(i, k)] = -Cn[IX(nr - 1, k)] / dr;
                else vrh[IX(i, k)] = Cr;
            }
        }
    }
}

/* rb.f:1679-1721 */
void vo_rhsu(const vo_params *p, const double *Cn, double *urh)
{
    const int nr = p->nr, np = p->np, ld = p->ld;
    const double dp = p->dp;
    const int ann = (p->geom == 1);
    for (int i = 1; i <= nr - 1; ++i) {
        const double r = ann ? p->r0 + (double)i * p->dr : 1.0;   /* ann.f:1715 */
        for (int k = 0; k <= np; ++k) {
            int km = k - 1;
            if (k == 0) km = np - 1;
            double Ch = (Cn[IX(i, k)] - Cn[IX(i, km)]) / dp;
            urh[IX(i, k)] = -Ch / r;
        }
    }
}

/* rb.f:1557-1626: wall vorticity, temperature walls, time-level rotation
 * and the temperature clamp.  Returns the number of clamped points. */
int vo_update(const vo_params *p, vo_fields *f)
{
    const int nr = p->nr, np = p->np, ld = p->ld;
    const double dr = p->dr;
    double *vel = f->vel, *Cn = f->Cn, *thn = f->thn;
    int nclamp = 0;
    for (int k = 0; k <= np; ++k) {
        Cn[IX(0, k)] = (3.0 * vel[IX(1, k)] - vel[IX(2, k)] / 3.0) / dr;
        Cn[IX(nr, k)] = (vel[IX(nr - 1, k)] / 3.0 - 3.0 * vel[IX(nr, k)]) / dr;
    }
    for (int k = 0; k <= np; ++k) {
        thn[IX(0, k)] = 0.0;
        thn[IX(nr + 1, k)] = 1.0;
    }
    for (int i = 0; i <= nr + 1; ++i)
        for (int k = 0; k <= np; ++k) {
            f->Cp[IX(i, k)] = f->C[IX(i, k)];
            f->C[IX(i, k)] = Cn[IX(i, k)];
            f->thp[IX(i, k)] = f->th[IX(i, k)];
            if (thn[IX(i, k)] > 1.0) { thn[IX(i, k)] = 0.99; ++nclamp; }
            if (thn[IX(i, k)] < 0.0) { thn[IX(i, k)] = 0.01; ++nclamp; }
            f->th[IX(i, k)] = thn[IX(i, k)];
        }
    return nclamp;
}

/* ------------------------------------------------------------ multigrid */

/* rb.f:1724-1793 */
int vo_mg_init(vo_mg *mg, int nr, int np, double dr, double dp, int nlev_cap)
{
    memset(mg, 0, sizeof(*mg));
    int nlev = 0, itest = nr, itest2 = nr;
    for (int i = 2; i <= 8; ++i) {
        itest = itest / 2;
        if (itest * 2 - itest2 == 0) {
            if (itest2 >= 4) nlev = nlev + 1;
            itest2 = itest;
        } else {
            break;
        }
    }
    if (nlev >= nlev_cap + 1) nlev = nlev_cap;
    if (nlev > VO_MAXLEV) nlev = VO_MAXLEV;
    mg->nlev = nlev;
    mg->nlev_cap = nlev_cap;
    mg->geom = 0;
    mg->r0 = 0.0;
    mg->nsweep = 1;
    mg->zero_seam = 1;
    for (int l = 1; l <= nlev; ++l) {
        int p2 = 1 << l;
        mg->n1[l - 1] = nr * 2 / p2;
        mg->n3[l - 1] = np * 2 / p2;
        mg->hr[l - 1] = dr * (double)p2 * 0.5;
        mg->hp[l - 1] = dp * (double)p2 * 0.5;
        mg->hhr[l - 1] = mg->hr[l - 1] * mg->hr[l - 1];
        mg->hhp[l - 1] = mg->hp[l - 1] * mg->hp[l - 1];
        size_t sz = (size_t)(mg->n1[l - 1] + 1) * (size_t)(mg->n3[l - 1] + 1);
        mg->v[l - 1] = (double *)calloc(sz, sizeof(double));
        mg->f[l - 1] = (double *)calloc(sz, sizeof(double));
    }
    mg->ld = nr + 2;
    mg->ncols = np + 1;
    size_t sz = (size_t)mg->ld * (size_t)mg->ncols;
    mg->psif = (double *)calloc(sz, sizeof(double));
    mg->psic = (double *)calloc(sz, sizeof(double));
    mg->rc = (double *)calloc(sz, sizeof(double));
    mg->rf = (double *)calloc(sz, sizeof(double));
    mg->tempa = (double *)calloc(sz, sizeof(double));
    return nlev;
}

void vo_mg_set_geom(vo_mg *mg, int geom, double r0, int nsweep, int zero_seam)
{
    mg->geom = geom;
    mg->r0 = r0;
    mg->nsweep = nsweep;
    mg->zero_seam = zero_seam;
}

void vo_mg_free(vo_mg *mg)
{
    for (int l = 0; l < mg->nlev; ++l) { free(mg->v[l]); free(mg->f[l]); }
    free(mg->psif); free(mg->psic); free(mg->rc); free(mg->rf); free(mg->tempa);
    memset(mg, 0, sizeof(*mg));
}

/* rb.f:2805-2829 / 2831-2845: 2D scratch <-> per-level store */
static void trans1(const vo_mg *mg, const double *a2, double *a1, int n1, int n3)
{
    const int ld = mg->ld;
    for (int k = 0; k <= n3; ++k)
        for (int i = 0; i <= n1; ++i)
            a1[(size_t)k * (n1 + 1) + i] = a2[IX(i, k)];
}
static void trans2(const vo_mg *mg, const double *a1, double *a2, int n1, int n3)
{
    const int ld = mg->ld;
    for (int k = 0; k <= n3; ++k)
        for (int i = 0; i <= n1; ++i)
            a2[IX(i, k)] = a1[(size_t)k * (n1 + 1) + i];
}

/* relax2 (rb.f:2063-2371; ann.f:2071-2359) for which == 2, relax3
 * (rb.f:2374-2600; ann.f:2365-2591) for 3.  klev is 1-based.  In the
 * annulus the operators carry the cylindrical terms with 1/r evaluated at the
 * level's own cell (v) or face (u) radii, and every call does 3 sweeps
 * (ann.f:2140, 2411).  Returns nconv. */
static int relax(vo_mg *mg, int which, int nrel, int klev, double tolcon)
{
    const int ld = mg->ld;
    const int L = klev - 1;
    const int n1 = mg->n1[L], n3 = mg->n3[L];
    const int ann = (mg->geom == 1);
    double *psif = mg->psif, *rf = mg->rf, *rc = mg->rc;
    int nconv = 0;
    double resmax = 0.0;
    trans2(mg, mg->f[L], rf, n1, n3);
    const double hhri = 1.0 / mg->hhr[L];
    const double hri = 1.0 / mg->hr[L];
    const double hhpi = 1.0 / mg->hhp[L];
    const int nn = n1 + 2;
    double *w = (double *)malloc(sizeof(double) * 14 * nn);
    double *pa = w, *pb = w + nn, *pc = w + 2 * nn, *pd = w + 3 * nn, *pe = w + 4 * nn;
    double *ta = w + 5 * nn, *tb = w + 6 * nn, *tc = w + 7 * nn, *tr = w + 8 * nn;
    double *ts = w + 9 * nn, *pr = w + 10 * nn, *ps = w + 11 * nn, *rvs = w + 12 * nn;
    double *rvi = w + 13 * nn;
    const int nsw = ann ? mg->nsweep : nrel;

    /* line coefficients (rb.f:2120-2156 at k=1; ann.f:2097-2126) */
    if (which == 2) {
        for (int i = 1; i <= n1; ++i) {
            double rval = ann ? mg->r0 + ((double)i - 0.5) * mg->hr[L] : 1.0;
            rvi[i] = 1.0 / rval;
            rvs[i] = rvi[i] * rvi[i];
            if (i == 1) {
                pa[i] = 0.0; pb[i] = 0.0;
                if (ann) {
                    pc[i] = -2.0 * hhri - 2.0 * hhpi * rvs[i] + rvs[i];
                    pd[i] = 5.0 / 3.0 * hhri + 4.0 / 3.0 * hri * rvi[i];
                } else {
                    pc[i] = -2.0 * hhri - 2.0 * hhpi * rvs[i];
                    pd[i] = 5.0 / 3.0 * hhri;
                }
                pe[i] = -0.2 * hhri;
            } else if (i == n1) {
                pa[i] = -0.2 * hhri;
                if (ann) {
                    pb[i] = 5.0 / 3.0 * hhri - 4.0 / 3.0 * hri * rvi[i];
                    pc[i] = -2.0 * hhri - 2.0 * hhpi * rvs[i] + rvs[i];
                } else {
                    pb[i] = 5.0 / 3.0 * hhri;
                    pc[i] = -2.0 * hhri - 2.0 * hhpi * rvs[i];
                }
                pd[i] = 0.0; pe[i] = 0.0;
            } else {
                pa[i] = 0.0;
                if (ann) {
                    pb[i] = hhri - 1.5 * hri * rvi[i];
                    pc[i] = -(2.0 * hhri + 2.0 * hhpi * rvs[i]) + rvs[i];
                    pd[i] = hhri + 1.5 * hri * rvi[i];
                } else {
                    pb[i] = hhri;
                    pc[i] = -(2.0 * hhri + 2.0 * hhpi * rvs[i]);
                    pd[i] = hhri;
                }
                pe[i] = 0.0;
            }
        }
    } else {
        for (int i = 1; i <= n1 - 1; ++i) {
            double rval = ann ? (double)i * mg->hr[L] + mg->r0 : 1.0;
            rvi[i] = 1.0 / rval;
            rvs[i] = rvi[i] * rvi[i];
            if (ann) {
                pa[i] = hhri - 1.5 * hri * rvi[i];
                pb[i] = -(2.0 * hhri - rvs[i] + 2.0 * hhpi * rvs[i]);
                pc[i] = hhri + 1.5 * hri * rvi[i];
            } else {
                pa[i] = hhri;
                pb[i] = -(2.0 * hhri + 2.0 * hhpi * rvs[i]);
                pc[i] = hhri;
            }
        }
    }

    for (int irel = 1; irel <= nsw; ++irel) {
        if (which == 2) {
            for (int k = 1; k <= n3; ++k) {
                int ka = k + 1, km = k - 1;
                if (k == n3) ka = 1;
                for (int i = 1; i <= n1; ++i) {
                    ta[i] = pa[i]; tb[i] = pb[i]; tc[i] = pc[i]; tr[i] = pd[i]; ts[i] = pe[i];
                    pr[i] = -(psif[IX(i, ka)] + psif[IX(i, km)]) * hhpi * rvs[i] + rf[IX(i, k)];
                }
                vo_pdma(ta + 1, tb + 1, tc + 1, tr + 1, ts + 1, pr + 1, ps + 1, n1);
                for (int i = 1; i <= n1; ++i) {
                    psif[IX(i, k)] = ps[i];
                    if (k == n3) psif[IX(i, 0)] = psif[IX(i, k)];
                }
            }
            resmax = 0.0;
            for (int k = 1; k <= n3; ++k)
                for (int i = 1; i <= n1; ++i) {
                    int ka = k + 1, km = k - 1;
                    if (k == n3) ka = 1;
                    double temp = (psif[IX(i, ka)] - 2.0 * psif[IX(i, k)] + psif[IX(i, km)]) * hhpi * rvs[i];
                    if (ann) temp = temp + psif[IX(i, k)] * rvs[i];
                    if (i == 1) {
                        temp = temp + (-2.0 * psif[IX(i, k)] + 5.0 / 3.0 * psif[IX(i + 1, k)]
                                       - 0.2 * psif[IX(i + 2, k)]) * hhri;
                        if (ann) temp = temp + 4.0 / 3.0 * psif[IX(i + 1, k)] * hri * rvi[i];
                    } else if (i == n1) {
                        temp = temp + (-2.0 * psif[IX(i, k)] + 5.0 / 3.0 * psif[IX(i - 1, k)]
                                       - 0.2 * psif[IX(i - 2, k)]) * hhri;
                        if (ann) temp = temp - 4.0 / 3.0 * psif[IX(i - 1, k)] * hri * rvi[i];
                    } else {
                        temp = temp + (psif[IX(i - 1, k)] - 2.0 * psif[IX(i, k)] + psif[IX(i + 1, k)]) * hhri;
                        if (ann) temp = temp + 1.5 * (psif[IX(i + 1, k)] - psif[IX(i - 1, k)]) * hri * rvi[i];
                    }
                    temp = rf[IX(i, k)] - temp;
                    rc[IX(i, k)] = temp;
                    temp = fabs(temp);
                    if (temp > resmax) resmax = temp;
                }
        } else {
            for (int k = 1; k <= n3; ++k) {
                int ka = k + 1, km = k - 1;
                if (k == n3) ka = 1;
                for (int i = 1; i <= n1 - 1; ++i) {
                    ta[i] = pa[i]; tb[i] = pb[i]; tc[i] = pc[i];
                    tr[i] = -(psif[IX(i, ka)] + psif[IX(i, km)]) * hhpi * rvs[i] + rf[IX(i, k)];
                }
                vo_tdma(ta + 1, tb + 1, tc + 1, ts + 1, tr + 1, n1 - 1);
                for (int i = 1; i <= n1 - 1; ++i) {
                    psif[IX(i, k)] = ts[i];
                    if (k == n3) psif[IX(i, 0)] = psif[IX(i, k)];
                }
            }
            resmax = 0.0;
            for (int k = 1; k <= n3; ++k)
                for (int i = 1; i <= n1 - 1; ++i) {
                    int ka = k + 1, km = k - 1;
                    if (k == n3) ka = 1;
                    double temp;
                    if (ann)
                        temp = (psif[IX(i + 1, k)] + psif[IX(i - 1, k)]) * hhri
                             + 1.5 * (psif[IX(i + 1, k)] - psif[IX(i - 1, k)]) * hri * rvi[i]
                             + psif[IX(i, k)] * rvs[i]
                             + (psif[IX(i, ka)] + psif[IX(i, km)]) * hhpi * rvs[i]
                             - psif[IX(i, k)] * (2.0 * hhri + 2.0 * hhpi * rvs[i]);
                    else
                        temp = (psif[IX(i + 1, k)] + psif[IX(i - 1, k)]) * hhri
                             + (psif[IX(i, ka)] + psif[IX(i, km)]) * hhpi * rvs[i]
                             - psif[IX(i, k)] * (2.0 * hhri + 2.0 * hhpi * rvs[i]);
                    temp = rf[IX(i, k)] - temp;
                    rc[IX(i, k)] = temp;
                    temp = fabs(temp);
                    if (temp > resmax) resmax = temp;
                }
        }
    }
    trans1(mg, psif, mg->v[L], n1, n3);
    if (klev == 1 && resmax < tolcon) nconv = 1;
    free(w);
    return nconv;
}

/* corser, rb.f:2603-2718: full weighting of the latest residual (the 3D
 * 27-point formula with the collapsed j index), coarse faces zeroed. */
static void corser(vo_mg *mg, int klev)
{
    const int ld = mg->ld;
    const int Lf = klev - 1, Lc = klev;
    const int n1f = mg->n1[Lf], n3f = mg->n3[Lf];
    const int n1c = mg->n1[Lc], n3c = mg->n3[Lc];
    double *rc = mg->rc, *rf = mg->rf, *psic = mg->psic, *psif = mg->psif;
    for (int k = 0; k <= n3f; ++k)           /* trans3 */
        for (int i = 0; i <= n1f; ++i)
            rf[IX(i, k)] = rc[IX(i, k)];
    for (int k = 1; k <= n3c - 1; ++k)
        for (int i = 1; i <= n1c - 1; ++i) {
            int ii = i + i, kk = k + k;
            int iip = ii + 1, iim = ii - 1, kkp = kk + 1, kkm = kk - 1;
            psic[IX(i, k)] = 0.0;
            psif[IX(i, k)] = 0.0;
            double temp = rf[IX(ii, kk)];
            temp = temp + temp + rf[IX(iip, kk)] + rf[IX(iim, kk)]
                 + rf[IX(ii, kk)] + rf[IX(ii, kk)]
                 + rf[IX(ii, kkp)] + rf[IX(ii, kkm)];
            temp = temp + temp + rf[IX(iip, kk)] + rf[IX(iip, kk)]
                 + rf[IX(iim, kk)] + rf[IX(iim, kk)]
                 + rf[IX(iip, kkp)] + rf[IX(iip, kkm)]
                 + rf[IX(iim, kkp)] + rf[IX(iim, kkm)]
                 + rf[IX(ii, kkp)] + rf[IX(ii, kkm)]
                 + rf[IX(ii, kkp)] + rf[IX(ii, kkm)];
            temp = temp + temp + rf[IX(iip, kkp)] + rf[IX(iip, kkm)]
                 + rf[IX(iip, kkp)] + rf[IX(iip, kkm)]
                 + rf[IX(iim, kkp)] + rf[IX(iim, kkm)]
                 + rf[IX(iim, kkp)] + rf[IX(iim, kkm)];
            temp = temp * 0.015625;
            rc[IX(i, k)] = temp;
        }
    for (int k = 0; k <= n3c; ++k) {
        rc[IX(0, k)] = 0.0; rc[IX(n1c, k)] = 0.0;
        psic[IX(0, k)] = 0.0; psic[IX(n1c, k)] = 0.0;
        psif[IX(0, k)] = 0.0; psif[IX(n1c, k)] = 0.0;
    }
    for (int i = 0; i <= n1c; ++i) {
        if (mg->zero_seam) { rc[IX(i, 0)] = 0.0; rc[IX(i, n3c)] = 0.0; }
        psic[IX(i, 0)] = 0.0; psic[IX(i, n3c)] = 0.0;
        psif[IX(i, 0)] = 0.0; psif[IX(i, n3c)] = 0.0;
    }
    trans1(mg, rc, mg->f[Lc], n1c, n3c);
}

/* finer, rb.f:2721-2803: bilinear prolongation of the coarse level klev,
 * added to the stored fine level klev-1. */
static void finer(vo_mg *mg, int klev)
{
    const int ld = mg->ld;
    const int Lc = klev - 1, Lf = klev - 2;
    const int n1c = mg->n1[Lc], n3c = mg->n3[Lc];
    const int n1f = mg->n1[Lf], n3f = mg->n3[Lf];
    double *psic = mg->psic, *psif = mg->psif, *tempa = mg->tempa;
    trans2(mg, mg->v[Lc], psic, n1c, n3c);
    for (int i = 0; i <= n1f; ++i)
        for (int k = 0; k <= n3f; ++k) {
            int i1 = i / 2, k1 = k / 2, i2 = i1 * 2, k2 = k1 * 2;
            double val;
            if (i2 == i) {
                if (k2 == k) val = psic[IX(i1, k1)];
                else val = (psic[IX(i1, k1)] + psic[IX(i1, k1 + 1)]) * 0.5;
            } else {
                if (k2 == k) val = (psic[IX(i1, k1)] + psic[IX(i1 + 1, k1)]) * 0.5;
                else val = (psic[IX(i1, k1)] + psic[IX(i1 + 1, k1)]
                            + psic[IX(i1, k1 + 1)] + psic[IX(i1 + 1, k1 + 1)]) * 0.25;
            }
            psif[IX(i, k)] = val;
        }
    trans2(mg, mg->v[Lf], tempa, n1f, n3f);
    for (int k = 0; k <= n3f; ++k)
        for (int i = 0; i <= n1f; ++i) {
            psif[IX(i, k)] = psif[IX(i, k)] + tempa[IX(i, k)];
            psic[IX(i, k)] = 0.0;
        }
}

/* poisson2 (rb.f:1796-1927) and poisson3 (rb.f:1930-2059) */
int vo_poisson_ref(vo_mg *mg, int which, double *psi, double *psip,
                   const double *rhs, int nr, int np, double tolcon,
                   int maxit, int *iters)
{
    const int ld = mg->ld;
    const int nlev = mg->nlev;
    const int imax = (which == 2) ? nr + 1 : nr;
    const int nrel1 = 1, nrel2 = 1;
    double *psif = mg->psif, *psic = mg->psic, *rc = mg->rc, *rf = mg->rf;
    for (int l = 0; l < nlev; ++l) {
        size_t sz = (size_t)(mg->n1[l] + 1) * (size_t)(mg->n3[l] + 1);
        memset(mg->v[l], 0, sz * sizeof(double));
        memset(mg->f[l], 0, sz * sizeof(double));
    }
    for (int i = 0; i <= imax; ++i)
        for (int k = 0; k <= np; ++k) {
            psic[IX(i, k)] = 0.0;
            psif[IX(i, k)] = 2.0 * psi[IX(i, k)] - psip[IX(i, k)];
            rc[IX(i, k)] = 0.0;
            rf[IX(i, k)] = rhs[IX(i, k)];
        }
    trans1(mg, rf, mg->f[0], mg->n1[0], mg->n3[0]);
    trans1(mg, psif, mg->v[0], mg->n1[0], mg->n3[0]);
    int nconv = 0;
    for (int it = 1; it <= maxit; ++it) {
        if (iters) *iters = it;
        if (it != 1) {
            for (int is = 0; is <= imax; ++is)
                for (int ks = 0; ks <= np; ++ks) {
                    psif[IX(is, ks)] = psip[IX(is, ks)] + 1.2 * (psif[IX(is, ks)] - psip[IX(is, ks)]);
                    psip[IX(is, ks)] = psif[IX(is, ks)];
                }
            trans1(mg, psif, mg->v[0], mg->n1[0], mg->n3[0]);
        }
        if (it == 1 || nlev == 1) {
            nconv = relax(mg, which, nrel1, 1, tolcon);
            for (int ii = 0; ii <= imax; ++ii)
                for (int kk = 0; kk <= np; ++kk)
                    psip[IX(ii, kk)] = psif[IX(ii, kk)];
        }
        if (nlev > 1) {
            for (int ilev = 1; ilev <= nlev - 1; ++ilev) {
                corser(mg, ilev);
                nconv = relax(mg, which, nrel1, ilev + 1, tolcon);
            }
            for (int ifmg = 0; ifmg <= nlev - 2; ++ifmg) {
                for (int ilev = nlev; ilev >= nlev - ifmg; --ilev) {
                    finer(mg, ilev);
                    nconv = relax(mg, which, nrel2, ilev - 1, tolcon);
                }
                for (int ilev = nlev - ifmg - 1; ilev <= nlev - 1; ++ilev) {
                    corser(mg, ilev);
                    nconv = relax(mg, which, nrel1, ilev + 1, tolcon);
                }
            }
            for (int ilev = nlev; ilev >= 2; --ilev) {
                finer(mg, ilev);
                nconv = relax(mg, which, nrel2, ilev - 1, tolcon);
            }
        }
        if (nconv == 1) {
            for (int ii = 0; ii <= imax; ++ii)
                for (int kk = 0; kk <= np; ++kk)
                    psip[IX(ii, kk)] = psi[IX(ii, kk)];
            int ihi = (which == 2) ? nr : nr - 1;
            for (int ii = 1; ii <= ihi; ++ii)
                for (int kk = 0; kk <= np; ++kk)
                    psi[IX(ii, kk)] = psif[IX(ii, kk)];
            int iwall = (which == 2) ? nr + 1 : nr;
            for (int kk = 0; kk <= np; ++kk) {
                psi[IX(0, kk)] = 0.0;
                psi[IX(iwall, kk)] = 0.0;
            }
            return 0;
        }
    }
    return 1;
}

/* ----------------------------------------------------- exact Poisson */

/* Row coefficients of the r part of the converged operators (the level-1
 * residuals of relax2/relax3): row i couples psi(i-2..i+2) with s[0..4];
 * rv2 is the factor of the phi second difference (1/r^2, 1 in the box). */
static void op_row(int which, int geom, int i, int nr, double dr, double r0,
                   double *s, double *rv2)
{
    const double hhri = 1.0 / (dr * dr), hri = 1.0 / dr;
    for (int o = 0; o < 5; ++o) s[o] = 0.0;
    const int ann = (geom == 1);
    if (which == 2) {
        const double rvi = ann ? 1.0 / (r0 + ((double)i - 0.5) * dr) : 1.0;
        const double rvs = rvi * rvi;
        *rv2 = rvs;
        if (ann) s[2] += rvs;
        if (i == 1) {
            s[2] += -2.0 * hhri; s[3] += 5.0 / 3.0 * hhri; s[4] += -0.2 * hhri;
            if (ann) s[3] += 4.0 / 3.0 * hri * rvi;
        } else if (i == nr) {
            s[2] += -2.0 * hhri; s[1] += 5.0 / 3.0 * hhri; s[0] += -0.2 * hhri;
            if (ann) s[1] -= 4.0 / 3.0 * hri * rvi;
        } else {
            s[1] += hhri; s[2] += -2.0 * hhri; s[3] += hhri;
            if (ann) { s[3] += 1.5 * hri * rvi; s[1] -= 1.5 * hri * rvi; }
        }
    } else {
        const double rvi = ann ? 1.0 / ((double)i * dr + r0) : 1.0;
        const double rvs = rvi * rvi;
        *rv2 = rvs;
        s[1] = hhri; s[2] = -2.0 * hhri; s[3] = hhri;
        if (ann) { s[2] += rvs; s[3] += 1.5 * hri * rvi; s[1] -= 1.5 * hri * rvi; }
        if (i == 1) s[1] = 0.0;
        if (i == nr - 1) s[3] = 0.0;
    }
}

/* Residual of the operator whose discrete solution poisson2/poisson3
 * converge to: relax2's residual (rb.f:2306-2357, ann.f:2301-2348) and
 * relax3's (rb.f:2553-2587, ann.f:2555-2583) at level 1, unknowns
 * k = 1..np (k = 0 == np). */
double vo_residual_g(int which, const double *psi, const double *rhs, double *out,
                     int nr, int np, int ld, double dr, double dp, int geom, double r0)
{
    const double hhpi = 1.0 / (dp * dp);
    double rmax = 0.0;
    const int ihi = (which == 2) ? nr : nr - 1;
    for (int k = 1; k <= np; ++k) {
        int ka = (k == np) ? 1 : k + 1, km = k - 1;
        for (int i = 1; i <= ihi; ++i) {
            double sr[5], rv2;
            op_row(which, geom, i, nr, dr, r0, sr, &rv2);
            double pkm = (km == 0) ? psi[IX(i, np)] : psi[IX(i, km)];
            double t = (psi[IX(i, ka)] - 2.0 * psi[IX(i, k)] + pkm) * hhpi * rv2;
            for (int o = -2; o <= 2; ++o) {
                int ii = i + o;
                if (ii < 1 || ii > ihi || sr[o + 2] == 0.0) continue;
                t += sr[o + 2] * psi[IX(ii, k)];
            }
            double r = rhs[IX(i, k)] - t;
            if (out) out[IX(i, k)] = r;
            if (fabs(r) > rmax) rmax = fabs(r);
        }
    }
    return rmax;
}

double vo_residual(int which, const double *psi, const double *rhs,
                   double *out, int nr, int np, int ld, double dr, double dp)
{
    return vo_residual_g(which, psi, rhs, out, nr, np, ld, dr, dp, 0, 0.0);
}

/* Separable exact solve: the operator is D_r + diag(rv2) (x) D_phi with
 * D_phi the periodic 3-point second difference on np points.  A real DFT in
 * phi diagonalises D_phi (eigenvalue -(2-2cos(2 pi m/np))/dp^2); each mode
 * is a banded (bandwidth 2) system in r solved by elimination without
 * pivoting (row diagonally dominant). */
int vo_poisson_exact_g(int which, double *psi, double *psip, const double *rhs,
                       int nr, int np, int ld, double dr, double dp, int geom, double r0)
{
    const int n = (which == 2) ? nr : nr - 1;
    const double hhpi = 1.0 / (dp * dp);
    const double twopi = 6.283185307179586476925286766559;
    const int nm = np / 2 + 1;
    double *cs = (double *)malloc(sizeof(double) * np);
    double *sn = (double *)malloc(sizeof(double) * np);
    for (int j = 0; j < np; ++j) { cs[j] = cos(twopi * j / np); sn[j] = sin(twopi * j / np); }
    double *fc = (double *)calloc((size_t)nm * (n + 1), sizeof(double));
    double *fsn = (double *)calloc((size_t)nm * (n + 1), sizeof(double));
    for (int m = 0; m < nm; ++m)
        for (int i = 1; i <= n; ++i) {
            long double sc = 0.0L, ss = 0.0L;
            for (int q = 0; q < np; ++q) {
                int j = (int)(((long)m * q) % np);
                double fv = rhs[IX(i, q + 1)];
                sc += (long double)fv * cs[j];
                ss += (long double)fv * sn[j];
            }
            fc[(size_t)m * (n + 1) + i] = (double)sc;
            fsn[(size_t)m * (n + 1) + i] = (double)ss;
        }
    double *bm = (double *)malloc(sizeof(double) * 5 * (n + 3));
    double *x = (double *)malloc(sizeof(double) * (n + 3));
    double *y = (double *)malloc(sizeof(double) * (n + 3));
    for (int m = 0; m < nm; ++m) {
        double lam = -(2.0 - 2.0 * cs[m % np]) * hhpi;
        for (int pass = 0; pass < 2; ++pass) {
            double *rv = pass == 0 ? fc : fsn;
            if (pass == 1 && (m == 0 || 2 * m == np)) continue;
#define BM(i, o) bm[(size_t)(i) * 5 + (o) + 2]
            for (int i = 0; i <= n + 2; ++i) for (int o = -2; o <= 2; ++o) BM(i, o) = 0.0;
            for (int i = 1; i <= n; ++i) {
                double sr[5], rv2;
                op_row(which, geom, i, nr, dr, r0, sr, &rv2);
                for (int o = -2; o <= 2; ++o)
                    if (i + o >= 1 && i + o <= n) BM(i, o) = sr[o + 2];
                BM(i, 0) += lam * rv2;
                y[i] = rv[(size_t)m * (n + 1) + i];
            }
            for (int i = 1; i <= n; ++i) {
                for (int rr = i + 1; rr <= i + 2 && rr <= n; ++rr) {
                    double a = BM(rr, i - rr);
                    if (a == 0.0) continue;
                    double fct = a / BM(i, 0);
                    for (int o = 0; o <= 2 && i + o <= n; ++o)
                        BM(rr, i + o - rr) -= fct * BM(i, o);
                    y[rr] -= fct * y[i];
                }
            }
            for (int i = n; i >= 1; --i) {
                double sacc = y[i];
                for (int o = 1; o <= 2 && i + o <= n; ++o) sacc -= BM(i, o) * x[i + o];
                x[i] = sacc / BM(i, 0);
            }
            for (int i = 1; i <= n; ++i) rv[(size_t)m * (n + 1) + i] = x[i];
#undef BM
        }
    }
    for (int kk = 0; kk <= np; ++kk)
        for (int i = 0; i <= ((which == 2) ? nr + 1 : nr); ++i)
            psip[IX(i, kk)] = psi[IX(i, kk)];
    for (int i = 1; i <= n; ++i)
        for (int q = 0; q < np; ++q) {
            long double sacc = 0.0L;
            for (int m = 0; m < nm; ++m) {
                int j = (int)(((long)m * q) % np);
                double wgt = (m == 0 || 2 * m == np) ? 1.0 : 2.0;
                sacc += (long double)wgt * (fc[(size_t)m * (n + 1) + i] * cs[j]
                                            + fsn[(size_t)m * (n + 1) + i] * sn[j]);
            }
            psi[IX(i, q + 1)] = (double)(sacc / np);
        }
    for (int i = 1; i <= n; ++i) psi[IX(i, 0)] = psi[IX(i, np)];
    int iwall = (which == 2) ? nr + 1 : nr;
    for (int kk = 0; kk <= np; ++kk) { psi[IX(0, kk)] = 0.0; psi[IX(iwall, kk)] = 0.0; }
    free(cs); free(sn); free(fc); free(fsn); free(bm); free(x); free(y);
    return 0;
}

int vo_poisson_exact(int which, double *psi, double *psip, const double *rhs,
                     int nr, int np, int ld, double dr, double dp)
{
    return vo_poisson_exact_g(which, psi, psip, rhs, nr, np, ld, dr, dp, 0, 0.0);
}

/* ------------------------------------------------------------ one step */

int vo_step(const vo_params *p, vo_fields *f, vo_mg *mg, double *vrh,
            double *urh, int first, int mode, double tolcon,
            double *trans, int *nclamp)
{
    int st = 0, it = 0;
    vo_cvorticity(p, f, first);
    vo_temperature(p, f, first);
    vo_heat(p, f->th, &trans[0], &trans[1]);
    vo_rhsv(p, f->Cn, vrh);
    /* rb.f stops after 500 outer iterations (rb.f:1843); ann.f runs up to
     * nvs = 10000 (ann.f:1821) */
    const int maxit = (p->geom == 1) ? 10000 : 500;
    if (mode == 0) st |= vo_poisson_ref(mg, 2, f->vel, f->velp, vrh, p->nr, p->np, tolcon, maxit, &it);
    else st |= vo_poisson_exact_g(2, f->vel, f->velp, vrh, p->nr, p->np, p->ld, p->dr, p->dp, p->geom, p->r0);
    vo_rhsu(p, f->Cn, urh);
    if (mode == 0) st |= vo_poisson_ref(mg, 3, f->uel, f->uelp, urh, p->nr, p->np, tolcon, maxit, &it);
    else st |= vo_poisson_exact_g(3, f->uel, f->uelp, urh, p->nr, p->np, p->ld, p->dr, p->dp, p->geom, p->r0);
    int nc = vo_update(p, f);
    if (nclamp) *nclamp = nc;
    return st;
}

/* ------------------------------------------------------------ seeding */

static unsigned long long splitmix64(unsigned long long *s)
{
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void vo_seed_temperature(const vo_params *p, double *th, double r0,
                         double amp, unsigned long long seed)
{
    const int nr = p->nr, np = p->np, ld = p->ld;
    unsigned long long s = seed;
    for (int k = 0; k < np; ++k)
        for (int i = 1; i <= nr; ++i) {
            double r = r0 + ((double)i - 0.5) * p->dr;
            double u = (double)(splitmix64(&s) >> 11) * (1.0 / 9007199254740992.0);
            th[IX(i, k)] = 2.0 * (r - r0) + amp * (2.0 * u - 1.0);
        }
    for (int i = 1; i <= nr; ++i) th[IX(i, np)] = th[IX(i, 0)];
    for (int k = 0; k <= np; ++k) { th[IX(0, k)] = 0.0; th[IX(nr + 1, k)] = 1.0; }
}
