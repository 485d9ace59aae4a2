/* vvm_oracle.c -- CPU oracle of the meridional (r,z) plane with swirl on a
 * stretched grid (test infrastructure; see vvm_oracle.h for the composite).
 * Compiled with -ffp-contract=off; the GPU evaluates the same expressions
 * with the same tables and must match bit for bit in the explicit step. */
#include "vvm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define F2(a, i, j) (a)[(size_t)(j) * (size_t)(P->nr + 1) + (size_t)(i)]

/* wfdif.f:59-85: weights C(q,k) of node q for derivative k at xx, stored
 * c[k*n + q]; the same recursion, 0-based */
void vm_fornberg(const double *x, int n, double xx, int m, double *c)
{
    memset(c, 0, (size_t)(m + 1) * n * sizeof(double));
    c[0] = 1.0;
    double c1 = 1.0, c4 = x[0] - xx;
    for (int i = 1; i < n; ++i) {
        const int mn = i < m ? i : m;
        double c2 = 1.0;
        const double c5 = c4;
        c4 = x[i] - xx;
        for (int j = 0; j < i; ++j) {
            const double c3 = x[i] - x[j];
            c2 = c2 * c3;
            if (j == i - 1) {
                for (int k = mn; k >= 1; --k) c[k * n + i] = c1 * (k * c[(k - 1) * n + i - 1] - c5 * c[k * n + i - 1]) / c2;
                c[i] = -c1 * c5 * c[i - 1] / c2;
            }
            for (int k = mn; k >= 1; --k) c[k * n + j] = (c4 * c[k * n + j] - k * c[(k - 1) * n + j]) / c3;
            c[j] = c4 * c[j] / c3;
        }
        c1 = c2;
    }
}

void vm_nodes(int n, double a, double b, double beta, double *x)
{
    for (int q = 0; q <= n; ++q) {
        const double xi = (double)q / (double)n;
        const double s = (beta == 0.0) ? xi : 0.5 * (1.0 + tanh(beta * (2.0 * xi - 1.0)) / tanh(beta));
        x[q] = a + (b - a) * s;
    }
    x[0] = a;
    x[n] = b;
}

/* Compact row at interior node i of nodes x: alpha f'(i-1) + f'(i) + beta
 * f'(i+1) = a f(i-1) + b f(i) + c f(i+1), fourth order on the stretched
 * spacing: the five Taylor conditions of orders 0..4, solved by Gaussian
 * elimination with partial pivoting.  out = (alpha, beta, a, b, c). */
static void compact_row(const double *x, int i, double out[5])
{
    const double hm = x[i] - x[i - 1], hp = x[i + 1] - x[i];
    /* unknowns (alpha, beta, a, b, c); rows: orders 0..4 */
    double M[5][6] = {
        {0.0, 0.0, 1.0, 1.0, 1.0, 0.0},
        {-1.0, -1.0, -hm, 0.0, hp, 1.0},
        {hm, -hp, hm * hm / 2.0, 0.0, hp * hp / 2.0, 0.0},
        {-hm * hm / 2.0, -hp * hp / 2.0, -hm * hm * hm / 6.0, 0.0, hp * hp * hp / 6.0, 0.0},
        {hm * hm * hm / 6.0, -hp * hp * hp / 6.0, hm * hm * hm * hm / 24.0, 0.0, hp * hp * hp * hp / 24.0, 0.0}};
    for (int col = 0; col < 5; ++col) {
        int piv = col;
        for (int r = col + 1; r < 5; ++r)
            if (fabs(M[r][col]) > fabs(M[piv][col])) piv = r;
        if (piv != col)
            for (int q = 0; q < 6; ++q) {
                const double t = M[col][q];
                M[col][q] = M[piv][q];
                M[piv][q] = t;
            }
        for (int r = col + 1; r < 5; ++r) {
            const double f = M[r][col] / M[col][col];
            for (int q = col; q < 6; ++q) M[r][q] = M[r][q] - f * M[col][q];
        }
    }
    for (int r = 4; r >= 0; --r) {
        double s = M[r][5];
        for (int q = r + 1; q < 5; ++q) s = s - M[r][q] * out[q];
        out[r] = s / M[r][r];
    }
}

/* the tables of a line of n+1 nodes: rows[5*(n+1)] (alpha, beta, a, b, c at
 * node i), end weights e0[5], en[5] (f'(0) from nodes 0..4, f'(n) from
 * nodes n-4..n), Thomas factors fact[i], bp[i] of the rows 1..n-1 (tdma's
 * elimination, velvor2dperiodicrbconv.f:3050-3054) */
typedef struct {
    int n;
    double *rows, *fact, *bp, e0[5], en[5];
} line_tab;

static void line_tables(int n, const double *x, line_tab *T)
{
    T->n = n;
    T->rows = (double *)calloc((size_t)5 * (n + 1), sizeof(double));
    T->fact = (double *)calloc((size_t)(n + 1), sizeof(double));
    T->bp = (double *)calloc((size_t)(n + 1), sizeof(double));
    for (int i = 1; i <= n - 1; ++i) compact_row(x, i, T->rows + 5 * i);
    double c[2 * 5];
    vm_fornberg(x, 5, x[0], 1, c);
    for (int q = 0; q < 5; ++q) T->e0[q] = c[5 + q];
    vm_fornberg(x + n - 4, 5, x[n], 1, c);
    for (int q = 0; q < 5; ++q) T->en[q] = c[5 + q];
    /* rows alpha_i, 1, beta_i; elimination as tdma */
    T->bp[1] = 1.0;
    for (int i = 2; i <= n - 1; ++i) {
        T->fact[i] = T->rows[5 * i + 0] / T->bp[i - 1];
        T->bp[i] = 1.0 - T->fact[i] * T->rows[5 * (i - 1) + 1];
    }
}

static void line_free(line_tab *T)
{
    free(T->rows);
    free(T->fact);
    free(T->bp);
}

/* one line: f[0..n] with stride s; out rows 1..n-1 (stride s) */
static void line_solve(const line_tab *T, const double *f, size_t s, double *out)
{
    const int n = T->n;
    const double *R = T->rows;
    const double d0 = T->e0[0] * f[0] + T->e0[1] * f[s] + T->e0[2] * f[2 * s] + T->e0[3] * f[3 * s] +
                      T->e0[4] * f[4 * s];
    const double dn = T->en[0] * f[(size_t)(n - 4) * s] + T->en[1] * f[(size_t)(n - 3) * s] +
                      T->en[2] * f[(size_t)(n - 2) * s] + T->en[3] * f[(size_t)(n - 1) * s] +
                      T->en[4] * f[(size_t)n * s];
    double rp = 0.0;
    for (int i = 1; i <= n - 1; ++i) {
        const double *q = R + 5 * i;
        double r = q[2] * f[(size_t)(i - 1) * s] + q[3] * f[(size_t)i * s] + q[4] * f[(size_t)(i + 1) * s];
        if (i == 1) r = r - q[0] * d0;
        if (i == n - 1) r = r - q[1] * dn;
        rp = (i == 1) ? r : r - T->fact[i] * rp;
        out[(size_t)i * s] = rp;
    }
    double y = rp / T->bp[n - 1];
    out[(size_t)(n - 1) * s] = y;
    for (int i = n - 2; i >= 1; --i) {
        y = (out[(size_t)i * s] - R[5 * i + 1] * y) / T->bp[i];
        out[(size_t)i * s] = y;
    }
}

void vm_compact_line(int n, const double *x, const double *f, double *out)
{
    line_tab T;
    line_tables(n, x, &T);
    line_solve(&T, f, 1, out);
    line_free(&T);
}

void vm_derivatives(const vm_params *P, const vm_fields *F, double *const d[6])
{
    const int nr = P->nr, nz = P->nz, ni = nr + 1;
    line_tab Tr, Tz;
    line_tables(nr, P->r, &Tr);
    line_tables(nz, P->z, &Tz);
    const double *src[3] = {F->w, F->v, F->th};
    for (int q = 0; q < 3; ++q) {
        for (int j = 1; j <= nz - 1; ++j) line_solve(&Tr, src[q] + (size_t)j * ni, 1, d[2 * q] + (size_t)j * ni);
        for (int i = 1; i <= nr - 1; ++i) line_solve(&Tz, src[q] + i, (size_t)ni, d[2 * q + 1] + i);
    }
    line_free(&Tr);
    line_free(&Tz);
}

/* three-point weights at interior node i of x: first derivative c1[3],
 * second c2[3] (nodes i-1, i, i+1) */
static void w3(const double *x, int i, double c1[3], double c2[3])
{
    double c[3 * 3];
    vm_fornberg(x + i - 1, 3, x[i], 2, c);
    for (int q = 0; q < 3; ++q) {
        c1[q] = c[3 + q];
        c2[q] = c[6 + q];
    }
}

/* the DuFort-Frankel step of rotann2.f:281-306 (vorticity), 330-341
 * (swirl), 370-385 (temperature) in velocity form: a field f advances by
 *   f_new = (AA f_p + S) / A,  A = 1/(2 P dt) - D/2,  AA = 1/(2 P dt) + D/2,
 * where D is the centre coefficient of the diffusion operator (both second
 * derivatives and, for w and v, the -1/r^2 term), S the rest of it
 * (neighbours and the first-derivative term) plus the sources, minus the
 * advection divided by P (P = Pr for w and v, 1 for th). */
void vm_explicit(const vm_params *P, vm_fields *F, double *const d[6])
{
    const int nr = P->nr, nz = P->nz;
    const double dt = P->dt, Pr = P->Pr, Ra = P->Ra, Re = P->Re;
    const double RRP = (Re != 0.0) ? Ra / Re / Pr : 0.0;   /* no rotation: no coupling */
    const double *w = F->w, *v = F->v, *th = F->th, *u = F->u, *ww = F->ww;
    for (int j = 1; j <= nz - 1; ++j) {
        double z1[3], z2[3];
        w3(P->z, j, z1, z2);
        for (int i = 1; i <= nr - 1; ++i) {
            double r1[3], r2[3];
            w3(P->r, i, r1, r2);
            const double r = P->r[i];
            const double uu = F2(u, i, j), wv = F2(ww, i, j);
            /* neighbour part of the diffusion operator and its centre */
#define OFF(f) (r2[0] * F2(f, i - 1, j) + r2[2] * F2(f, i + 1, j) + z2[0] * F2(f, i, j - 1) + \
                z2[2] * F2(f, i, j + 1) + (r1[0] * F2(f, i - 1, j) + r1[1] * F2(f, i, j) + r1[2] * F2(f, i + 1, j)) / r)
            const double Dc = r2[1] + z2[1];
            const double vz = z1[0] * F2(v, i, j - 1) + z1[1] * F2(v, i, j) + z1[2] * F2(v, i, j + 1);
            const double thz = z1[0] * F2(th, i, j - 1) + z1[1] * F2(th, i, j) + z1[2] * F2(th, i, j + 1);
            const double vv = F2(v, i, j), tt = F2(th, i, j);
            {   /* azimuthal vorticity, rotann2.f:281-306 */
                const double D = Dc - 1.0 / (r * r);
                const double A = 1.0 / (2.0 * Pr * dt) - D / 2.0, AA = 1.0 / (2.0 * Pr * dt) + D / 2.0;
                const double adv = (uu * F2(d[0], i, j) + wv * F2(d[1], i, j) - uu * F2(w, i, j) / r) / Pr;
                const double src = 2.0 / Pr * vv * vz / r + 2.0 * Re * vz - Ra * r * thz -
                                   2.0 * RRP * (vv * thz + tt * vz);
                F2(F->wn, i, j) = (AA * F2(F->wp, i, j) + OFF(w) + src - adv) / A;
            }
            {   /* swirl, rotann2.f:330-341 */
                const double D = Dc - 1.0 / (r * r);
                const double A = 1.0 / (2.0 * Pr * dt) - D / 2.0, AA = 1.0 / (2.0 * Pr * dt) + D / 2.0;
                const double adv = (uu * F2(d[2], i, j) + wv * F2(d[3], i, j) + uu * vv / r) / Pr;
                const double src = -2.0 * Re * uu + 2.0 * RRP * tt * uu;
                F2(F->vn, i, j) = (AA * F2(F->vp, i, j) + OFF(v) + src - adv) / A;
            }
            {   /* temperature, rotann2.f:370-385 */
                const double A = 1.0 / (2.0 * dt) - Dc / 2.0, AA = 1.0 / (2.0 * dt) + Dc / 2.0;
                const double adv = uu * F2(d[4], i, j) + wv * F2(d[5], i, j);
                F2(F->thn, i, j) = (AA * F2(F->thp, i, j) + OFF(th) - adv) / A;
            }
#undef OFF
        }
    }
    /* the solves see the current wall vorticity (the new one follows from
     * the velocities they return) */
    for (int i = 0; i <= nr; ++i) {
        F2(F->wn, i, 0) = F2(F->w, i, 0);
        F2(F->wn, i, nz) = F2(F->w, i, nz);
    }
    for (int j = 0; j <= nz; ++j) {
        F2(F->wn, 0, j) = F2(F->w, 0, j);
        F2(F->wn, nr, j) = F2(F->w, nr, j);
    }
}

void vm_rhs(const vm_params *P, vm_fields *F)
{
    const int nr = P->nr, nz = P->nz;
    const double *w = F->wn;
    for (int j = 1; j <= nz - 1; ++j) {
        double z1[3], z2[3];
        w3(P->z, j, z1, z2);
        for (int i = 1; i <= nr - 1; ++i) {
            double r1[3], r2[3];
            w3(P->r, i, r1, r2);
            const double wz = z1[0] * F2(w, i, j - 1) + z1[1] * F2(w, i, j) + z1[2] * F2(w, i, j + 1);
            const double wr = r1[0] * F2(w, i - 1, j) + r1[1] * F2(w, i, j) + r1[2] * F2(w, i + 1, j);
            F2(F->fu, i, j) = wz;
            F2(F->fw, i, j) = -wr - F2(w, i, j) / P->r[i];
        }
    }
}

void vm_apply(const vm_params *P, int which, const double *psi, double *out)
{
    const int nr = P->nr, nz = P->nz;
    for (int j = 1; j <= nz - 1; ++j) {
        double z1[3], z2[3];
        w3(P->z, j, z1, z2);
        for (int i = 1; i <= nr - 1; ++i) {
            double r1[3], r2[3];
            w3(P->r, i, r1, r2);
            const double r = P->r[i];
            const double fm = (i > 1) ? F2(psi, i - 1, j) : 0.0, fp = (i < nr - 1) ? F2(psi, i + 1, j) : 0.0;
            const double gm = (j > 1) ? F2(psi, i, j - 1) : 0.0, gp = (j < nz - 1) ? F2(psi, i, j + 1) : 0.0;
            const double f0 = F2(psi, i, j);
            double L = r2[0] * fm + r2[1] * f0 + r2[2] * fp + (r1[0] * fm + r1[1] * f0 + r1[2] * fp) / r +
                       z2[0] * gm + z2[1] * f0 + z2[2] * gp;
            if (which == 1) L = L - f0 / (r * r);
            F2(out, i, j) = L;
        }
    }
}

void vm_update(const vm_params *P, vm_fields *F)
{
    const int nr = P->nr, nz = P->nz, ni = nr + 1;
    double c[3 * 3];
    /* wall vorticity: omega = -w_r at the cylinders, u_z at the lids (the
     * velocities vanish on the walls), three-point one-sided weights */
    vm_fornberg(P->r, 3, P->r[0], 1, c);
    const double a0 = c[4], a1 = c[5];
    vm_fornberg(P->r + nr - 2, 3, P->r[nr], 1, c);
    const double b1 = c[3], b0 = c[4];   /* weights of nodes nr-2, nr-1 */
    for (int j = 1; j <= nz - 1; ++j) {
        F2(F->wn, 0, j) = -(a0 * F2(F->ww, 1, j) + a1 * F2(F->ww, 2, j));
        F2(F->wn, nr, j) = -(b1 * F2(F->ww, nr - 2, j) + b0 * F2(F->ww, nr - 1, j));
    }
    vm_fornberg(P->z, 3, P->z[0], 1, c);
    const double e0 = c[4], e1 = c[5];
    vm_fornberg(P->z + nz - 2, 3, P->z[nz], 1, c);
    const double g1 = c[3], g0 = c[4];
    for (int i = 0; i <= nr; ++i) {
        F2(F->wn, i, 0) = (i == 0 || i == nr) ? 0.0 : e0 * F2(F->u, i, 1) + e1 * F2(F->u, i, 2);
        F2(F->wn, i, nz) = (i == 0 || i == nr) ? 0.0 : g1 * F2(F->u, i, nz - 2) + g0 * F2(F->u, i, nz - 1);
    }
    /* swirl zero on the walls; th = 0 / 1 at the cylinders (rotann2.f:
     * 399-404); insulated lids: the one-sided three-point derivative of th
     * vanishes */
    vm_fornberg(P->z, 3, P->z[0], 1, c);
    const double t0 = c[3], t1 = c[4], t2 = c[5];
    vm_fornberg(P->z + nz - 2, 3, P->z[nz], 1, c);
    const double s2 = c[3], s1 = c[4], s0 = c[5];
    for (int i = 0; i <= nr; ++i) {
        F2(F->vn, i, 0) = 0.0;
        F2(F->vn, i, nz) = 0.0;
    }
    for (int j = 0; j <= nz; ++j) {
        F2(F->vn, 0, j) = 0.0;
        F2(F->vn, nr, j) = 0.0;
    }
    for (int i = 1; i <= nr - 1; ++i) {
        F2(F->thn, i, 0) = -(t1 * F2(F->thn, i, 1) + t2 * F2(F->thn, i, 2)) / t0;
        F2(F->thn, i, nz) = -(s1 * F2(F->thn, i, nz - 1) + s2 * F2(F->thn, i, nz - 2)) / s0;
    }
    for (int j = 0; j <= nz; ++j) {
        F2(F->thn, 0, j) = 0.0;
        F2(F->thn, nr, j) = 1.0;
    }
    const size_t n = (size_t)ni * (size_t)(nz + 1);
    for (size_t q = 0; q < n; ++q) {
        F->wp[q] = F->w[q];
        F->w[q] = F->wn[q];
        F->vp[q] = F->v[q];
        F->v[q] = F->vn[q];
        F->thp[q] = F->th[q];
        F->th[q] = F->thn[q];
    }
}

void vm_heat(const vm_params *P, const vm_fields *F, double out[2])
{
    const int nr = P->nr, nz = P->nz;
    double c[3 * 3];
    vm_fornberg(P->r, 3, P->r[0], 1, c);
    const double a0 = c[3], a1 = c[4], a2 = c[5];
    vm_fornberg(P->r + nr - 2, 3, P->r[nr], 1, c);
    const double b2 = c[3], b1 = c[4], b0 = c[5];
    double s0 = 0.0, s1 = 0.0;
    for (int j = 0; j <= nz; ++j) {
        const double wz = 0.5 * ((j > 0 ? P->z[j] - P->z[j - 1] : 0.0) + (j < nz ? P->z[j + 1] - P->z[j] : 0.0));
        s0 = s0 + wz * (a0 * F2(F->th, 0, j) + a1 * F2(F->th, 1, j) + a2 * F2(F->th, 2, j));
        s1 = s1 + wz * (b2 * F2(F->th, nr - 2, j) + b1 * F2(F->th, nr - 1, j) + b0 * F2(F->th, nr, j));
    }
    out[0] = s0;
    out[1] = s1;
}
