/*
 * vv_oracle.h -- CPU oracle for the velocity-vorticity time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * program monashubuntudir_velvor2dperiodicrbconvdir/velvor2dperiodicrbconv.f
 * (abbreviated rb.f below).  It exists so that the HIP library can be checked
 * against it; nothing in the product links it.  Only tests/, the smoke entry
 * point and the CPU-baseline leg of bench.py load it.
 *
 * Array convention (same as the reference's column-major COMMON arrays with a
 * leading dimension chosen by the caller): a 2D field F(i,k), i = 0..nr+1
 * (wall-bounded direction "r"), k = 0..np (periodic direction "phi",
 * k = 0 and k = np are the same physical column) lives at F[k*ld + i].
 *
 * Every routine follows the reference's evaluation order literally so that
 * with -ffp-contract=off the results are bit-identical to the flang build of
 * the reference (checked in tests/test_oracle_vs_reference.py).
 */
#ifndef VV_ORACLE_H
#define VV_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vo_params {
    int nr, np, ld;
    double dt, dr, dp;     /* dp = box width / np (rb.f:130-133)          */
    double Ra, Pr;
    /* geometry: 0 = Cartesian periodic box (rb.f, r == 1),
     *           1 = rotating annulus plane (velvor2dnewsch3.4.f, "ann.f"):
     *               r = r0 + i*dr at vorticity rows, cylindrical terms,
     *               centrifugal buoyancy with Re, Robert 0.05 */
    int geom;
    double r0, Re, robert;
} vo_params;

/* State of one run: the reference's COMMON /sol1/ /sol2/ /sol3/ (rb.f:69-76).
 * A, B, An, Bn, Ap are scratch loaned by the reference to the compact
 * derivatives (rb.f:625-626, 646-647, 670-671). */
typedef struct vo_fields {
    double *C, *Cp, *Cn;
    double *th, *thp, *thn;
    double *uel, *vel, *wel, *uelp, *velp;
    double *A, *B, *An, *Bn, *Ap;
} vo_fields;

/* Multigrid workspace: rb.f COMMON /backup/ /arrays/ /step/ /ints/
 * (rb.f:1736-1739).  The flat per-level store v,f of the reference is kept
 * as one array per level; the shared 2D scratch arrays are kept shared. */
#define VO_MAXLEV 8
typedef struct vo_mg {
    int nlev, ld, ncols;
    int n1[VO_MAXLEV], n3[VO_MAXLEV];
    double hr[VO_MAXLEV], hp[VO_MAXLEV], hhr[VO_MAXLEV], hhp[VO_MAXLEV];
    double *v[VO_MAXLEV], *f[VO_MAXLEV];
    double *psif, *psic, *rc, *rf, *tempa;
    int nlev_cap;          /* 4 in rb.f (rb.f:1766), 5 in ann.f (ann.f:1778) */
    int geom;              /* 0 box, 1 annulus */
    double r0;
    int nsweep;            /* sweeps per relax: nrel (rb.f) or 3 (ann.f:2140) */
    int zero_seam;         /* corser zeroes rc at k=0,n3 (rb.f:2688-2689) */
} vo_mg;

/* banded solvers, 1-based semantics on 0-based arrays of length n */
void vo_tdma(double *a, double *b, double *c, double *y, double *r, int n);
void vo_ptdma(double *a, double *b, double *c, double *y, double *r, int n);
void vo_pdma(double *a, double *b, double *c, double *d, double *e,
             double *r, double *x, int n);

/* explicit physics (first != 0 selects the forward-Euler first-step form) */
void vo_cvorticity(const vo_params *p, vo_fields *f, int first);
void vo_temperature(const vo_params *p, vo_fields *f, int first);
void vo_heat(const vo_params *p, const double *th, double *trans1, double *trans2);
/* Kinetic energy of the velocities (no counterpart in the reference; the
 * build's definition): 1/2 dr dp sum over k = 0..np-1 of
 *   sum_{i=1..nr} w_T(i) vel(i,k)^2 + sum_{i=1..nr-1} w_C(i) uel(i,k)^2,
 * the unknowns of poisson2 / poisson3, with w = 1 in the box and the row
 * radius in the annulus (w_T = r0 + (i-1/2) dr, w_C = r0 + i dr). */
double vo_energy(const vo_params *p, const double *uel, const double *vel);
void vo_rhsv(const vo_params *p, const double *Cn, double *vrh);
void vo_rhsu(const vo_params *p, const double *Cn, double *urh);
int  vo_update(const vo_params *p, vo_fields *f);

/* Poisson */
int  vo_mg_init(vo_mg *mg, int nr, int np, double dr, double dp, int nlev_cap);
/* geometry of the Poisson operators (set after vo_mg_init) */
void vo_mg_set_geom(vo_mg *mg, int geom, double r0, int nsweep, int zero_seam);
void vo_mg_free(vo_mg *mg);
/* literal FMG/line-GS of the reference; which = 2 (v, cell rows 1..nr) or
 * 3 (u, face rows 1..nr-1).  tolcon is the absolute max-norm stop
 * (2e-6 in the reference).  Returns 0 when converged, 1 when maxit outer
 * iterations passed (the reference executes STOP after 500). */
int  vo_poisson_ref(vo_mg *mg, int which, double *psi, double *psip,
                    const double *rhs, int nr, int np, double tolcon,
                    int maxit, int *iters);
/* exact discrete solve of the same operator (separable: real DFT in phi,
 * banded elimination in r); same in/out conventions as vo_poisson_ref. */
int  vo_poisson_exact(int which, double *psi, double *psip, const double *rhs,
                      int nr, int np, int ld, double dr, double dp);
/* residual of the converged operator: out = rhs - L psi on the unknowns,
 * returns max |out| */
double vo_residual(int which, const double *psi, const double *rhs,
                   double *out, int nr, int np, int ld, double dr, double dp);
/* the same two routines for either geometry (r0 used when geom == 1) */
int  vo_poisson_exact_g(int which, double *psi, double *psip, const double *rhs,
                        int nr, int np, int ld, double dr, double dp, int geom, double r0);
double vo_residual_g(int which, const double *psi, const double *rhs, double *out,
                     int nr, int np, int ld, double dr, double dp, int geom, double r0);

/* one full time step of the reference time loop (rb.f:341-382):
 * mode 0 = literal reference Poisson (tolcon absolute),
 * mode 1 = exact Poisson.  vrh/urh are caller-provided scratch fields.
 * outputs: trans[2] (heat), clamp count in *nclamp. */
int  vo_step(const vo_params *p, vo_fields *f, vo_mg *mg, double *vrh,
             double *urh, int first, int mode, double tolcon,
             double *trans, int *nclamp);

/* seeded initial condition of the periodic Rayleigh-Benard box:
 * T = 2(r-r0) on cells + U(-amp,amp) from splitmix64(seed) drawn for
 * k = 0..np-1 (outer) and i = 1..nr (inner); k = np copies k = 0;
 * walls T(0)=0, T(nr+1)=1.  Fills th only. */
void vo_seed_temperature(const vo_params *p, double *th, double r0,
                         double amp, unsigned long long seed);

#ifdef __cplusplus
}
#endif
#endif
