/* vv3_oracle.h -- CPU oracle (test infrastructure only) for the 3D rotating-
 * annulus time step of rot3dmg.f: a literal C restatement of its explicit
 * routines and its three Poisson residual operators, plus an exact solve of
 * those operators for tiny grids.  Only the tests, the smoke check and the
 * benchmark's CPU leg use it; no product code links it.
 *
 * Layout: X(i,j,k) at X[(k*ldj + j)*ldi + i], i = 0..nr+1, j = 0..nz+1,
 * k = 0..np (k = 0 and k = np are the same azimuth; the reference keeps both).
 */
#ifndef VV3_ORACLE_H
#define VV3_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nr, nz, np, ldi, ldj;
    double dt, dr, dz, dp, Ra, Re, Ptl, Gr, r0, rmax;
    int model;   /* 0: rot3dmg.f; 1: rot3dmghor.f (stationary horizontal
                    annulus: gravity in the r-theta plane through Gr, no
                    Coriolis terms, T = 1 inside / 0 outside, heat flux at
                    both cylinders near mid-axis) */
} v3_params;

typedef struct {
    double *An, *A, *Ap, *Bn, *B, *Bp, *Cn, *C, *Cp;
    double *thn, *th, *thp;
    double *uel, *vel, *wel, *uelp, *velp, *welp;
    double *urh, *vrh, *wrh;
} v3_fields;

void v3_avorticity(const v3_params *P, v3_fields *F);   /* rot3dmg.f:341-468 */
void v3_bvorticity(const v3_params *P, v3_fields *F);   /* rot3dmg.f:477-593 */
void v3_cvorticity(const v3_params *P, v3_fields *F);   /* rot3dmg.f:601-728 */
void v3_temperature(const v3_params *P, v3_fields *F);  /* rot3dmg.f:736-874 */
double v3_heat(const v3_params *P, const v3_fields *F); /* rot3dmg.f:882-905; trans1 */
/* heat of the model: out[0] = trans1, out[1] = trans2 (rot3dmghor.f:913-946,
 * the two cylinders between nz/2-3 and nz/2+3; 0 for rot3dmg.f) */
void v3_heat2(const v3_params *P, const v3_fields *F, double out[2]);
void v3_update(const v3_params *P, v3_fields *F);       /* rot3dmg.f:964-1018 */
/* Kinetic energy of the velocities (no counterpart in the reference; the
 * build's definition): 1/2 dr dz dp sum over k = 0..np-1 of
 *   sum_{j=1..nz, i=1..nr-1}   (r0 + i dr)       uel^2
 * + sum_{j=1..nz, i=1..nr}     (r0 + (i-1/2) dr) vel^2
 * + sum_{j=1..nz-1, i=1..nr}   (r0 + (i-1/2) dr) wel^2,
 * each over the unknowns of its Poisson solve (poisson3, 2, 1). */
double v3_energy(const v3_params *P, const v3_fields *F);
/* Local Nusselt numbers at mid-height (sect360v2mg.f:318-325): for each
 * plane k = 0..np at j = nz/2,
 *   inner[k] = -r0 ln(r0) (-184 T0 + 225 T1 - 50 T2 + 9 T3) / (60 dr),
 *   outer[k] = -rmax ln(r0) (184 T(nr+1) - 225 T(nr) + 50 T(nr-1) - 9 T(nr-2)) / (60 dr),
 * in the reference's operation order (ln(r0) from the C library). */
void v3_local_nu(const v3_params *P, const v3_fields *F, double *inner, double *outer);
/* Azimuthal Fourier-mode volume integrals of the temperature after
 * rot3dmghor.f:954-991 (subroutine aaaaaaaa), n = 1..nmax:
 *   c[n-1] = dp dr dz sum_{k=1..np} sum_{i=0..nr+1} sum_{j=0..nz+1}
 *            f_i f_j (r0 + i dr) T(i,j,k) cos(n k dp),   s[n-1] with sin,
 * f = 1/2 on the rows i = 0, nr+1 and j = 0, nz+1, 1 elsewhere.  The
 * reference routine reads r0 from an uninitialised local (it lacks COMMON
 * /radius/) and carries each plane's running sum, scaled by dr dz, into the
 * next plane; the run's r0 and one sum per plane are used here. */
void v3_modes(const v3_params *P, const v3_fields *F, int nmax, double *c, double *s);
void v3_rhsw(const v3_params *P, v3_fields *F);         /* rot3dmg.f:1026-1048 */
void v3_rhsv(const v3_params *P, v3_fields *F);         /* rot3dmg.f:1066-1093 */
void v3_rhsu(const v3_params *P, v3_fields *F);         /* rot3dmg.f:1111-1134 */

/* residual f - L psi of operator which (1: w, 2: v, 3: u) at the finest
 * level, into res (same layout; zero outside the unknowns); returns the
 * max-norm.  rot3dmg.f:1695-1723 (w), 1927-1965 (v), 2130-2165 (u). */
double v3_residual(const v3_params *P, int which, const double *psi, const double *rhs, double *res);

/* exact solution of L psi = rhs on the unknowns of operator `which`, by a
 * real DFT in phi (long double sums) and a dense LU of each mode's (r, z)
 * system; psi is zero outside the unknowns, psi(.,.,0) = psi(.,.,np).
 * Returns 0, or -1 if a mode's system is singular.  For tiny grids only
 * ((unknowns per plane)^2 doubles of workspace). */
int v3_poisson_exact(const v3_params *P, int which, double *psi, const double *rhs);

#ifdef __cplusplus
}
#endif
#endif
