/* vvt_oracle.h -- CPU oracle (test infrastructure only) of the 3D
 * horizontal-convection tank step.  Only the tests, the smoke check and the
 * benchmark's CPU leg use it; no product code links it.
 *
 * The tank has no single reference program: it composes rot3dmg.f's 3D
 * velocity-vorticity step (its staggering, DuFort-Frankel form, stretching
 * terms, wall-vorticity formulas, right-hand sides and residual operators,
 * rot3dmg.f:341-1134, 1695-2165) in Cartesian form (no metric terms, no
 * centrifugal buoyancy), with the compact-scheme advection derivatives of the
 * 2D codes along all three axes (velvor2dperiodicrbconv.f:607-628 and
 * 1080-1102) and the tank boundaries of velvor2dhorconvtank.f (insulated
 * walls, the linear surface temperature, stress-free side walls:
 * velvor2dhorconvtank.f:1720-1745; the wall heat fluxes of its heat routine,
 * 1574-1622).
 *
 * Axes: i = x ("r" of rot3dmg.f, walls at both ends), j = z (vertical:
 * gravity along -z through Gr, rotation about z through Re, the heated
 * surface at j = nz+1), k = y (the long horizontal axis, side walls).
 * Staggering as rot3dmg.f: th, u, w, B sit at y-cells k = 1..np with ghost
 * planes k = 0 and np+1; v, A, C at y-faces k = 0..np with the walls at
 * k = 0 and k = np.
 * Layout: X(i,j,k) at X[(k*ldj + j)*ldi + i], i = 0..nr+1, j = 0..nz+1,
 * k = 0..np+1.
 */
#ifndef VVT_ORACLE_H
#define VVT_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nr, nz, np, ldi, ldj;
    double dt, dr, dz, dp, Re, Ptl, Gr;
} vt_params;

typedef struct {
    double *An, *A, *Ap, *Bn, *B, *Bp, *Cn, *C, *Cp;
    double *thn, *th, *thp;
    double *uel, *vel, *wel, *uelp, *velp, *welp;
    double *urh, *vrh, *wrh;
} vt_fields;

/* the twelve advection derivatives: d[3*f + a] for field f = A, B, C, th
 * and axis a = x, y, z, each at the field's own points */
typedef struct {
    double *d[12];
} vt_derivs;

/* Compact-scheme first derivative of one line (1-4-1 rows, tdma of
 * velvor2dperiodicrbconv.f:3046-3064).  f holds the line's nodes 0..m+1,
 * out receives rows 1..m (out[1..m]).
 * type 0: the nodes are equally spaced h apart (node 0 and m+1 are wall or
 *         ghost values one spacing away), end rows as the Cr rows of
 *         velvor2dperiodicrbconv.f:613-620;
 * type 1: nodes 0 and m+1 are wall values half a spacing away, end rows as
 *         its thr rows (1085-1097, with their truncated decimals). */
void vt_compact_line(int type, int m, double h, const double *f, double *out);
/* the elimination tables of a line (fact, bp, c; 1-based, size m+1) */
void vt_compact_tables(int type, int m, double *fact, double *bp, double *c);

void vt_derivatives(const vt_params *P, const vt_fields *F, vt_derivs *D);
void vt_avorticity(const vt_params *P, vt_fields *F, const vt_derivs *D);
void vt_bvorticity(const vt_params *P, vt_fields *F, const vt_derivs *D);
void vt_cvorticity(const vt_params *P, vt_fields *F, const vt_derivs *D);
void vt_temperature(const vt_params *P, vt_fields *F, const vt_derivs *D);
/* heat(trans1, trans2) of velvor2dhorconvtank.f:1574-1622 on the tank:
 * the fluxes through the bottom (trans1) and the surface (trans2) over the
 * warm half k = np/2 .. np-1, integrated over x */
void vt_heat(const vt_params *P, const vt_fields *F, double out[2]);
void vt_rhsw(const vt_params *P, vt_fields *F);
void vt_rhsv(const vt_params *P, vt_fields *F);
void vt_rhsu(const vt_params *P, vt_fields *F);
/* wall vorticity, temperature boundaries, ghost planes; then the time
 * levels rotate (Ap <- A <- An, ... as rot3dmg.f:1003-1015) */
void vt_update(const vt_params *P, vt_fields *F);

/* unknowns of operator which (1: w, 2: v, 3: u): i = 1..i1, j = 1..j1,
 * k = 1..k1 */
void vt_ranges(const vt_params *P, int which, int *i1, int *j1, int *k1);
/* L psi on the unknowns of `which` (psi outside the unknowns is ignored:
 * walls are zero, the y ghosts of u and w mirror the first and last cells) */
void vt_apply(const vt_params *P, int which, const double *psi, double *out);
/* residual rhs - L psi on the unknowns (zero elsewhere); returns its
 * max-norm */
double vt_residual(const vt_params *P, int which, const double *psi, const double *rhs, double *res);
/* output conventions of a solve: walls zero, y ghosts of u and w copied from
 * the first and last cells */
void vt_psi_bounds(const vt_params *P, int which, double *psi);

/* kinetic energy 1/2 dr dz dp (sum u^2 + v^2 + w^2) over the unknowns */
double vt_energy(const vt_params *P, const vt_fields *F);

#ifdef __cplusplus
}
#endif
#endif
