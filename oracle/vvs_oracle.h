/* vvs_oracle.h -- CPU oracle (test infrastructure only) of the 3D step of
 * sect360v2mg.f: the rotating annulus of rot3dmg.f with a radial barrier,
 * so that phi is wall-bounded (the barrier's two faces are the planes
 * k = 0 and k = np, the "360 degree sector").  Only the tests, the smoke
 * check and the benchmark's CPU leg use it; no product code links it.
 *
 * sect360v2mg.f does not compile here (its include files sect360mg.ina and
 * sect360mg.inb are absent), so each routine is restated from the text; the
 * file:line of each follows its name.  Operation order is the reference's
 * (compiled with -ffp-contract=off).
 *
 * Staggering in phi (sect360v2mg.f:206-223, 964-980): th, u, w and B sit at
 * cells k = 1..np with the planes k = 0 and k = np+1 on the barrier's faces;
 * v, A and C sit at faces k = 0..np, the barrier faces being k = 0 and np.
 * Layout: X(i,j,k) at X[(k*ldj + j)*ldi + i], i = 0..nr+1, j = 0..nz+1,
 * k = 0..np+1 (np+2 planes).  The parameters are rot3dmg.f's (v3_params,
 * model ignored).
 */
#ifndef VVS_ORACLE_H
#define VVS_ORACLE_H
#include "vv3_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

void vs_avorticity(const v3_params *P, v3_fields *F);   /* sect360v2mg.f:343-452 */
void vs_bvorticity(const v3_params *P, v3_fields *F);   /* sect360v2mg.f:460-569 */
void vs_cvorticity(const v3_params *P, v3_fields *F);   /* sect360v2mg.f:577-684 */
void vs_temperature(const v3_params *P, v3_fields *F);  /* sect360v2mg.f:692-831 */
double vs_heat(const v3_params *P, const v3_fields *F); /* sect360v2mg.f:840-881; trans1 */
void vs_update(const v3_params *P, v3_fields *F);       /* sect360v2mg.f:936-996 */
void vs_rhsw(const v3_params *P, v3_fields *F);         /* sect360v2mg.f:999-1031 */
void vs_rhsv(const v3_params *P, v3_fields *F);         /* sect360v2mg.f:1034-1063 */
void vs_rhsu(const v3_params *P, v3_fields *F);         /* sect360v2mg.f:1066-1100 */

/* unknowns of operator which (1: w, 2: v, 3: u): i = 1..i1, j = 1..j1,
 * k = 1..k1 */
void vs_ranges(const v3_params *P, int which, int *i1, int *j1, int *k1);
/* residual f - L psi of operator which at the finest level into res (same
 * layout, zero outside the unknowns; may be NULL); returns the max-norm.
 * sect360v2mg.f:1817-1848 (w), 2034-2065 (v), 2233-2261 (u).  psi is read
 * on the unknowns only: the walls the operator reaches are zero. */
double vs_residual(const v3_params *P, int which, const double *psi, const double *rhs, double *res);
/* exact solution of L x = rhs on the unknowns of `which` by a dense LU of
 * the whole operator (tiny grids only: (i1 j1 k1)^2 doubles); x is written
 * on the unknowns, the rest of x is left alone.  Returns 0, or -1 if the
 * system is singular or memory runs out. */
int vs_poisson_exact(const v3_params *P, int which, double *x, const double *rhs);
/* poisson1/2/3 (sect360v2mg.f:1357-1460, 1462-1557, 1559-1661) with the
 * exact solve in place of the multigrid iteration: psip <- psi on the
 * range the reference copies, psi <- the solution on the unknowns, then the
 * walls and the mirrored ghost values the reference writes. */
int vs_poisson(const v3_params *P, int which, double *psi, double *psip, const double *rhs);

#ifdef __cplusplus
}
#endif
#endif
