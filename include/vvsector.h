/* vvsector.h -- C ABI of the MI355X 3D step of sect360v2mg.f: the rotating
 * annulus of rot3dmg.f with a radial barrier, so that phi is bounded by two
 * walls (the barrier's faces).  FP64 throughout; status codes and error text
 * are shared with vvmi.h (VV_OK, VV_EINVAL, ..., vv_last_error()).
 *
 * The run takes rot3dmg.f's input deck (vv3_config of vvmi3.h; its model
 * field is ignored) and its field ids (enum vv3_field).  Fields are
 * (np+2) x (nz+2) x (nr+2) doubles, i fastest, then j, then k:
 * X(i,j,k) at host[(k*ldj + j)*ldi + i], i = 0..nr+1, j = 0..nz+1,
 * k = 0..np+1, as the reference's arrays.  th, u, w and B sit at the cells
 * k = 1..np with the planes k = 0 and np+1 on the barrier's faces; v, A and
 * C at the faces k = 0..np, the barrier's faces being k = 0 and k = np
 * (sect360v2mg.f:206-223, 964-980).
 *
 * Every call on a run holds the run's lock.
 */
#ifndef VVSECTOR_H
#define VVSECTOR_H
#include "vvmi3.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vvs_run vvs_run;

/* largest np (the phi transforms are dense tables of np^2 doubles) and nz
 * (a plane's heat-flux terms are summed in LDS) */
#define VVS_NP_MAX 2048
#define VVS_NZ_MAX 4096

/* Allocates the fields (zero) and the Poisson solver on `device` (the
 * reference builds its multigrid levels in poissoninit,
 * sect360v2mg.f:1302).  Needs nr >= 4, 4 <= nz <= VVS_NZ_MAX and
 * 4 <= np <= VVS_NP_MAX (VV_EINVAL otherwise). */
int vvs_create(const vv3_config *cfg, int device, vvs_run **out);
int vvs_destroy(vvs_run *run);

/* whole-field copies (ldi >= nr+2, ldj >= nz+2; 0 selects nr+2 / nz+2) */
int vvs_put(vvs_run *run, int field, const double *host, int ldi, int ldj);
int vvs_get(vvs_run *run, int field, double *host, int ldi, int ldj);

/* Avorticity, Bvorticity, Cvorticity (sect360v2mg.f:343, 460, 577): An on
 * k = 1..np-1, Bn on k = 1..np with the barrier closures of its rows k = 1
 * and k = np, Cn on k = 1..np-1.  which = 1 / 2 / 3 runs one routine, 0 all
 * three in one launch. */
int vvs_vorticity(vvs_run *run, int which);
/* temperature (sect360v2mg.f:692): thn on i = 1..nr, j = 0..nz+1,
 * k = 0..np+1, the barrier's faces insulated.  As the reference's loop
 * does, the row k = np advects with the phi derivative of the row k = np-1
 * and its one-sided phi derivative in place of the z derivative
 * (sect360v2mg.f:787 assigns it to thz). */
int vvs_temperature(vvs_run *run);
/* heat (sect360v2mg.f:840): the inner-wall flux trans1 over every plane
 * k = 0..np+1 */
int vvs_heat(vvs_run *run, double *trans1);
/* Local Nusselt numbers at mid-height of both cylinders, one per plane
 * k = 0..np (inner and outer hold np+1 values), as the reference writes
 * them with its fields (sect360v2mg.f:318-325), at j = nz/2:
 *   inner[k] = -r0 ln(r0) (-184 T0 + 225 T1 - 50 T2 + 9 T3) / (60 dr),
 *   outer[k] = -rmax ln(r0) (184 T(nr+1) - 225 T(nr) + 50 T(nr-1) - 9 T(nr-2)) / (60 dr). */
int vvs_local_nu(vvs_run *run, double *inner, double *outer);
/* rhsw / rhsv / rhsu (sect360v2mg.f:999, 1034, 1066): which = 1 / 2 / 3 */
int vvs_rhs(vvs_run *run, int which);
/* poisson1 / poisson2 / poisson3 (sect360v2mg.f:1357, 1462, 1559): solves
 * the operator of relax1/2/3's residual (sect360v2mg.f:1817, 2034, 2233)
 * for wel / vel / uel from wrh / vrh / urh exactly; psip <- old psi, then
 * the walls and the mirrored ghost values the reference writes.  The
 * reference iterates to an absolute residual of 1e-5 (tolcon,
 * sect360v2mg.f:1668); this solve is exact to rounding. */
int vvs_poisson(vvs_run *run, int which);
/* update (sect360v2mg.f:936): wall vorticity on the end walls, the
 * cylinders and the barrier's faces (sect360v2mg.f:964-974), wall
 * temperatures, then the time levels rotate */
int vvs_update(vvs_run *run);
/* n steps of the reference time loop (sect360v2mg.f:235-249); trans1 of
 * the last step in *trans1 (may be NULL) */
int vvs_step(vvs_run *run, int n, double *trans1);

#ifdef __cplusplus
}
#endif
#endif
