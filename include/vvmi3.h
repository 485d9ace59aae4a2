/* vvmi3.h -- C ABI of the MI355X 3D annulus time step (rot3dmg.f: 3D
 * velocity-vorticity model on a staggered grid, periodic in phi, walls in r
 * and z; and its stationary horizontal-annulus variant rot3dmghor.f, model
 * VV3_MODEL_HOR).  FP64 throughout; status codes and error text are shared
 * with vvmi.h (VV_OK, VV_EINVAL, ..., vv_last_error()).
 *
 * Fields are (np+1) x (nz+2) x (nr+2) doubles, i fastest, then j, then k:
 * X(i,j,k) at host[(k*(nz+2) + j)*(nr+2) + i], i = 0..nr+1, j = 0..nz+1,
 * k = 0..np, as the reference's (0:kr,0:kz,0:kp) arrays with kr = nr+1,
 * kz = nz+1.  k = 0 and k = np are the same azimuth (both are kept).
 *
 * Every call on a run holds the run's lock, so calls on one run from
 * several threads are serialised; calls on different runs run concurrently.
 */
#ifndef VVMI3_H
#define VVMI3_H
#include "vvmi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vv3_run vv3_run;

/* which program's physics the run steps */
enum vv3_model {
    VV3_MODEL_ROT3DMG = 0,   /* rot3dmg.f: rotating annulus (Coriolis and
                                centrifugal terms through Re, axial gravity
                                through Gr), T = 0 inside / 1 outside      */
    VV3_MODEL_HOR = 1        /* rot3dmghor.f: stationary horizontal annulus,
                                gravity in the r-theta plane through Gr
                                (rot3dmghor.f:479, 606, 749), no Coriolis
                                terms, T = 1 inside / 0 outside
                                (rot3dmghor.f:1042), heat flux through both
                                cylinders near mid-axis (rot3dmghor.f:913) */
};

/* the scalars of rot3dmg.f's input deck (rot3dmg.f:59-86) */
typedef struct {
    int    nr, nz, np;          /* intervals in r, z, phi (np even)          */
    double dt;
    double r0, rmax, zmax;      /* inner and outer radius, height            */
    double Ra, Re, Ptl, Gr;     /* Rayleigh, Reynolds, Prandtl, Grashof      */
    double dr, dz, dp;          /* grid spacings as the caller computed them
                                   (COMMON /num/); 0 derives them as
                                   rot3dmg.f:91-93 does                       */
    int    model;               /* enum vv3_model (0: rot3dmg.f)             */
} vv3_config;

enum vv3_field {
    VV3_AN = 0, VV3_A, VV3_AP, VV3_BN, VV3_B, VV3_BP, VV3_CN, VV3_C, VV3_CP,
    VV3_THN, VV3_TH, VV3_THP, VV3_UEL, VV3_VEL, VV3_WEL, VV3_UELP, VV3_VELP,
    VV3_WELP, VV3_URH, VV3_VRH, VV3_WRH, VV3_NFIELD
};

/* Allocates the fields and the Poisson solver on `device` (the reference
 * builds its multigrid levels in poissoninit, rot3dmg.f:1153).  The three
 * Poisson operators are solved exactly and separably: a real DFT in phi, the
 * eigenbasis of the z rows, a banded solve in r per (mode, z eigenvalue).
 * Needs nr, nz >= 4 (nz >= 6 for VV3_MODEL_HOR) and an even np,
 * 4 <= np <= 8192 (VV_EINVAL otherwise). */
int vv3_create(const vv3_config *cfg, int device, vv3_run **out);
int vv3_destroy(vv3_run *run);

/* whole-field copies; the host array is X(i,j,k) at
 * host[(k*ldj + j)*ldi + i] (ldi >= nr+2, ldj >= nz+2; 0 selects nr+2 / nz+2),
 * so a COMMON array of rot3dmg.f, real*8 X(0:kr,0:kz,0:kp), is passed with
 * ldi = kr+1, ldj = kz+1 (rot3dmg.f:23-29) */
int vv3_put(vv3_run *run, int field, const double *host, int ldi, int ldj);
int vv3_get(vv3_run *run, int field, double *host, int ldi, int ldj);

/* Avorticity, Bvorticity, Cvorticity (rot3dmg.f:341, 477, 601): An, Bn, Cn
 * on their interior points, with the hybrid upwind switch of the reference.
 * which = 1 / 2 / 3 runs one routine, 0 all three in one launch (they read
 * only the old levels, so the order does not matter) */
int vv3_vorticity(vv3_run *run, int which);
/* temperature (rot3dmg.f:736): thn on i = 1..nr, j = 0..nz+1 */
int vv3_temperature(vv3_run *run);
/* heat (rot3dmg.f:882): inner-wall heat flux trans1; for VV3_MODEL_HOR
 * the inner-cylinder flux of heat(trans1,trans2) (rot3dmghor.f:913), whose
 * outer-cylinder flux vv3_fluxes returns */
int vv3_heat(vv3_run *run, double *trans1);
/* the wall fluxes of the last heat evaluation (vv3_heat or vv3_step):
 * trans1, and trans2 of rot3dmghor.f (0 for rot3dmg.f); either may be NULL */
int vv3_fluxes(vv3_run *run, double *trans1, double *trans2);
/* Kinetic energy of the current velocities (no counterpart in the
 * reference):
 *   ke = 1/2 dr dz dp sum_{k=0..np-1} [ sum_{j=1..nz,   i=1..nr-1} (r0 + i dr) uel^2
 *                                     + sum_{j=1..nz,   i=1..nr} (r0 + (i-1/2) dr) vel^2
 *                                     + sum_{j=1..nz-1, i=1..nr} (r0 + (i-1/2) dr) wel^2 ],
 * each over the unknowns of its Poisson solve: 1/2 of the volume integral
 * of |u|^2 by the midpoint rule. */
int vv3_energy(vv3_run *run, double *ke);
/* Local Nusselt numbers at mid-height of both cylinders, one per plane
 * k = 0..np (inner and outer hold np+1 values): the formula of
 * sect360v2mg.f:318-325 at j = nz/2,
 *   inner[k] = -r0 ln(r0) (-184 T0 + 225 T1 - 50 T2 + 9 T3) / (60 dr),
 *   outer[k] = -rmax ln(r0) (184 T(nr+1) - 225 T(nr) + 50 T(nr-1) - 9 T(nr-2)) / (60 dr). */
int vv3_local_nu(vv3_run *run, double *inner, double *outer);
/* Azimuthal Fourier-mode volume integrals of the temperature after
 * rot3dmghor.f:954-991 (subroutine aaaaaaaa), modes n = 1..nmax
 * (1 <= nmax <= 16):
 *   c[n-1] = dp dr dz sum_{k=1..np} sum_{i=0..nr+1} sum_{j=0..nz+1}
 *            f_i f_j (r0 + i dr) T(i,j,k) cos(n k dp),  s[n-1] with sin,
 * f = 1/2 on the wall rows i = 0, nr+1 and j = 0, nz+1.  The reference
 * routine reads r0 from an uninitialised local and carries each plane's
 * running sum into the next plane; the run's r0 and one sum per plane are
 * used here. */
int vv3_modes(vv3_run *run, int nmax, double *c, double *s);
/* rhsw / rhsv / rhsu (rot3dmg.f:1026, 1066, 1111): which = 1 / 2 / 3 */
int vv3_rhs(vv3_run *run, int which);
/* poisson1 / poisson2 / poisson3 (rot3dmg.f:1216, 1325, 1434): solves the
 * operator of relax1/2/3's residual (rot3dmg.f:1695, 1927, 2130) for wel /
 * vel / uel from wrh / vrh / urh, exactly; psip <- old psi, walls 0 as the
 * reference writes them.  The reference iterates to an absolute residual of
 * 1e-5 (tolcon, rot3dmg.f:1562); this solve is exact to rounding. */
int vv3_poisson(vv3_run *run, int which);
/* update (rot3dmg.f:964): wall vorticity and wall temperatures, then the
 * time levels rotate (Ap <- A <- An, ... for A, B, C, th) */
int vv3_update(vv3_run *run);
/* n steps of the reference time loop (rot3dmg.f:240-253); trans1 of the
 * last step is returned in *trans1 (may be NULL) */
int vv3_step(vv3_run *run, int n, double *trans1);

#ifdef __cplusplus
}
#endif
#endif
