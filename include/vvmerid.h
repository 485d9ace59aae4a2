/* vvmerid.h -- C ABI of the MI355X meridional-plane step with swirl on a
 * stretched grid (libvvmi.so).  FP64 throughout; status codes and error text
 * are those of vvmi.h (VV_OK, VV_EINVAL, ..., vv_last_error()).
 *
 * The step is the velocity-vorticity form of rotann2.f's meridional (r,z)
 * plane of a rotating annulus: azimuthal vorticity, swirl and temperature
 * advanced by its DuFort-Frankel leapfrog (vorticity 268, velocity 324,
 * temperature 379; the Coriolis terms through Re, centrifugal buoyancy
 * through Ra, the couplings Ra/(Re Pr), rotann2.f:291-306, 342-345), its
 * boundaries (update 497: vorticity from the wall velocities, th = 0 inside
 * and 1 outside, rotann2.f:395-404), its poisson (598) as two exact velocity
 * solves, its heat (468) -- on nodes that the caller gives (tanh clustering
 * in practice; the reference's grid.f:723-858 makes stretched nodes by
 * geometric expansion), with the weights of its wfdif.f:59-85 (Fornberg) for
 * every three-point derivative and a fourth-order compact scheme on the
 * uneven nodes for the advection derivatives.  The equations:
 *   (1/Pr)(w_t + u w_r + ww w_z - u w/r) = lap w - w/r^2 + (2/Pr) v v_z / r
 *        + 2 Re v_z - Ra r th_z - 2 Ra/(Re Pr) (v th_z + th v_z)
 *   (1/Pr)(v_t + u v_r + ww v_z + u v/r) = lap v - v/r^2 - 2 Re u
 *        + 2 Ra/(Re Pr) th u
 *   th_t + u th_r + ww th_z = lap th
 *   lap u - u/r^2 = w_z,   lap ww = -w_r - w/r
 * (w the azimuthal vorticity, v the swirl, u and ww the radial and axial
 * velocities; u = v = ww = 0 on the walls; the lids insulated).
 *
 * Fields are (nz+1) x (nr+1) doubles, X(i,j) at host[j*(nr+1) + i], node
 * (r[i], z[j]), i = 0..nr, j = 0..nz.  Every call on a run holds the run's
 * lock.
 */
#ifndef VVMERID_H
#define VVMERID_H
#include "vvmi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vvm_run vvm_run;

typedef struct {
    int    nr, nz;        /* intervals in r and z: 8 <= nr, nz <= 1024          */
    double dt;
    double Ra, Re, Pr;    /* Rayleigh, Re = sqrt(Ta) (Re = 0: no rotation), Prandtl */
} vvm_config;

enum vvm_field {
    VVM_W = 0, VVM_WP, VVM_WN,      /* azimuthal vorticity: current, previous, new */
    VVM_V, VVM_VP, VVM_VN,          /* swirl                                        */
    VVM_TH, VVM_THP, VVM_THN,       /* temperature                                  */
    VVM_U, VVM_WW,                  /* radial and axial velocity                    */
    VVM_FU, VVM_FW,                 /* their Poisson right-hand sides               */
    VVM_NFIELD
};

/* Allocates the fields and the solver on `device`.  r[0..nr] and z[0..nz]
 * are the nodes, strictly increasing, r[0] > 0.  Both velocity operators
 * are solved exactly: the eigenbasis of the z rows (one decomposition, at
 * create), one tridiagonal solve along r per z eigenvalue. */
int vvm_create(const vvm_config *cfg, const double *r, const double *z, int device, vvm_run **out);
int vvm_destroy(vvm_run *run);

/* whole-field copies (enum vvm_field) */
int vvm_put(vvm_run *run, int field, const double *host);
int vvm_get(vvm_run *run, int field, double *host);

/* The compact first derivatives (w_r, w_z, v_r, v_z, th_r, th_z) at the
 * inner nodes; vvm_get_derivative copies derivative d (0..5, that order). */
int vvm_derivatives(vvm_run *run);
int vvm_get_derivative(vvm_run *run, int d, double *host);
/* vorticity, velocity and temperature of rotann2.f (one DuFort-Frankel
 * step of w, v, th into wn, vn, thn at the inner nodes; the walls of wn take
 * the current wall vorticity), on the last vvm_derivatives */
int vvm_explicit(vvm_run *run);
/* fu = w_z, fw = -w_r - w/r at the inner nodes, from wn */
int vvm_rhs(vvm_run *run);
/* which = 1: u from fu; 2: ww from fw (exact solve, walls zero) */
int vvm_poisson(vvm_run *run, int which);
/* max |L psi - f| / max |f| over the inner nodes of the last solve of
 * `which` */
int vvm_residual(vvm_run *run, int which, double *relres);
/* walls of the new levels, then the time levels rotate */
int vvm_update(vvm_run *run);
/* n whole steps in the order derivatives, explicit, rhs, poisson 1 and 2,
 * update (replayed from captured HIP graphs of one and of ten steps) */
int vvm_step(vvm_run *run, int n);
/* enable = 0: vvm_step launches the kernels of each step one by one (the
 * same kernels; for tools that trace launches); 1: graphs (the default) */
int vvm_graphs(vvm_run *run, int enable);
/* heat(trans) of rotann2.f:468: the fluxes dth/dr through the inner (out[0])
 * and outer (out[1]) cylinder, trapezoid-integrated over z */
int vvm_heat(vvm_run *run, double out[2]);
/* kinetic energy 1/2 sum (u^2 + v^2 + ww^2) r with trapezoid weights (no
 * counterpart in the reference) */
int vvm_energy(vvm_run *run, double *ke);

/* Per-kernel device time of the following calls (HIP events around each
 * launch, the steps launched one by one; no counterpart in the reference):
 * launches, summed milliseconds and the summed algorithmic work -- bytes of
 * the fields read and written once each, or, for the transform GEMMs,
 * floating-point operations (2 M N K). */
enum vvm_kernel_id {
    VVM_K_DERIV = 0,   /* compact derivatives (r lines in LDS, z lines one per thread) */
    VVM_K_EXPLICIT,    /* DuFort-Frankel step of w, v, th                               */
    VVM_K_RHS,         /* Poisson right-hand sides                                      */
    VVM_K_ZFWD,        /* z eigenbasis, forward (GEMM; flops)                           */
    VVM_K_RSOLVE,      /* tridiagonal solves along r (cyclic reduction in LDS)          */
    VVM_K_ZINV,        /* z eigenbasis, inverse (GEMM; flops)                           */
    VVM_K_UPDATE,      /* walls and the level rotation                                  */
    VVM_K_COUNT
};
int vvm_profile(vvm_run *run, int enable);
int vvm_profile_read(vvm_run *run, int kernel, long long *launches, double *ms, double *work);

#ifdef __cplusplus
}
#endif
#endif
