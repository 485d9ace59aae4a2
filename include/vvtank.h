/* vvtank.h -- C ABI of the MI355X 3D horizontal-convection tank step
 * (libvvmi.so).  FP64 throughout; status codes and error text are those of
 * vvmi.h (VV_OK, VV_EINVAL, ..., vv_last_error()).
 *
 * The tank has no single reference program; its step composes:
 *   - rot3dmg.f's 3D velocity-vorticity step (rot3dmg.f:241-253): its
 *     staggered grid, DuFort-Frankel vorticity and temperature updates with
 *     the vortex stretching terms (Avorticity 341, Bvorticity 477,
 *     Cvorticity 601, temperature 736), its rotation terms (the Coriolis
 *     part of RHS2, rot3dmg.f:456-458) and axial gravity through Gr, its
 *     wall vorticity (update, 964) and right-hand sides (rhsw/v/u,
 *     1026-1134), its three Poisson operators (1695-2165) -- in Cartesian
 *     form (no metric terms, no centrifugal buoyancy);
 *   - the compact-scheme advection derivatives of the 2D codes along all
 *     three axes (velvor2dperiodicrbconv.f:607-628, 1080-1102), as batched
 *     tridiagonal solves (its tdma, 3046-3064);
 *   - the tank boundaries of velvor2dhorconvtank.f:1720-1745: insulated
 *     bottom and walls, the surface temperature linear along the tank
 *     (thn = k/np), stress-free side walls (zero vorticity), and its
 *     heat(trans1,trans2) (1574-1622) at the bottom and the surface.
 *
 * Axes: i = x across the tank (walls at both ends, nr intervals of lx/nr),
 * j = z vertical (gravity along -z, rotation about z; nz intervals of
 * lz/nz; the surface is the top, j = nz+1), k = y along the tank (side
 * walls; np intervals of ly/np).  Fields are (np+2) x (nz+2) x (nr+2)
 * doubles, X(i,j,k) at host[(k*ldj + j)*ldi + i]: th, u, w and B (= omega_y)
 * live at the y-cells k = 1..np with ghost planes k = 0 and np+1; v, A
 * (= omega_x) and C (= omega_z) at the y-faces k = 0..np, the side walls
 * being k = 0 and k = np.
 *
 * Every call on a run holds the run's lock.
 */
#ifndef VVTANK_H
#define VVTANK_H
#include "vvmi.h"
#include "vvmi3.h"   /* the field identifiers: enum vv3_field */

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vvt_run vvt_run;

typedef struct {
    int    nr, nz, np;    /* intervals in x, z, y: nr, nz >= 6, np >= 6       */
    double dt;
    double lx, lz, ly;    /* tank width, depth and length                     */
    double Re;            /* rotation: Re = sqrt(Ta) (rot3dmg.f's Re)         */
    double Ptl;           /* Prandtl number                                   */
    double Gr;            /* Grashof number; Gr Ptl is the Rayleigh number    */
} vvt_config;

/* Allocates the fields, the derivative fields and the Poisson solver on
 * `device`: each of the three operators is solved exactly and separably
 * (cosine or sine modes along y, the eigenbasis of its z rows, a banded
 * solve along x per mode pair).  The y modes are taken by FFTs in LDS when
 * np is a power of two from 8 to 2048, by matrix-core GEMMs otherwise. */
int vvt_create(const vvt_config *cfg, int device, vvt_run **out);
int vvt_destroy(vvt_run *run);

/* whole-field copies (enum vv3_field); ldi >= nr+2, ldj >= nz+2 (0 selects
 * nr+2 / nz+2); np+2 planes */
int vvt_put(vvt_run *run, int field, const double *host, int ldi, int ldj);
int vvt_get(vvt_run *run, int field, double *host, int ldi, int ldj);

/* The compact first derivatives of A, B, C and th along x, y and z at each
 * field's own points (12 batched line solves).  vvt_get_derivative copies
 * derivative d = 3 f + a (f: A, B, C, th; a: x, y, z) to the host. */
int vvt_derivatives(vvt_run *run);
int vvt_get_derivative(vvt_run *run, int d, double *host);
/* Avorticity, Bvorticity, Cvorticity and temperature on the derivatives of
 * the last vvt_derivatives call: An, Bn, Cn, thn at their interior points */
int vvt_vorticity(vvt_run *run);
int vvt_temperature(vvt_run *run);
/* heat(trans1,trans2) of velvor2dhorconvtank.f:1574-1622: the fluxes
 * through the bottom (trans1) and the surface (trans2) over the warm half
 * k = np/2..np-1, integrated over x; either pointer may be NULL */
int vvt_heat(vvt_run *run, double *trans1, double *trans2);
/* rhsw / rhsv / rhsu (which = 1 / 2 / 3) */
int vvt_rhs(vvt_run *run, int which);
/* poisson1 / 2 / 3: psip <- psi (whole field), psi <- the exact solution
 * of the operator on its unknowns, zero walls, ghost planes of u and w
 * mirrored */
int vvt_poisson(vvt_run *run, int which);
/* wall vorticity and temperature boundaries, then the time levels rotate */
int vvt_update(vvt_run *run);
/* n whole steps (the order of rot3dmg.f:241-253); the fluxes of the last */
int vvt_step(vvt_run *run, int n, double *trans1, double *trans2);
/* kinetic energy 1/2 dr dz dp sum (u^2 + v^2 + w^2) over the unknowns of
 * the three solves (no counterpart in the reference) */
int vvt_energy(vvt_run *run, double *ke);

/* Per-kernel device time of the following calls (HIP events around each
 * launch on the run's stream; no counterpart in the reference).  For each
 * kernel: launches, summed milliseconds and the summed algorithmic work of
 * those launches -- bytes of the fields read and written once each, or, for
 * the transform GEMMs, floating-point operations (2 M N K). */
enum vvt_kernel_id {
    VVT_K_DERIV_YZ = 0,  /* compact derivatives along y and z (one thread per line) */
    VVT_K_DERIV_X,       /* compact derivatives along x (lines staged in LDS)       */
    VVT_K_AVORT,         /* not used: the A, B, C and temperature updates run       */
    VVT_K_BVORT,         /*   fused, as VVT_K_EXPLICIT                              */
    VVT_K_CVORT,         /*                                                         */
    VVT_K_TEMP,          /*                                                         */
    VVT_K_HEAT,          /* heat fluxes                                             */
    VVT_K_RHS,           /* right-hand sides                                        */
    VVT_K_PSI_PRE,       /* psip <- psi, walls cleared                              */
    VVT_K_YFWD,          /* y modes, forward (GEMM; work in flops)                  */
    VVT_K_ZFWD,          /* z eigenbasis, forward (GEMM; flops)                     */
    VVT_K_XSOLVE,        /* banded solves along x                                   */
    VVT_K_ZINV,          /* z eigenbasis, inverse (GEMM; flops)                     */
    VVT_K_YINV,          /* y modes, inverse (GEMM; flops)                          */
    VVT_K_UPDATE,        /* boundaries and the rotation copy                        */
    VVT_K_YFFT_FWD,      /* y modes, forward, by FFT (np a power of two; bytes)     */
    VVT_K_YFFT_INV,      /* y modes, inverse, by FFT (bytes)                        */
    VVT_K_EXPLICIT,      /* A, B, C and temperature updates (one pass over the      */
                         /*   fields they read)                                     */
    VVT_K_COUNT
};
int vvt_profile(vvt_run *run, int enable);
int vvt_profile_read(vvt_run *run, int kernel, long long *launches, double *ms, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
