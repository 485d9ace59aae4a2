/* vvm_oracle.h -- CPU oracle (test infrastructure only) of the meridional
 * (r,z) plane of a rotating annulus with swirl, on a stretched grid.  Only
 * the tests, the smoke check and the benchmark's CPU leg use it; no product
 * code links it.
 *
 * The reference holds this physics as a stream-function/vorticity program on
 * a uniform grid (rotann2.f: azimuthal vorticity, swirl and temperature,
 * rotann2.f:268-420; Coriolis terms through Re, centrifugal buoyancy through
 * Ra and the coupling terms Ra/(Re Pr), rotann2.f:291-305, 333-341), and
 * stretched grids only as a grid generator (grid.f:723-858) and Fornberg's
 * finite-difference weights (wfdif.f:59-85).  The composite restated here is
 * its velocity-vorticity form:
 *   (1/Pr)(w_t + u w_r + w w_z - u w/r) = lap w - w/r^2 + (2/Pr) v v_z / r
 *        + 2 Re v_z - Ra r th_z - 2 Ra/(Re Pr) (v th_z + th v_z)
 *   (1/Pr)(v_t + u v_r + w v_z + u v/r) = lap v - v/r^2 - 2 Re u
 *        + 2 Ra/(Re Pr) th u
 *   th_t + u th_r + w th_z = lap th
 *   lap u - u/r^2 = omega_z,   lap w = -omega_r - omega/r
 * (omega the azimuthal vorticity, v the swirl, u and w the radial and axial
 * velocities), advanced by the DuFort-Frankel leapfrog of rotann2.f
 * (rotann2.f:281-306) with the advection derivatives by a fourth-order
 * compact scheme on the stretched nodes, the other derivatives by Fornberg's
 * three-point weights, walls: u = v = w = 0, th = 0 inside and 1 outside
 * (rotann2.f:395-404), insulated top and bottom.
 *
 * Layout: F(i,j) at F[j*(nr+1) + i], i = 0..nr (r nodes), j = 0..nz.
 */
#ifndef VVM_ORACLE_H
#define VVM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Fornberg's weights (wfdif.f:59-85): c[k*n + q] is the weight of node q for
 * the k-th derivative at xx, k = 0..m, from the n nodes x[0..n-1] */
void vm_fornberg(const double *x, int n, double xx, int m, double *c);
/* tanh-clustered nodes s(q/n) = (1 + tanh(beta (2 q/n - 1))/tanh(beta))/2,
 * q = 0..n, scaled to [a, b] (beta = 0: uniform) */
void vm_nodes(int n, double a, double b, double beta, double *x);

typedef struct {
    int nr, nz;
    double dt, Ra, Re, Pr;
    const double *r, *z;   /* nodes r[0..nr], z[0..nz] */
} vm_params;

typedef struct {
    double *w, *wp, *wn;      /* azimuthal vorticity (current, previous, new) */
    double *v, *vp, *vn;      /* swirl */
    double *th, *thp, *thn;   /* temperature */
    double *u, *ww;           /* radial and axial velocity */
    double *fu, *fw;          /* their Poisson right-hand sides */
} vm_fields;

/* compact first derivative of one line of n+1 nodes x[0..n] at the nodes
 * 1..n-1 (out[1..n-1]); the end derivatives from 5-point Fornberg weights */
void vm_compact_line(int n, const double *x, const double *f, double *out);
/* the six advection derivatives d[0..5] = (w_r, w_z, v_r, v_z, th_r, th_z)
 * at the interior nodes */
void vm_derivatives(const vm_params *P, const vm_fields *F, double *const d[6]);
/* one DuFort-Frankel step of w, v, th at the interior nodes into wn, vn,
 * thn (the first step of a run starts with the previous levels equal to
 * the current ones); the walls of wn take the current wall vorticity */
void vm_explicit(const vm_params *P, vm_fields *F, double *const d[6]);
/* Poisson right-hand sides fu = omega_z, fw = -omega_r - omega/r at the
 * interior nodes, from wn */
void vm_rhs(const vm_params *P, vm_fields *F);
/* L psi on the interior nodes for which = 1 (u: lap - 1/r^2) or 2 (w:
 * lap); walls zero */
void vm_apply(const vm_params *P, int which, const double *psi, double *out);
/* walls of the new levels (wall vorticity from the new velocities, th = 0 /
 * 1 at the cylinders, insulated top and bottom, v = 0), then the levels
 * rotate: p <- current <- new */
void vm_update(const vm_params *P, vm_fields *F);
/* heat flux through the inner and outer cylinder:
 * out[0] = sum_j (trapezoid weight in z) * dth/dr at r0, out[1] at rmax */
void vm_heat(const vm_params *P, const vm_fields *F, double out[2]);

#ifdef __cplusplus
}
#endif
#endif
