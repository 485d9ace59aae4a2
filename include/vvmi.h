/*
 * vvmi.h -- C ABI of the MI355X velocity-vorticity time-step library
 * (libvvmi.so, HIP for gfx950).
 *
 * The library replaces, behind a thin ISO_C_BINDING shim, the hot path of
 * the thesis codes: the explicit compact-scheme DuFort-Frankel step of
 * vorticity and temperature, the multigrid Poisson solves for the velocity
 * components, and the wall heat-flux diagnostic.  Every entry point below
 * names the reference routine it stands in for (paths are relative to the
 * reference repository; rb.f = monashubuntudir_velvor2dperiodicrbconvdir/
 * velvor2dperiodicrbconv.f).
 *
 * Conventions
 *   - Host arrays use the reference's column-major layout F(0:kr,1,0:kp):
 *     element (i,k) at F[k*ld + i] with ld = kr+1 (the caller's PARAMETER).
 *   - i = 0..nr+1 runs across the walls ("r"), k = 0..np along the periodic
 *     direction ("phi"); k = 0 and k = np are the same physical column.
 *   - All arithmetic is IEEE binary64.
 *   - Every function returns VV_OK (0) or a nonzero status; no C++
 *     exception crosses the boundary.  vv_last_error() returns a message.
 *   - A vv_run owns device storage for one simulation.  Every call on a run
 *     holds the run's lock, so calls on one run from several threads are
 *     serialised (the reference's main program calls rhsv/poisson2 and
 *     rhsu/poisson3 in two OpenMP sections, rb.f:370-380, which then reach
 *     one run from two threads); calls on different runs run concurrently.
 */
#ifndef VVMI_H
#define VVMI_H

#ifdef __cplusplus
extern "C" {
#endif

#define VV_OK            0
#define VV_NOT_CONVERGED 1    /* Poisson cycle cap hit (reference: STOP after 500 sweeps, rb.f:1843) */
#define VV_EINVAL       -1    /* bad argument (sizes, null pointers)        */
#define VV_EHIP         -2    /* HIP runtime error                          */
#define VV_ENOMEM       -3    /* device allocation failed                   */
#define VV_ENONFINITE   -4    /* NaN/Inf met in a Poisson right-hand side   */

/* field identifiers for vv_put / vv_get (the reference COMMON arrays,
 * rb.f:69-76) */
enum vv_field {
    VV_C = 0, VV_CP, VV_CN,        /* vorticity, previous, new           */
    VV_TH, VV_THP, VV_THN,         /* temperature, previous, new         */
    VV_UEL, VV_VEL, VV_UELP, VV_VELP,
    VV_VRH, VV_URH,                /* Poisson right-hand sides            */
    VV_MEAN_T2, VV_MEAN_T, VV_MEAN_U2, VV_MEAN_V2,   /* running means     */
    VV_NFIELDS
};

/* Geometry and physics of one run: the reference's COMMON /num/ /val/
 * /radius/ (rb.f:72-75) plus the knobs it hard-codes.
 * geom = VV_GEOM_BOX: periodic Cartesian box of rb.f (r == 1 in all metric
 *        terms, dp = width/np);
 * geom = VV_GEOM_ANNULUS: rotating-annulus (r,phi) plane of
 *        iarchomedir_.../velvor2dnewsch3.4.f ("ann.f"): r = r0 + i*dr,
 *        cylindrical terms, centrifugal buoyancy 2Ra(..)/Re/Pr
 *        (ann.f:981-987), dp = 2 pi/np (ann.f:124), width ignored. */
#define VV_GEOM_BOX     0
#define VV_GEOM_ANNULUS 1
typedef struct vv_config {
    int    nr, np;        /* intervals across / along (8 <= nr <= 2048,
                             8 <= np <= 16383: lines are held in LDS; any np:
                             phi coarsening halves np while it is even; the
                             direct level's transforms are mixed-radix FFTs
                             (radix 2 and the odd factor) for even column
                             counts up to 8192, dense transforms (2 c^2
                             weights) otherwise, for at most 4096 columns:
                             an np whose direct level would be larger, e.g.
                             odd np above 4095 or 16382 = 2 x 8191, is
                             rejected) */
    double dt;
    double r0, rmax;      /* walls; dr = (rmax-r0)/nr                       */
    double width;         /* periodic length; dp = width/np (rb.f:132-133) */
    double Ra, Pr;
    double robert;        /* Robert filter coefficient (0.02 rb.f:1043, 0.05 ann.f:1055) */
    double rtol;          /* Poisson stop: inf-norm residual <= rtol*|rhs| */
    int    max_cycles;    /* V-cycle cap per solve                          */
    int    nu1, nu2;      /* zebra line sweeps before/after coarse grid     */
    int    stats;         /* 1: running means + CFL every step (rb.f:409-458) */
    int    geom;          /* VV_GEOM_BOX or VV_GEOM_ANNULUS                 */
    double Re;            /* rotational Reynolds number (annulus only)     */
    int    direct_cols;   /* phi coarsening stops at <= this many columns,
                             where the level is solved exactly (real DFT in
                             phi + one line solve per mode); >= np makes the
                             whole solve direct.  0 selects the default:
                             direct for np <= 8192 (one exact solve per
                             Poisson call, measured 1.8-3.4x faster than the
                             V-cycle at every grid from 65 x 33 to
                             8193 x 1025), a V-cycle down to 128 columns
                             above.                                         */
    int    precision;     /* VV_PREC_FP64 (0): the V-cycle in FP64;
                             VV_PREC_MIXED (1): FP64 defect correction around
                             an FP32 V-cycle -- the residual f - Lu, the
                             update u += e, the coarsest (direct) level, the
                             stop test and every reduction in FP64; the zebra
                             sweeps, the restriction and the prolongation of
                             the correction equation in FP32.  Applies to
                             V-cycles (direct_cols < np); a direct solve of
                             the whole plane is FP64 either way.            */
    int    lines;         /* the compact schemes' line solves (tdma, ptdma):
                             VV_LINES_REFERENCE (0): the reference's
                             recurrences in its order, bit for bit;
                             VV_LINES_SCAN (1): one wavefront per line, the
                             same elimination factors with the affine
                             recurrences composed by a scan across the
                             lanes -- equal to rounding, not bit for bit
                             (lines of up to 2048 rows; longer ones take
                             the reference form).  No counterpart in the
                             reference.                                     */
} vv_config;

#define VV_PREC_FP64  0
#define VV_PREC_MIXED 1
#define VV_LINES_REFERENCE 0
#define VV_LINES_SCAN      1

/* per-step diagnostics (device-resident until read) */
typedef struct vv_step_info {
    double trans1, trans2;     /* wall heat flux, rb.f heat() (rb.f:1452)  */
    double cflu, cflv;         /* rb.f:421-458                              */
    double relres_v, relres_u; /* final relative residual of each solve    */
    int    cycles_v, cycles_u; /* V-cycles used                             */
    int    nclamp;             /* temperature clamps (rb.f:1614-1621)       */
    int    status;
} vv_step_info;

typedef struct vv_run vv_run;

const char *vv_last_error(void);
int vv_device_count(int *n);

/* Create a run on a device.  Replaces the static COMMON storage and
 * poissoninit (rb.f:1724): the multigrid hierarchy is built here.
 * nlev (out, may be NULL) receives the number of multigrid levels. */
int vv_create(const vv_config *cfg, int device, vv_run **out, int *nlev);
int vv_destroy(vv_run *run);
/* Re-arms a run for a new configuration of the same grid, geometry and
 * solver (nr, np, geom, r0, rmax, width, direct_cols, nu1, nu2, precision;
 * VV_EINVAL otherwise): the physics (dt, Ra, Pr, Re, robert), the stop and
 * the statistics switch take the new values, every field and the solver
 * state return to those of a fresh vv_create, and the tables are kept.  No
 * counterpart in the reference; for callers that keep runs resident across
 * requests (the served sweep). */
int vv_reset(vv_run *run, const vv_config *cfg);

/* Copy one field between a host array (leading dimension ld >= nr+2,
 * np+1 columns) and the run. */
int vv_put(vv_run *run, int field, const double *host, int ld);
int vv_get(vv_run *run, int field, double *host, int ld);

/* The reference start-up copy Xp = X (rb.f:315-329) and step counter reset. */
int vv_start(vv_run *run);

/* Explicit parts of the time loop on the device-resident state.
 * first != 0 selects the forward-Euler first-step forms.
 *   vv_cvorticity  -> Cvorticity1 / Cvorticity   (rb.f:584, rb.f:786)
 *   vv_temperature -> temperature1 / temperature (rb.f:1056, rb.f:1220)
 *   vv_heat        -> heat(trans1,trans2)        (rb.f:1452)
 *   vv_rhs         -> rhsv(vrh) and rhsu(urh)    (rb.f:1629, rb.f:1679)
 *   vv_update      -> update                     (rb.f:1557)            */
int vv_cvorticity(vv_run *run, int first);
int vv_temperature(vv_run *run, int first);
int vv_heat(vv_run *run, double *trans1, double *trans2);
int vv_rhs(vv_run *run);
int vv_update(vv_run *run, int *nclamp);

/* Kinetic energy of the current velocities.  No counterpart in the
 * reference (its diagnostics are the wall heat fluxes); the definition is
 *   ke = 1/2 dr dp sum_{k=0..np-1} [ sum_{i=1..nr}   wT(i) vel(i,k)^2
 *                                  + sum_{i=1..nr-1} wC(i) uel(i,k)^2 ]
 * over the unknowns of poisson2 / poisson3, with w = 1 in the box and the
 * row radius in the annulus (wT = r0 + (i-1/2) dr, wC = r0 + i dr), i.e.
 * 1/2 of the integral of |u|^2 over the domain by the midpoint rule. */
int vv_energy(vv_run *run, double *ke);

/* Poisson solves with the reference's in/out conventions: initial guess
 * 2*psi - psip, on return psip = previous psi, psi = solution with zero
 * wall rows and column 0 equal to column np.
 *   which = 2 -> poisson2(vel,velp,vrh,...)  (rb.f:1796)
 *   which = 3 -> poisson3(uel,uelp,urh,...)  (rb.f:1930)
 *   which = 5 -> both, batched in the same launches                      */
int vv_poisson(vv_run *run, int which, int *cycles, double *relres);

/* Host-array drop-in of poisson2/poisson3 with the reference argument
 * list: psi, psip, rhs are host arrays with leading dimension ld.
 * Replaces poisson2 (rb.f:1796) and poisson3 (rb.f:1930). */
int vv_poisson_host(vv_run *run, int which, double *psi, double *psip,
                    const double *rhs, int ld, int *cycles, double *relres);

/* nsteps complete time steps (rb.f:341-470 loop body, the first step of a
 * run in its forward-Euler form) without host copies; info (may be NULL)
 * receives the diagnostics of the last step. */
int vv_step(vv_run *run, int nsteps, vv_step_info *info);
/* The relative residual after each cycle of the last solve of v and of u,
 * up to max entries each (0 past a solve's cycles; at most 64 are kept);
 * ncycles[0], ncycles[1] = their cycle counts.  No counterpart in the
 * reference. */
int vv_solve_history(vv_run *run, double *relres_v, double *relres_u, int max, int *ncycles);

/* Diagnostics of the steps run so far: trans1/trans2 per step (rb.f:384-388
 * multiplies them by (rmax-r0)/width to get the Nusselt numbers).
 * Copies min(n, steps) pairs of the most recent steps. */
int vv_heat_history(vv_run *run, double *trans, int n);

/* Timing of the last vv_step call on the run's stream (milliseconds). */
int vv_last_step_ms(vv_run *run, double *ms);

/* Per-kernel timing with HIP events on the run's stream (no counterpart in
 * the reference; used by the benchmark for the roofline figure).
 * vv_profile(run, 1) clears and starts collection, 0 stops it.
 * vv_profile_read returns, for one kernel id, the launches, the summed
 * device time (ms) and the summed algorithmic bytes of those launches. */
enum vv_kernel_id {
    VV_K_COMPACT = 0,  /* compact derivatives (Thomas / cyclic line solves)      */
    VV_K_DF,           /* DuFort-Frankel update of C and T + Robert filter      */
    VV_K_HEAT,         /* wall heat flux                                          */
    VV_K_RHS,          /* Poisson right-hand sides and initial guess             */
    VV_K_ZEBRA,        /* multigrid line sweeps (restriction and prolongation    */
                       /* fused into the first sweep of a level)                 */
    VV_K_RESTRICT,     /* restriction onto the coarsest level                    */
    VV_K_PROLONG,      /* synthesis of the coarsest-level correction             */
    VV_K_NORM,         /* residual max-norm and stop test                        */
    VV_K_DIRECT,       /* spectral solve of the coarsest level                   */
    VV_K_OUT,          /* solution -> velocity fields                            */
    VV_K_UPDATE,       /* wall update, clamp, running means, CFL                 */
    VV_K_COUNT
};
int vv_profile(vv_run *run, int enable);
int vv_profile_read(vv_run *run, int kernel, long long *launches, double *ms, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
