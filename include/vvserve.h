/* vvserve.h -- C ABI of the request loop that serves many velocity-vorticity
 * runs on one GPU at once (libvvmi.so).
 *
 * No counterpart in the reference: the thesis programs run one simulation
 * per process.  This is the served parameter sweep: requests (a run's
 * configuration, initial temperature and step count) arrive while others
 * are on the GPU; a request is admitted when a slot is free and its device
 * footprint fits the memory budget (first come, first served); each
 * admitted run steps on its own stream from its own worker thread, so the
 * work of different runs overlaps on the device.  A run's results do not
 * depend on what else is running: every run has its own fields, multigrid
 * hierarchy and stream.  Runs stay resident between requests: a request of
 * the same grid, geometry and solver as an idle run re-arms it (vv_reset)
 * instead of creating one; idle runs (at most two per slot) are destroyed
 * when the memory budget needs their footprint.
 */
#ifndef VVSERVE_H
#define VVSERVE_H
#include "vvmi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vv_server vv_server;

typedef struct {
    vv_config cfg;       /* the run (as for vv_create)                              */
    int nsteps;          /* time steps to run (vv_start, then vv_step)              */
    int ld;              /* leading dimension of th0 / th_out / vel_out (>= nr+2)   */
    const double *th0;   /* initial temperature (np+1 columns of ld); copied at
                            submission; NULL leaves the run's zero field           */
    double *th_out;      /* optional: final temperature, written before the result
                            is posted (caller-owned, must stay valid)              */
    double *vel_out;     /* optional: final v velocity, likewise                    */
    void *user;          /* returned with the result                                */
} vv_request;

typedef struct {
    long long id;        /* as returned by vv_server_submit                        */
    int status;          /* VV_OK, VV_NOT_CONVERGED or a negative error code       */
    int steps;
    double trans1, trans2;              /* wall fluxes of the last step            */
    double t_submit, t_admit, t_done;   /* seconds since the server was created    */
    void *user;
    char error[128];     /* vv_last_error() of the failing call, if status < 0      */
} vv_result;

typedef struct {
    int queued, running, done;
    int peak_running;
    double mem_in_use, peak_mem;        /* admitted footprint of the running and
                                           the resident idle runs, bytes          */
    int resident;        /* idle runs kept for later requests                      */
    int reused, created; /* requests served by a resident run / by a new one       */
} vv_server_stats;

/* Device footprint a run of this configuration is admitted with (bytes): an
 * upper estimate of its fields, multigrid hierarchy and tables. */
double vv_run_footprint(const vv_config *cfg);

/* slots worker threads on `device`; mem_budget bytes of admitted footprint */
int vv_server_create(int device, int slots, double mem_budget, vv_server **out);
/* queues a request; returns its id (>= 0) or a negative status (VV_EINVAL if
 * the request alone exceeds the budget or its sizes are bad) */
long long vv_server_submit(vv_server *s, const vv_request *req);
/* takes the next completed result; waits up to timeout_s seconds (< 0:
 * forever); returns VV_OK, or 1 if none completed in time */
int vv_server_wait(vv_server *s, vv_result *res, double timeout_s);
int vv_server_stats_read(vv_server *s, vv_server_stats *st);
/* finishes the queued and running requests, joins the workers, frees */
int vv_server_destroy(vv_server *s);

#ifdef __cplusplus
}
#endif
#endif
