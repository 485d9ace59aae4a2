/* vvband.h -- batched banded line solvers of the reference on device arrays
 * (libvvmi.so, HIP for gfx950).
 *
 * The thesis codes solve their compact-scheme and line-relaxation systems
 * with three small routines, called once per line from loops over the other
 * index (rb.f = monashubuntudir_velvor2dperiodicrbconvdir/
 * velvor2dperiodicrbconv.f):
 *   tdma(a,b,c,y,r,n)      rb.f:3046-3064  tridiagonal, Thomas, no pivoting;
 *                          a(1) and c(n) are not read
 *   ptdma(a,b,c,y,r,n)     rb.f:3066-3116  cyclic tridiagonal: a(1) couples
 *                          row 1 to row n and c(n) row n to row 1 (n >= 3)
 *   pdma(a,b,c,d,e,r,x,n)  rb.f:3118-3164  pentadiagonal: a, b below the
 *                          diagonal c, d, e above it (n >= 2)
 * Here a whole batch of such lines is one launch, one thread per line, with
 * the reference's operations in the reference's order, so every solution is
 * bit-for-bit the reference routine's.  The coefficient and right-hand-side
 * arrays are read only (the reference overwrites b and r, or a, b, r, or
 * c, d, e, r, as its own scratch; the modified rows go to a device scratch
 * buffer instead).
 *
 * Layout: nsys systems of n rows; row j (0-based) of system s of every array
 * is at p[s*ss + j*sr] (ss, sr >= 1 elements).  ss = 1, sr = nsys (lines
 * interleaved) gives coalesced loads.  All pointers are device pointers on
 * the current device; x may equal r.  d and e are read by VV_BAND_PDMA only
 * (NULL otherwise).  stream is a hipStream_t (NULL: the null stream); the
 * call returns after the launch, without synchronising.
 * Status codes and vv_last_error() are those of vvmi.h.
 */
#ifndef VVBAND_H
#define VVBAND_H
#include "vvmi.h"

#ifdef __cplusplus
extern "C" {
#endif

enum vv_band_kind {
    VV_BAND_TDMA = 0,     /* tdma  (rb.f:3046-3064), n >= 1 */
    VV_BAND_PTDMA = 1,    /* ptdma (rb.f:3066-3116), n >= 3 */
    VV_BAND_PDMA = 2      /* pdma  (rb.f:3118-3164), n >= 2 */
};

/* Solves the batch.  The scratch rows (2 words per row for tdma, 3 for
 * ptdma and pdma) are allocated on the stream for the call and freed after
 * it (stream-ordered).  Replaces the per-line loops around tdma / ptdma /
 * pdma, e.g. rb.f:607-628 (Cr by tdma) and rb.f:631-649 (Ch by ptdma). */
int vv_band_solve(int kind, int n, int nsys, long long ss, long long sr,
                  const double *a, const double *b, const double *c,
                  const double *d, const double *e, const double *r,
                  double *x, void *stream);

/* The same solve without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The scratch rows live in one static device array per GPU of
 * VV_BAND_NOALLOC_WORDS doubles (64 MiB), held from the first call on that
 * GPU until the process ends; a batch needing more (words per row x n x nsys)
 * is rejected with VV_EINVAL.  Every call on a GPU uses that GPU's one array,
 * so calls that overlap on one GPU (different streams, or different host
 * threads without ordering) race: order them on one stream.  Each call writes
 * the scratch rows before it reads them, so nothing carries over between
 * calls and nothing needs clearing. */
#define VV_BAND_NOALLOC_WORDS 8388608
int vv_band_solve_noalloc(int kind, int n, int nsys, long long ss, long long sr,
                          const double *a, const double *b, const double *c,
                          const double *d, const double *e, const double *r,
                          double *x, void *stream);

#ifdef __cplusplus
}
#endif
#endif
