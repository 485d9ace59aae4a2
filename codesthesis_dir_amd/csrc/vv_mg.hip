// vv_mg.hip -- multigrid Poisson solver for the velocity components.
//
// The converged discrete solution is fixed by the reference's residual
// operators (rb.f relax2 residual rb.f:2306-2357 for v, relax3 residual
// rb.f:2553-2587 for u; the annulus forms ann.f:2301-2348, 2555-2583): a
// periodic 3-point second difference along phi (times 1/r^2 in the annulus)
// plus the r operator D_r, which for v carries the one-sided closure rows
// (-2, 5/3, -1/5)/dr^2 at both walls.  That closure is not preserved under
// r-coarsening (a point-smoother V-cycle with rediscretised coarse grids
// diverges on it), so the cycle here never coarsens r:
//   - smoother: zebra r-line Gauss-Seidel; every line is solved exactly by
//     parallel cyclic reduction in LDS, with the reduction coefficients of
//     the level's line operator precomputed on the host (the closure rows
//     are first folded into a tridiagonal system); both colours of a sweep
//     run in one launch (k_mg_zebra);
//   - coarsening: phi only, full weighting / linear interpolation, down to
//     direct_cols columns or fewer, or to the first odd column count; that
//     level is solved exactly: the operator is invariant along phi, so a
//     real DFT over the level's columns decouples it into one line system
//     per Fourier mode (k_mg_fft_fwd or k_mg_dft<0>: project; k_mg_modes:
//     line solves; k_mg_fft_inv or k_mg_dft<1>: synthesise).
// Both velocity problems of a run are batched into every launch
// (blockIdx.y = problem).  Kernels return at once when their problem has
// converged, so several cycles can be queued between host checks.

#include "vv_internal.h"
#include "vv_fft.h"
#include "vv_reduce.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

// The solver's arithmetic is not tied to the reference's operation order
// (its converged solution is fixed by the operators, and the stop test is a
// tolerance), so multiply-adds are fused here.  The explicit-step kernels
// (vv_explicit.hip) keep every rounding of the reference.
#pragma clang fp contract(fast)

namespace vv {

namespace {

constexpr int MAXN = 2048;     // rows per line held in LDS
constexpr int MAXS = 11;       // cyclic-reduction steps for n <= MAXN

// Line smoothing modes of k_mg_zebra.
enum { ZB_PLAIN = 0, ZB_RESTRICT = 1, ZB_PROLONG = 2 };

// Level and problem descriptors are kernel arguments (passed by value), so
// no dependent load precedes a kernel's first data load.  T is the type of
// the smoothed levels: double, or float for the inner cycle of the mixed
// precision defect correction (the direct level is double either way).
template <class T>
struct KLevT {
    T *u, *uo, *f;          // current iterate, output buffer of a sweep, rhs
    const T *lu;
    const double *twid;
    const double2 *ftw;     // direct level: FFT twiddles (null: dense transforms)
    double *xm;
    T *dlt;                 // black-line changes of the last pre-sweep (level L-1)
    double cphi;
    int ncol, nstep, tabsz;
};
template <class T>
struct KProbT {
    const T *sten, *rphi;
    int n;
    KLevT<T> L, F, Cc;      // the launch's level, the finer level, the coarser level
};
template <class T>
struct KArgsT {
    KProbT<T> P[MG_MAXPROB];
    const int *done, *cycles;
    int mode, lev0;
    int store_delta;        // k_mg_zebra: store the black-line changes in L.dlt
};
using KLev = KLevT<double>;
using KProb = KProbT<double>;
using KArgs = KArgsT<double>;

__device__ __forceinline__ size_t cix(int j, int q, int n) { return (size_t)q * n + j; }

// residual f - (D_r u + cphi*rphi*(u_{q+1} - 2u + u_{q-1})) at (j, q)
template <class T>
__device__ __forceinline__ T resid_at(const KProbT<T> &P, const KLevT<T> &L, int j, int q)
{
    const int n = P.n, ncol = L.ncol;
    const T *u = L.u;
    const T *s = P.sten + 5 * j;
    T lu = 0;
#pragma unroll
    for (int o = -2; o <= 2; ++o) {
        const int jj = j + o;
        if (jj >= 0 && jj < n && s[o + 2] != 0.0) lu += s[o + 2] * u[cix(jj, q, n)];
    }
    const T uc = u[cix(j, q, n)];
    if (ncol > 1) {
        const int qm = (q == 0) ? ncol - 1 : q - 1;
        const int qp = (q == ncol - 1) ? 0 : q + 1;
        lu += (T)L.cphi * P.rphi[j] * (u[cix(j, qp, n)] - (T)2 * uc + u[cix(j, qm, n)]);
    }
    return L.f[cix(j, q, n)] - lu;
}

// Restricted residual of the finer level F at coarse line qc, row j (full
// weighting along phi, centred on fine line 2qc+1).
template <class T>
__device__ __forceinline__ T restricted_at(const KProbT<T> &P, const KLevT<T> &F, int j, int qc)
{
    const int q0 = 2 * qc, q1 = 2 * qc + 1, q2 = (2 * qc + 2) % F.ncol;
    const T r0 = resid_at(P, F, j, q0);
    const T r1 = resid_at(P, F, j, q1);
    const T r2 = resid_at(P, F, j, q2);
    return (T)0.25 * r0 + (T)0.5 * r1 + (T)0.25 * r2;
}

// The reduction coefficients of one line operator for the rows a thread owns
// (rows tid + k*nt, k < R): alpha and beta of each of the S steps and the
// final 1/diagonal, in registers.
template <int R, class T>
__device__ __forceinline__ void pcr_coef(const T *__restrict__ tab, int S, int n, T a[R][MAXS],
                                         T b[R][MAXS], T di[R])
{
    const int tid = threadIdx.x, nt = blockDim.x;
    const T *alpha = tab + 2, *beta = tab + 2 + (size_t)S * n, *dinv = tab + 2 + 2 * (size_t)S * n;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * nt;
        const bool in = i < n;
#pragma unroll
        for (int t = 0; t < MAXS; ++t) {
            a[k][t] = (in && t < S) ? alpha[(size_t)t * n + i] : (T)0;
            b[k][t] = (in && t < S) ? beta[(size_t)t * n + i] : (T)0;
        }
        di[k] = in ? dinv[i] : (T)0;
    }
}

// Solves NL r-lines that share one operator, in LDS, by parallel cyclic
// reduction with the coefficients a, b, di of pcr_coef.  rhs[l][k] is the
// thread's right-hand side of line l for row tid + k*nt; with FOLD the
// closure rows are folded here (tab[0], tab[1]), otherwise the caller has
// folded them.  Solutions return in x[l][k].  The caller separates two calls
// that reuse buf by a barrier.
template <int R, int NL, bool FOLD, class T>
__device__ __forceinline__ void pcr_solve(const T *__restrict__ tab, int S, int n, T (*buf)[2][MAXN],
                                          T rhs[NL][R], T x[NL][R], const T a[R][MAXS],
                                          const T b[R][MAXS], const T di[R])
{
    const int tid = threadIdx.x, nt = blockDim.x;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * nt;
#pragma unroll
        for (int l = 0; l < NL; ++l)
            if (i < n) buf[l][0][i] = rhs[l][k];
    }
    __syncthreads();
    if (FOLD && n > 2) {   // r0 -= g0 r1, r_{n-1} -= gN r_{n-2}
        const bool o0 = (tid == 0), oN = (tid == (n - 1) % nt);
        T r0[NL], rN[NL];
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            r0[l] = o0 ? buf[l][0][0] - tab[0] * buf[l][0][1] : (T)0;
            rN[l] = oN ? buf[l][0][n - 1] - tab[1] * buf[l][0][n - 2] : (T)0;
        }
        __syncthreads();
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            if (o0) buf[l][0][0] = r0[l];
            if (oN) buf[l][0][n - 1] = rN[l];
        }
        __syncthreads();
    }
    int cur = 0;
#pragma unroll
    for (int t = 0; t < MAXS; ++t) {
        if (t < S) {            // S is uniform over the block
            const int s = 1 << t;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const int i = tid + k * nt;
                if (i < n) {
#pragma unroll
                    for (int l = 0; l < NL; ++l) {
                        T v = buf[l][cur][i];
                        if (i >= s) v += a[k][t] * buf[l][cur][i - s];
                        if (i + s < n) v += b[k][t] * buf[l][cur][i + s];
                        buf[l][cur ^ 1][i] = v;
                    }
                }
            }
            __syncthreads();
            cur ^= 1;
        }
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * nt;
#pragma unroll
        for (int l = 0; l < NL; ++l) x[l][k] = (i < n) ? buf[l][cur][i] * di[k] : (T)0;
    }
}

// pcr_coef + pcr_solve for a single use of an operator
template <int R, int NL, bool FOLD, class T>
__device__ __forceinline__ void pcr_lines(const T *__restrict__ tab, int S, int n,
                                          T (*buf)[2][MAXN], T rhs[NL][R], T x[NL][R])
{
    T a[R][MAXS], b[R][MAXS], di[R];
    pcr_coef<R, T>(tab, S, n, a, b, di);
    pcr_solve<R, NL, FOLD, T>(tab, S, n, buf, rhs, x, a, b, di);
}

// One zebra sweep (both colours) of one level, from buffer u to buffer uo.
// Workgroup bx works through G consecutive line groups (G = ncol/2 over the
// grid size, 1 on small levels), loading the level's reduction coefficients
// once for all of them.  Group g owns black (odd) line qb = 2(bx G + g)+1 and
// red (even) line qr0 = qb-1; it also solves the red line qr1 = qb+1 that the black line
// needs (the workgroup that owns qr1 computes the same values from the same
// inputs), so both colours run in one launch:
//   red lines qr0, qr1 from the old black lines (u), in one reduction pass
//   (the lines share the level's operator); black line qb from the new red
//   values, which stay in registers.
// Modes: ZB_RESTRICT (first pre-sweep of a coarse level): f is the restricted
// residual of the finer level (stored for the owned lines) and the old
// iterate is the zero initial correction; ZB_PROLONG (first post-sweep): the
// old black lines are read with the coarser level's correction added (a line
// solve overwrites its own line, so only these reads need it).
// store_delta (last pre-sweep above the coarsest level): the change of each
// black line is kept in L.dlt.  After the sweep the black lines satisfy their
// line equations and a red line e has the residual
//   r_e = -c rphi (delta_{e-1} + delta_{e+1}),
// so the restriction onto the coarsest level needs only these changes.
template <int R, class T>
__global__ void __launch_bounds__(1024) k_mg_zebra(KArgsT<T> a)
{
    const int p = blockIdx.y;
    const KProbT<T> &P = a.P[p];
    const KLevT<T> &L = P.L;
    const int n = P.n, ncol = L.ncol;
    // With one row per thread the coefficients are loaded once for all line
    // groups; with two rows per thread they leave no registers for that (the
    // launch keeps one group per workgroup and each solve loads them).
    constexpr bool HOIST = (R == 1);
    const int G = HOIST ? (ncol / 2 + (int)gridDim.x - 1) / (int)gridDim.x : 1;
    const int g0 = (int)blockIdx.x * G;
    if (ncol < 2 || 2 * g0 + 1 >= ncol) return;
    const int done = a.done[p];            // its latency overlaps the loads below
    T ca[R][MAXS], cb[R][MAXS], cdi[R];
    if constexpr (HOIST) pcr_coef<R, T>(L.lu, L.nstep, n, ca, cb, cdi);
    if (done) return;                       // uniform over the workgroup
    for (int g = 0; g < G; ++g) {
    const int qb = 2 * (g0 + g) + 1;
    if (qb >= ncol) break;                  // uniform
    if (g > 0) __syncthreads();             // the previous group's last reads of buf
    const int qr0 = qb - 1;
    const int qr1 = (qb + 1 == ncol) ? 0 : qb + 1;
    const int qbm = (qr0 == 0) ? ncol - 1 : qr0 - 1;          // black left of qr0
    const int qbp = (qr1 == ncol - 1) ? 0 : qr1 + 1;          // black right of qr1
    __shared__ T buf[2][2][MAXN];
    const int tid = threadIdx.x, nt = blockDim.x;
    const T c = (T)L.cphi;
    const int mode = a.mode;
    const T *u = L.u;
    const T *e = P.Cc.u;
    T rr[2][R], xr[2][R], rb[1][R], xb[1][R], fb[R], fr0[R], ubo[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * nt;
        rr[0][k] = rr[1][k] = 0;
        fb[k] = fr0[k] = ubo[k] = 0;
        if (i >= n) continue;
        const T rp = P.rphi[i];
        if (mode == ZB_RESTRICT) {
            const T f0 = restricted_at(P, P.F, i, qr0);
            const T f1 = restricted_at(P, P.F, i, qr1);
            fb[k] = restricted_at(P, P.F, i, qb);
            fr0[k] = f0;
            rr[0][k] = f0;
            rr[1][k] = f1;
        } else if (mode == ZB_PROLONG) {
            const T um0 = u[cix(i, qbm, n)] + e[cix(i, qbm >> 1, n)];
            const T ub = u[cix(i, qb, n)] + e[cix(i, qb >> 1, n)];
            const T up1 = u[cix(i, qbp, n)] + e[cix(i, qbp >> 1, n)];
            rr[0][k] = L.f[cix(i, qr0, n)] - c * rp * (um0 + ub);
            rr[1][k] = L.f[cix(i, qr1, n)] - c * rp * (ub + up1);
            fb[k] = L.f[cix(i, qb, n)];
            ubo[k] = ub;
        } else {
            const T ub = u[cix(i, qb, n)];
            rr[0][k] = L.f[cix(i, qr0, n)] - c * rp * (u[cix(i, qbm, n)] + ub);
            rr[1][k] = L.f[cix(i, qr1, n)] - c * rp * (ub + u[cix(i, qbp, n)]);
            fb[k] = L.f[cix(i, qb, n)];
            ubo[k] = ub;
        }
    }
    if (mode == ZB_RESTRICT) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int i = tid + k * nt;
            if (i < n) {
                L.f[cix(i, qr0, n)] = fr0[k];
                L.f[cix(i, qb, n)] = fb[k];
            }
        }
    }
    if constexpr (HOIST) pcr_solve<R, 2, true, T>(L.lu, L.nstep, n, buf, rr, xr, ca, cb, cdi);
    else pcr_lines<R, 2, true, T>(L.lu, L.nstep, n, buf, rr, xr);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * nt;
        rb[0][k] = (i < n) ? fb[k] - c * P.rphi[i] * (xr[0][k] + xr[1][k]) : (T)0;
    }
    if constexpr (HOIST) pcr_solve<R, 1, true, T>(L.lu, L.nstep, n, buf, rb, xb, ca, cb, cdi);
    else pcr_lines<R, 1, true, T>(L.lu, L.nstep, n, buf, rb, xb);
    const bool sd = a.store_delta;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * nt;
        if (i < n) {
            L.uo[cix(i, qr0, n)] = xr[0][k];
            L.uo[cix(i, qb, n)] = xb[0][k];
            if (sd) L.dlt[cix(i, qb >> 1, n)] = xb[0][k] - ubo[k];
        }
    }
    }   // line groups
}

// The same sweep with one line group per workgroup and the coefficients
// loaded by each solve: for lines of more than 1024 rows (two rows per
// thread), where holding the coefficients across both solves spills.
template <int R, class T>
__global__ void __launch_bounds__(1024) k_mg_zebra1(KArgsT<T> a)
{
    const int p = blockIdx.y;
    const KProbT<T> &P = a.P[p];
    const KLevT<T> &L = P.L;
    const int n = P.n, ncol = L.ncol;
    const int qb = 2 * blockIdx.x + 1;
    if (ncol < 2 || qb >= ncol) return;
    const int done = a.done[p];            // its latency overlaps the loads below
    const int qr0 = qb - 1;
    const int qr1 = (qb + 1 == ncol) ? 0 : qb + 1;
    const int qbm = (qr0 == 0) ? ncol - 1 : qr0 - 1;          // black left of qr0
    const int qbp = (qr1 == ncol - 1) ? 0 : qr1 + 1;          // black right of qr1
    __shared__ T buf[2][2][MAXN];
    const int tid = threadIdx.x, nt = blockDim.x;
    const T c = (T)L.cphi;
    const int mode = a.mode;
    const T *u = L.u;
    const T *e = P.Cc.u;
    T rr[2][R], xr[2][R], rb[1][R], xb[1][R], fb[R], fr0[R], ubo[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * nt;
        rr[0][k] = rr[1][k] = 0;
        fb[k] = fr0[k] = ubo[k] = 0;
        if (i >= n) continue;
        const T rp = P.rphi[i];
        if (mode == ZB_RESTRICT) {
            const T f0 = restricted_at(P, P.F, i, qr0);
            const T f1 = restricted_at(P, P.F, i, qr1);
            fb[k] = restricted_at(P, P.F, i, qb);
            fr0[k] = f0;
            rr[0][k] = f0;
            rr[1][k] = f1;
        } else if (mode == ZB_PROLONG) {
            const T um0 = u[cix(i, qbm, n)] + e[cix(i, qbm >> 1, n)];
            const T ub = u[cix(i, qb, n)] + e[cix(i, qb >> 1, n)];
            const T up1 = u[cix(i, qbp, n)] + e[cix(i, qbp >> 1, n)];
            rr[0][k] = L.f[cix(i, qr0, n)] - c * rp * (um0 + ub);
            rr[1][k] = L.f[cix(i, qr1, n)] - c * rp * (ub + up1);
            fb[k] = L.f[cix(i, qb, n)];
            ubo[k] = ub;
        } else {
            const T ub = u[cix(i, qb, n)];
            rr[0][k] = L.f[cix(i, qr0, n)] - c * rp * (u[cix(i, qbm, n)] + ub);
            rr[1][k] = L.f[cix(i, qr1, n)] - c * rp * (ub + u[cix(i, qbp, n)]);
            fb[k] = L.f[cix(i, qb, n)];
            ubo[k] = ub;
        }
    }
    if (done) return;                       // uniform over the workgroup
    if (mode == ZB_RESTRICT) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int i = tid + k * nt;
            if (i < n) {
                L.f[cix(i, qr0, n)] = fr0[k];
                L.f[cix(i, qb, n)] = fb[k];
            }
        }
    }
    pcr_lines<R, 2, true, T>(L.lu, L.nstep, n, buf, rr, xr);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * nt;
        rb[0][k] = (i < n) ? fb[k] - c * P.rphi[i] * (xr[0][k] + xr[1][k]) : (T)0;
    }
    pcr_lines<R, 1, true, T>(L.lu, L.nstep, n, buf, rb, xb);
    const bool sd = a.store_delta;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * nt;
        if (i < n) {
            L.uo[cix(i, qr0, n)] = xr[0][k];
            L.uo[cix(i, qb, n)] = xb[0][k];
            if (sd) L.dlt[cix(i, qb >> 1, n)] = xb[0][k] - ubo[k];
        }
    }
}

// Restriction onto the coarsest level L from the black-line changes of the
// last pre-sweep of level L-1 (see k_mg_zebra): coarse point k sits on black
// line 2k+1, whose red neighbours carry -c rphi (delta_{k-1} + delta_k) and
// -c rphi (delta_k + delta_{k+1}), and the black lines have no residual, so
// full weighting gives f_L = -c/4 rphi (delta_{k-1} + 2 delta_k + delta_{k+1}).
// T: the type of the finer level's changes; the coarsest level's f (fp)
// is double.
template <class T>
__global__ void __launch_bounds__(256) k_mg_restrict(KArgsT<T> a, double *f0, double *f1)
{
    const int p = blockIdx.y;
    const KProbT<T> &P = a.P[p];
    const int n = P.n, C = P.F.ncol / 2;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * C) return;
    const int done = a.done[p];
    const int k = (int)(t / n), j = (int)(t - (long long)k * n);
    const int km = (k == 0) ? C - 1 : k - 1, kp = (k == C - 1) ? 0 : k + 1;
    const T *d = P.F.dlt;
    const double v = -0.25 * P.F.cphi * (double)P.rphi[j] *
                     ((double)d[cix(j, km, n)] + 2.0 * (double)d[cix(j, k, n)] + (double)d[cix(j, kp, n)]);
    if (done) return;
    (p ? f1 : f0)[cix(j, k, n)] = v;
}

// Mixed precision: the defect of the FP64 iterate on level 0, f - Lu, as the
// FP32 right-hand side of the correction equation, whose iterate starts at
// zero (FP32 level-0 buffers f32, u32 of each problem).
struct MixPtr {
    float *f[MG_MAXPROB], *u[MG_MAXPROB];
    const double *src[MG_MAXPROB];
    double *dst[MG_MAXPROB];
    long long cnt[MG_MAXPROB];
};
__global__ void __launch_bounds__(256) k_mg_defect(KArgs a, MixPtr m)
{
    const int p = blockIdx.y;
    if (a.done[p]) return;
    const KProb &P = a.P[p];
    const long long cnt = (long long)P.n * P.L.ncol;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < cnt;
         t += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(t / P.n), j = (int)(t - (long long)q * P.n);
        m.f[p][t] = (float)resid_at(P, P.L, j, q);
        m.u[p][t] = 0.0f;
    }
}
// the direct level's FP64 correction -> its FP32 copy (src -> u)
__global__ void __launch_bounds__(256) k_mg_to32(const int *done, MixPtr m)
{
    const int p = blockIdx.y;
    if (done[p]) return;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < m.cnt[p];
         t += (long long)gridDim.x * blockDim.x)
        m.u[p][t] = (float)m.src[p][t];
}
// u += e: the FP32 correction of level 0 added to the FP64 iterate (u -> dst)
__global__ void __launch_bounds__(256) k_mg_correct(const int *done, MixPtr m)
{
    const int p = blockIdx.y;
    if (done[p]) return;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < m.cnt[p];
         t += (long long)gridDim.x * blockDim.x)
        m.dst[p][t] += (double)m.u[p][t];
}

// Exact solve of the coarsest level with C = ncol columns: the operator is
// invariant along phi, so the real DFT over the C columns decouples it into
// C line systems D_r + lambda_m diag(rphi), one per real Fourier line ml
// (ml = 0: m = 0; ml = 2m-1 / 2m: cos / sin of mode m; for even C,
// ml = C-1: the Nyquist mode m = C/2).  Three kernels: k_mg_dft<0> projects the
// residual on the lines, k_mg_modes solves each line by parallel cyclic
// reduction, k_mg_dft<1> synthesises the correction.  Below level 0 the
// correction starts from zero, so the residual is f.  On level 0 (the whole
// solve is direct) the first cycle solves for u from f (the initial guess is
// not needed by an exact solve) and later cycles project f - Lu: iterative
// refinement.

constexpr int DFT_TN = 64;     // rows per workgroup
constexpr int DFT_TM = 16;     // output lines (columns) per workgroup
constexpr int DFT_TK = 64;     // inner-dimension chunk staged in LDS

// Dense transform over the C columns of DFT_TN rows:
//   INV = 0: xm[o][i] = sum_k Wf[o][k] y[k][i], y = f (or the residual);
//   INV = 1: u[o][i] (+)= (1/C) sum_k Wi[o][k] xm[k][i].
// The weights (Wf | Wi, C x C each) are tabulated on the host; each
// workgroup stages DFT_TK x DFT_TN inputs and DFT_TM x DFT_TK weights in
// LDS per chunk.  Wave w computes outputs o0 + 4w .. o0 + 4w + 3 of row
// i0 + lane.
template <int INV>
__global__ void __launch_bounds__(256) k_mg_dft(KArgs a)
{
    const int p = blockIdx.y;
    const KProb &P = a.P[p];
    const KLev &L = P.L;
    const int n = P.n, C = L.ncol;
    const int i0 = blockIdx.x * DFT_TN, o0 = blockIdx.z * DFT_TM;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wavefront index (uniform)
    const int TM = C < DFT_TM ? C : DFT_TM;
    const int done = a.done[p];
    const bool from_res = !INV && a.lev0 && a.cycles[p] > 0;
    // below level 0 the correction is computed from zero in every cycle
    const bool store = INV && (!a.lev0 || a.cycles[p] == 0);
    __shared__ double Ys[DFT_TK][DFT_TN];
    __shared__ double Ws[DFT_TM][DFT_TK];
    const double *W = INV ? L.twid + (size_t)C * C : L.twid;
    const double *Y = INV ? L.xm : L.f;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    // staging: every thread issues all its loads into registers before the
    // LDS stores, so the loads of a chunk are in flight together
    constexpr int YPT = DFT_TK * DFT_TN / 256, WPT = DFT_TM * DFT_TK / 256;
    const int gi = i0 + lane;
    for (int k0 = 0; k0 < C; k0 += DFT_TK) {
        const int tk = (C - k0 < DFT_TK) ? C - k0 : DFT_TK;
        double yv[YPT], wv[WPT];
        if (from_res) {
#pragma unroll
            for (int r = 0; r < YPT; ++r) {
                const int kk = r * 4 + w;
                yv[r] = (kk < tk && gi < n) ? resid_at(P, L, gi, k0 + kk) : 0.0;
            }
        } else {
#pragma unroll
            for (int r = 0; r < YPT; ++r) {
                const int kk = r * 4 + w;
                yv[r] = (kk < tk && gi < n) ? Y[(size_t)(k0 + kk) * n + gi] : 0.0;
            }
        }
#pragma unroll
        for (int r = 0; r < WPT; ++r) {
            const int idx = threadIdx.x + r * 256;
            const int oo = idx / DFT_TK, kk = idx - oo * DFT_TK;
            wv[r] = (oo < TM && o0 + oo < C && kk < tk) ? W[(size_t)(o0 + oo) * C + k0 + kk] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < YPT; ++r) Ys[r * 4 + w][lane] = yv[r];
#pragma unroll
        for (int r = 0; r < WPT; ++r) {
            const int idx = threadIdx.x + r * 256;
            Ws[idx / DFT_TK][idx % DFT_TK] = wv[r];
        }
        __syncthreads();
        // wave w accumulates outputs 4w..4w+3; rows of Ws beyond TM are zero
#pragma unroll 8
        for (int kk = 0; kk < tk; ++kk) {
            const double y = Ys[kk][lane];
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[j] += Ws[4 * w + j][kk] * y;
        }
        __syncthreads();
    }
    const int i = gi;
    if (done || i >= n) return;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int oo = 4 * w + j;
        const int o = o0 + oo;
        if (oo >= TM || o >= C) continue;
        if (!INV) {
            L.xm[(size_t)o * n + i] = acc[j];
        } else {
            double *d = L.u + cix(i, o, n);
            const double v = acc[j] / (double)C;
            if (store) *d = v;
            else *d += v;
        }
    }
}

// Line solves of the coarsest level: workgroup w handles Fourier mode
// m = w; the cos and sin lines of 0 < m < C/2 share the mode's operator
// (one coefficient table), m = 0 and the Nyquist mode m = C/2 have one line.
// Reads the projections from xm and overwrites them with the solutions.
template <int R>
__global__ void __launch_bounds__(1024) k_mg_modes(KArgs a)
{
    const int p = blockIdx.y;
    const KProb &P = a.P[p];
    const KLev &L = P.L;
    const int n = P.n, C = L.ncol;
    const int m = blockIdx.x;
    if (m > C / 2) return;
    const int done = a.done[p];
    const bool pair = (m > 0 && 2 * m < C);
    const int ml0 = (m == 0) ? 0 : (pair ? 2 * m - 1 : C - 1);
    __shared__ double buf[2][2][MAXN];
    const int tid = threadIdx.x, nt = blockDim.x;
    double rhs[2][R], x[2][R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * nt;
        rhs[0][k] = (i < n) ? L.xm[(size_t)ml0 * n + i] : 0.0;
        rhs[1][k] = (i < n && pair) ? L.xm[(size_t)(ml0 + 1) * n + i] : 0.0;
    }
    if (done) return;
    const double *tab = L.lu + (size_t)m * L.tabsz;
    if (pair) pcr_lines<R, 2, true, double>(tab, L.nstep, n, buf, rhs, x);
    else pcr_lines<R, 1, true, double>(tab, L.nstep, n, buf, rhs, x);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * nt;
        if (i < n) {
            L.xm[(size_t)ml0 * n + i] = x[0][k];
            if (pair) L.xm[(size_t)(ml0 + 1) * n + i] = x[1][k];
        }
    }
}

// The direct level's transforms as FFTs in LDS (C even, 8 <= C <= 8192):
// radix 2 when C is a power of two (R = 0); otherwise C = b 2^bits with b odd,
// b radix-2 sub-lines per line and the in-place combination of fft_combine
// (R outputs per thread).  A workgroup transforms LP pairs of lines (rows i
// of the level, two real lines per complex line), the problem is blockIdx.y.  The
// projection writes the real lines of the modes, ml = 0 (m = 0), 2m-1 and 2m
// (cos and sin of 0 < m < C/2), C-1 (m = C/2), i.e. X[ml][i] = sum_k W[ml][k]
// Y[k][i] as the dense transform; Y is the level's right-hand side or, in a
// refinement cycle on level 0, its residual.  The synthesis forms
// (1/C) sum_m w_m (C_m cos + S_m sin) (w = 1 at m = 0, C/2, else 2) and
// stores it into, or on refinement adds it to, the level's iterate.
template <int R>
__global__ void __launch_bounds__(256) k_mg_fft_fwd(KArgs a, int bits, int b, int LP)
{
    extern __shared__ double2 z[];
    const int p = blockIdx.y;
    const KProb &P = a.P[p];
    const KLev &L = P.L;
    const int n = P.n, C = L.ncol, half = C >> 1, nl = 2 * LP;
    const int ib = blockIdx.x * nl;
    if (ib >= n || a.done[p]) return;   // uniform over the workgroup
    const bool from_res = a.lev0 && a.cycles[p] > 0;
    for (int e = threadIdx.x; e < nl * C; e += blockDim.x) {
        const int l = e % nl, k = e / nl;
        const int i = ib + l;
        double v = 0.0;
        if (i < n) v = from_res ? resid_at(P, L, i, k) : L.f[(size_t)k * n + i];
        const int pos = R ? mr_pos(k, b, C / b, bits) : (int)bitrev(k, bits);
        double *zz = reinterpret_cast<double *>(&z[(l >> 1) * C + pos]);
        zz[l & 1] = v;
    }
    __syncthreads();
    fft_stages<false>(z, L.ftw, C / b, bits, LP * b);
    if (R) fft_combine<false, R ? R : 1>(z, L.ftw + (C / b) / 2, C, C / b, b, LP);
    for (int e = threadIdx.x; e < nl * (half + 1); e += blockDim.x) {
        const int l = e % nl, m = e / nl;
        const int i = ib + l;
        if (i >= n) continue;
        const double2 zm = z[(l >> 1) * C + m], zn = z[(l >> 1) * C + (m == 0 ? 0 : C - m)];
        // A = (Z_m + conj Z_{C-m})/2, B = (Z_m - conj Z_{C-m})/(2i)
        double re, im;
        if ((l & 1) == 0) { re = 0.5 * (zm.x + zn.x); im = 0.5 * (zm.y - zn.y); }
        else { re = 0.5 * (zm.y + zn.y); im = -0.5 * (zm.x - zn.x); }
        if (m == 0) {
            L.xm[i] = re;
        } else if (m == half) {
            L.xm[(size_t)(C - 1) * n + i] = re;
        } else {
            L.xm[(size_t)(2 * m - 1) * n + i] = re;    // sum y cos
            L.xm[(size_t)(2 * m) * n + i] = -im;       // sum y sin
        }
    }
}

template <int R>
__global__ void __launch_bounds__(256) k_mg_fft_inv(KArgs a, int bits, int b, int LP)
{
    extern __shared__ double2 z[];
    const int p = blockIdx.y;
    const KProb &P = a.P[p];
    const KLev &L = P.L;
    const int n = P.n, C = L.ncol, half = C >> 1, nl = 2 * LP;
    const int ib = blockIdx.x * nl;
    if (ib >= n || a.done[p]) return;   // uniform over the workgroup
    // Z_m = X_m + i Y_m of the pair, X_m = C_m - i S_m (m <= C/2), X_{C-m} = conj X_m
    for (int e = threadIdx.x; e < LP * C; e += blockDim.x) {
        const int pr = e % LP, m = e / LP;
        const int mm = (m <= half) ? m : C - m;
        const double sg = (m <= half) ? -1.0 : 1.0;
        double xr = 0.0, xi = 0.0, yr = 0.0, yi = 0.0;
        for (int h2 = 0; h2 < 2; ++h2) {
            const int i = ib + 2 * pr + h2;
            if (i >= n) continue;
            double c, sn = 0.0;
            if (mm == 0) {
                c = L.xm[i];
            } else if (mm == half) {
                c = L.xm[(size_t)(C - 1) * n + i];
            } else {
                c = L.xm[(size_t)(2 * mm - 1) * n + i];
                sn = L.xm[(size_t)(2 * mm) * n + i];
            }
            if (h2 == 0) { xr = c; xi = sg * sn; } else { yr = c; yi = sg * sn; }
        }
        const int pos = R ? mr_pos(m, b, C / b, bits) : (int)bitrev(m, bits);
        z[pr * C + pos] = make_double2(xr - yi, xi + yr);
    }
    __syncthreads();
    fft_stages<true>(z, L.ftw, C / b, bits, LP * b);
    if (R) fft_combine<true, R ? R : 1>(z, L.ftw + (C / b) / 2, C, C / b, b, LP);
    const double inv = 1.0 / C;
    const bool store = !a.lev0 || a.cycles[p] == 0;
    for (int e = threadIdx.x; e < nl * C; e += blockDim.x) {
        const int l = e % nl, q = e / nl;
        const int i = ib + l;
        if (i >= n) continue;
        const double2 v = z[(l >> 1) * C + q];
        const double x = ((l & 1) ? v.y : v.x) * inv;
        double *d = L.u + cix(i, q, n);
        if (store) *d = x;
        else *d += x;
    }
}

// max-norm of the level-0 residual: one point per thread, one partial per
// workgroup (no atomics: same-address atomics from hundreds of workgroups
// cost more than a second launch)
__global__ void __launch_bounds__(256) k_mg_norm(KArgs a, double *npart)
{
    const int p = blockIdx.y;
    const KProb &P = a.P[p];
    const KLev &L = P.L;
    if (a.done[p]) return;
    const long long cnt = (long long)P.n * L.ncol;
    const long long stride = (long long)gridDim.x * blockDim.x;
    double m = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < cnt; t += stride) {
        const int q = (int)(t / P.n), j = (int)(t - (long long)q * P.n);
        m = nanmax(m, fabs(resid_at(P, L, j, q)));
    }
    __shared__ double sh[4];
    m = block_max(m, sh);
    if (threadIdx.x == 0) npart[p * MG_NORM_MAXB + blockIdx.x] = m;
}

struct StopArgs {
    int *done, *cycles;
    unsigned long long *fnorm;
    double *relres;
    double *hist;      // [MG_MAXPROB][MG_HIST]: relres after each cycle
    const double *npart, *fpart;
    int nbn, nbr, maxc;
    double rtol;
    int *check;    // deferred solves: the sticky word (else null)
};

// Stop test of one cycle (one workgroup): combines the residual partials of
// k_mg_norm and, at the first cycle of a solve, the |f| partials of the
// right-hand-side kernel; converged when the relative residual is <= rtol
// (a NaN residual also stops), capped at maxc cycles.  All loads are issued
// before the reductions.
__global__ void __launch_bounds__(256) k_mg_stop(StopArgs a)
{
    __shared__ double sh[4];
    int dn[MG_MAXPROB], cy[MG_MAXPROB];
    double fo[MG_MAXPROB], mr[MG_MAXPROB], mf[MG_MAXPROB];
#pragma unroll
    for (int p = 0; p < MG_MAXPROB; ++p) {
        dn[p] = a.done[p];
        cy[p] = a.cycles[p];
        fo[p] = __longlong_as_double((long long)a.fnorm[p]);
        mr[p] = 0.0;
        mf[p] = 0.0;
        for (int b = threadIdx.x; b < a.nbn; b += blockDim.x) mr[p] = nanmax(mr[p], a.npart[p * MG_NORM_MAXB + b]);
        for (int b = threadIdx.x; b < a.nbr; b += blockDim.x) mf[p] = nanmax(mf[p], a.fpart[p * a.nbr + b]);
    }
#pragma unroll
    for (int p = 0; p < MG_MAXPROB; ++p) {
        mr[p] = block_max(mr[p], sh);
        mf[p] = block_max(mf[p], sh);
    }
    if (threadIdx.x != 0) return;
#pragma unroll
    for (int p = 0; p < MG_MAXPROB; ++p) {
        if (dn[p]) continue;
        double fn = fo[p];
        if (cy[p] == 0) {
            fn = mf[p];
            a.fnorm[p] = (unsigned long long)__double_as_longlong(fn);
        }
        const double rel = mr[p] / (fn > 0.0 ? fn : 1.0);
        const int c = cy[p] + 1;
        a.cycles[p] = c;
        a.relres[p] = rel;
        if (c <= MG_HIST) a.hist[p * MG_HIST + c - 1] = rel;
        if (!(rel > a.rtol) || c >= a.maxc) a.done[p] = 1;
    }
    // deferred solves: a problem whose relative residual is above rtol (or
    // NaN; bit 0) or whose right-hand side norm is not finite (bit 1) sets
    // the sticky word, which the host reads once per vv_step call
    if (a.check) {
        int s = *a.check;
        for (int p = 0; p < 2; ++p) {
            if (!(a.relres[p] <= a.rtol)) s |= 1;
            if (!isfinite(__longlong_as_double((long long)a.fnorm[p]))) s |= 2;
        }
        *a.check = s;
    }
}

// level 0 -> fields: psip <- psi, psi <- solution (rb.f:1894-1913)
__global__ void __launch_bounds__(256) k_mg_out(const double *uv, const double *uu, double *vel, double *velp,
                                                double *uel, double *uelp, int nr, int np, int ld, int which)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int rows = nr + 2;
    if (t >= (long long)rows * (np + 1)) return;
    const int k = (int)(t / rows), i = (int)(t - (long long)k * rows);
    const size_t o = (size_t)k * ld + i;
    const int q = (k == 0) ? np - 1 : k - 1;
    if (which & 1) {
        const double old = vel[o];
        velp[o] = old;
        double nv = 0.0;
        if (i >= 1 && i <= nr) nv = uv[cix(i - 1, q, nr)];
        vel[o] = nv;
    }
    if ((which & 2) && i <= nr) {
        const double old = uel[o];
        uelp[o] = old;
        double nu = 0.0;
        if (i >= 1 && i <= nr - 1) nu = uu[cix(i - 1, q, nr - 1)];
        uel[o] = nu;
    }
}

inline int nblk(long long n, int wg) { return (int)((n + wg - 1) / wg); }

}  // namespace

// Parallel-cyclic-reduction tables of a line operator with rows
// s2 x_{i-2} + s1 x_{i-1} + d x_i + p1 x_{i+1} + p2 x_{i+2} = r_i, where only
// row 0 has p2 and only row n-1 has s2.  Those two couplings are removed
// first with rows 1 and n-2 (factors g0, gN); then the tridiagonal system is
// reduced in S = ceil(log2 n) steps: step t with stride s = 2^t replaces
// row i by row i + alpha*row(i-s) + beta*row(i+s).  The rhs follows the same
// linear combinations on the device; x_i = r_i^(S) * dinv_i.
int make_pcr(int n, const std::vector<double> &s2, const std::vector<double> &s1,
             const std::vector<double> &d, const std::vector<double> &p1,
             const std::vector<double> &p2, std::vector<double> &tab)
{
    std::vector<double> a(s1), b(d), c(p1);
    double g0 = 0.0, gN = 0.0;
    if (n > 2) {
        g0 = p2[0] / c[1];
        b[0] = b[0] - g0 * a[1];
        c[0] = c[0] - g0 * b[1];
        gN = s2[n - 1] / a[n - 2];
        a[n - 1] = a[n - 1] - gN * b[n - 2];
        b[n - 1] = b[n - 1] - gN * c[n - 2];
    }
    a[0] = 0.0;
    c[n - 1] = 0.0;
    int S = 0;
    while ((1 << S) < n) ++S;
    tab.assign(2 + 2 * (size_t)S * n + n, 0.0);
    tab[0] = g0;
    tab[1] = gN;
    double *alpha = tab.data() + 2, *beta = tab.data() + 2 + (size_t)S * n;
    std::vector<double> an(n), bn(n), cn(n);
    for (int t = 0; t < S; ++t) {
        const int s = 1 << t;
        for (int i = 0; i < n; ++i) {
            const double A = (i - s >= 0) ? -a[i] / b[i - s] : 0.0;
            const double Bc = (i + s < n) ? -c[i] / b[i + s] : 0.0;
            alpha[(size_t)t * n + i] = A;
            beta[(size_t)t * n + i] = Bc;
            an[i] = (i - s >= 0) ? A * a[i - s] : 0.0;
            cn[i] = (i + s < n) ? Bc * c[i + s] : 0.0;
            bn[i] = b[i] + ((i - s >= 0) ? A * c[i - s] : 0.0) + ((i + s < n) ? Bc * a[i + s] : 0.0);
        }
        a.swap(an);
        b.swap(bn);
        c.swap(cn);
    }
    double *dinv = tab.data() + 2 + 2 * (size_t)S * n;
    for (int i = 0; i < n; ++i) dinv[i] = 1.0 / b[i];
    return S;
}

int mg_build(vv_run *r)
{
    const int nr = r->nr, np = r->np;
    const double hhri = 1.0 / (r->dr * r->dr);
    if (nr > MAXN) { set_error("nr exceeds the line length held in LDS (2048)"); return VV_EINVAL; }
    const int Cd = r->cfg.direct_cols > 0 ? r->cfg.direct_cols : 1;
    int nlev = 1;
    // phi coarsening halves the column count while it is even
    int cdir = np;
    for (; cdir > Cd && cdir % 2 == 0; cdir /= 2) ++nlev;
    if (nlev > MG_MAXLEV) { set_error("too many multigrid levels"); return VV_EINVAL; }
    // a direct level without FFTs (fewer than 8 columns, odd, or above 8192)
    // holds 2 c^2 dense transform weights: bounded at c = 4096 (268 MB)
    if (!fft_cols(cdir) && cdir > MG_DENSE_MAXC) {
        set_error("the direct Poisson level would have " + std::to_string(cdir) +
                  " columns without an FFT (odd np, or direct_cols above 8192): at most " +
                  std::to_string(MG_DENSE_MAXC) + " columns are solved with dense transforms");
        return VV_EINVAL;
    }
    const int nvec[2] = {nr, nr - 1};
    // host-side tables first, then one pool
    struct LevHost { std::vector<double> tab; int S; double cphi; int ncol; int direct; size_t tabsz; };
    std::vector<LevHost> lh[2];
    std::vector<double> sten_h[2], rphi_h[2];
    size_t total = 0;
    for (int p = 0; p < 2; ++p) {
        const int n = nvec[p];
        // r operator stencil (i-2..i+2) and phi row factor: the level-1
        // residual operators of relax2 (rb.f:2306-2357, ann.f:2301-2348) and
        // relax3 (rb.f:2553-2587, ann.f:2555-2583)
        std::vector<double> s2(n, 0), s1(n, 0), dg(n, 0), p1(n, 0), p2(n, 0), rphi(n, 1.0);
        const bool ann = (r->cfg.geom == VV_GEOM_ANNULUS);
        const double hri = 1.0 / r->dr;
        for (int j = 0; j < n; ++j) {
            const int i = j + 1;
            double rvi = 1.0;
            if (ann) rvi = (p == 0) ? 1.0 / (r->cfg.r0 + ((double)i - 0.5) * r->dr)
                                    : 1.0 / ((double)i * r->dr + r->cfg.r0);
            const double rvs = rvi * rvi;
            rphi[j] = rvs;
            dg[j] = -2.0 * hhri + (ann ? rvs : 0.0);
            if (p == 0 && j == 0) {
                p1[j] = 5.0 / 3.0 * hhri + (ann ? 4.0 / 3.0 * hri * rvi : 0.0);
                p2[j] = -0.2 * hhri;
            } else if (p == 0 && j == n - 1) {
                s1[j] = 5.0 / 3.0 * hhri - (ann ? 4.0 / 3.0 * hri * rvi : 0.0);
                s2[j] = -0.2 * hhri;
            } else {
                if (j > 0) s1[j] = hhri - (ann ? 1.5 * hri * rvi : 0.0);
                if (j < n - 1) p1[j] = hhri + (ann ? 1.5 * hri * rvi : 0.0);
            }
        }
        sten_h[p].resize(5 * (size_t)n);
        for (int j = 0; j < n; ++j) {
            sten_h[p][5 * j + 0] = s2[j]; sten_h[p][5 * j + 1] = s1[j]; sten_h[p][5 * j + 2] = dg[j];
            sten_h[p][5 * j + 3] = p1[j]; sten_h[p][5 * j + 4] = p2[j];
        }
        rphi_h[p] = rphi;
        total += 6 * (size_t)n;
        double hp = r->dp;
        for (int l = 0, c = np; l < nlev; ++l, c /= 2, hp *= 2.0) {
            LevHost L;
            L.ncol = c;
            L.cphi = 1.0 / (hp * hp);
            L.direct = (l == nlev - 1);
            if (!L.direct) {
                std::vector<double> d2(dg);
                for (int j = 0; j < n; ++j) d2[j] = dg[j] - 2.0 * L.cphi * rphi[j];
                L.S = make_pcr(n, s2, s1, d2, p1, p2, L.tab);
                L.tabsz = L.tab.size();
                total += 3 * (size_t)n * c + L.tab.size();
                if (l == nlev - 2) total += (size_t)n * (c / 2);   // black-line changes
            } else {
                // one table per Fourier mode m = 0..c/2: the phi second
                // difference on c periodic points has eigenvalue
                // (2 cos(2 pi m/c) - 2)/hp^2
                const double pi = std::acos(-1.0);
                for (int m = 0; m <= c / 2; ++m) {
                    const double lam = (2.0 * std::cos(2.0 * pi * m / c) - 2.0) * L.cphi;
                    std::vector<double> d2(dg), t;
                    for (int j = 0; j < n; ++j) d2[j] = dg[j] + lam * rphi[j];
                    L.S = make_pcr(n, s2, s1, d2, p1, p2, t);
                    L.tabsz = t.size();
                    L.tab.insert(L.tab.end(), t.begin(), t.end());
                }
                // transform weights: Wf[ml][k] projects column k on real
                // line ml; Wi[q][ml] synthesises column q (times 1/c on
                // the device).  Angles are reduced mod c before the cos/sin.
                auto w = [&](int ml, int k, bool inv) {
                    if (ml == 0) return 1.0;
                    if (c % 2 == 0 && ml == c - 1) return (k & 1) ? -1.0 : 1.0;   // Nyquist
                    const int mm = (ml + 1) / 2;
                    const double ang = 2.0 * pi * (double)(((long long)mm * k) % c) / c;
                    const double v = (ml & 1) ? std::cos(ang) : std::sin(ang);
                    return inv ? 2.0 * v : v;
                };
                // (dense weights only where the level has no FFT: fewer than
                // 8 columns, odd or above 8192; 2 c^2 words would be 537 MB
                // at c = 4096)
                if (!fft_cols(c)) {
                    for (int ml = 0; ml < c; ++ml)
                        for (int k = 0; k < c; ++k) L.tab.push_back(w(ml, k, false));
                    for (int q = 0; q < c; ++q)
                        for (int ml = 0; ml < c; ++ml) L.tab.push_back(w(ml, q, true));
                }
                total += 3 * (size_t)n * c + L.tab.size();
            }
            lh[p].push_back(std::move(L));
        }
    }
    VV_HIP(hipMalloc(&r->mg_pool, total * sizeof(double)));
    std::vector<double> host(total, 0.0);
    size_t off = 0;
    r->mg.nprob = 2;
    for (int p = 0; p < 2; ++p) {
        const int n = nvec[p];
        MgProblemDev &P = r->mg.prob[p];
        P.n = n;
        P.nlev = nlev;
        std::memcpy(host.data() + off, sten_h[p].data(), 5 * (size_t)n * sizeof(double));
        P.sten = r->mg_pool + off;
        off += 5 * (size_t)n;
        std::memcpy(host.data() + off, rphi_h[p].data(), (size_t)n * sizeof(double));
        P.rphi = r->mg_pool + off;
        off += n;
        for (int l = 0; l < nlev; ++l) {
            MgLevelDev &L = P.lev[l];
            const LevHost &H = lh[p][l];
            L.ncol = H.ncol;
            L.cphi = H.cphi;
            L.nstep = H.S;
            L.direct = H.direct;
            L.tabsz = H.tabsz;
            L.u = r->mg_pool + off; off += (size_t)n * H.ncol;
            L.f = r->mg_pool + off; off += (size_t)n * H.ncol;
            L.u2 = nullptr;
            if (!H.direct) { L.u2 = r->mg_pool + off; off += (size_t)n * H.ncol; }
            L.dlt = nullptr;
            if (l == nlev - 2) { L.dlt = r->mg_pool + off; off += (size_t)n * (H.ncol / 2); }
            std::memcpy(host.data() + off, H.tab.data(), H.tab.size() * sizeof(double));
            L.lu = r->mg_pool + off;
            L.twid = H.direct ? L.lu + (size_t)(H.ncol / 2 + 1) * H.tabsz : nullptr;
            L.ftw = nullptr;
            off += H.tab.size();
            L.xm = nullptr;
            if (H.direct) { L.xm = r->mg_pool + off; off += (size_t)n * H.ncol; }
        }
    }
    VV_HIP(hipMemcpyAsync(r->mg_pool, host.data(), total * sizeof(double), hipMemcpyHostToDevice,
                          r->stream));
    std::vector<float> host32;
    if (r->cfg.precision == VV_PREC_MIXED && nlev > 1) {
        // the FP32 hierarchy of the inner cycle: stencil and row factors,
        // per smoothed level u32, u2_32, f32 (n c each), the black-line
        // changes of level nlev-2 and the PCR tables; the direct level's
        // FP32 copy u32 of its correction
        size_t t32 = 0;
        for (int p = 0; p < 2; ++p) {
            const int n = nvec[p];
            t32 += 6 * (size_t)n;
            for (int l = 0; l < nlev; ++l) {
                const LevHost &H = lh[p][l];
                if (H.direct) { t32 += (size_t)n * H.ncol; continue; }
                t32 += 3 * (size_t)n * H.ncol + H.tab.size();
                if (l == nlev - 2) t32 += (size_t)n * (H.ncol / 2);
            }
        }
        VV_HIP(hipMalloc(&r->mg_pool32, t32 * sizeof(float)));
        host32.assign(t32, 0.0f);
        size_t o = 0;
        auto put = [&](const double *src, size_t cnt) {
            for (size_t q = 0; q < cnt; ++q) host32[o + q] = (float)src[q];
            float *d = r->mg_pool32 + o;
            o += cnt;
            return d;
        };
        for (int p = 0; p < 2; ++p) {
            const int n = nvec[p];
            MgProblemDev &P = r->mg.prob[p];
            P.sten32 = put(sten_h[p].data(), 5 * (size_t)n);
            P.rphi32 = put(rphi_h[p].data(), (size_t)n);
            for (int l = 0; l < nlev; ++l) {
                MgLevelDev &L = P.lev[l];
                const LevHost &H = lh[p][l];
                L.u32 = r->mg_pool32 + o; o += (size_t)n * H.ncol;
                if (H.direct) continue;
                L.u2_32 = r->mg_pool32 + o; o += (size_t)n * H.ncol;
                L.f32 = r->mg_pool32 + o; o += (size_t)n * H.ncol;
                if (l == nlev - 2) { L.dlt32 = r->mg_pool32 + o; o += (size_t)n * (H.ncol / 2); }
                L.lu32 = put(H.tab.data(), H.tab.size());
            }
        }
        VV_HIP(hipMemcpyAsync(r->mg_pool32, host32.data(), t32 * sizeof(float), hipMemcpyHostToDevice,
                              r->stream));
    }
    {   // FFT twiddles of the direct level (both problems have its columns)
        const int C = r->mg.prob[0].lev[nlev - 1].ncol;
        std::vector<double> tw;
        if (fft_cols(C)) {
            // radix-2 twiddles of the n1 = 2^bits point sub-lines, then (C not
            // a power of two) the C twiddles of the mixed-radix combination
            const double pi = std::acos(-1.0);
            const int n1 = C & -C;
            for (int t = 0; t < n1 / 2; ++t) {
                tw.push_back(std::cos(2.0 * pi * t / n1));
                tw.push_back(-std::sin(2.0 * pi * t / n1));
            }
            if (n1 != C)
                for (int t = 0; t < C; ++t) {
                    tw.push_back(std::cos(2.0 * pi * t / C));
                    tw.push_back(-std::sin(2.0 * pi * t / C));
                }
            VV_HIP(hipMalloc(&r->mg_ftw, tw.size() * sizeof(double)));
            VV_HIP(hipMemcpyAsync(r->mg_ftw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice,
                                  r->stream));
            for (int p = 0; p < 2; ++p) r->mg.prob[p].lev[nlev - 1].ftw = r->mg_ftw;
        }
        VV_HIP(hipStreamSynchronize(r->stream));   // the host tables go out of scope
    }
    r->mg.done = r->mg_flags;
    r->mg.cycles = r->mg_flags + 2;
    r->mg.fnorm = r->mg_bits + 2;
    r->mg.relres = r->mg_relres;
    VV_HIP(hipMalloc(&r->mg_hist, 2 * MG_HIST * sizeof(double)));
    VV_HIP(hipMemsetAsync(r->mg_hist, 0, 2 * MG_HIST * sizeof(double), r->stream));
    r->mg.hist = r->mg_hist;
    r->mg.npart = r->npart;
    r->mg.fpart = r->fpart;
    r->mg.nbr = r->nbx;
    VV_HIP(hipStreamSynchronize(r->stream));
    r->nlev_v = r->nlev_u = nlev;
    return VV_OK;
}

static int zebra_threads(int n)
{
    int t = ((n + 63) / 64) * 64;
    return t > 1024 ? 1024 : t;
}

// descriptor of level L whose current iterate is buffer `cur` (0: u, 1: u2)
static KLev klev(const MgLevelDev &L, int cur)
{
    KLev k;
    k.u = cur ? L.u2 : L.u;
    k.uo = cur ? L.u : L.u2;
    k.f = L.f; k.lu = L.lu; k.twid = L.twid; k.xm = L.xm; k.dlt = L.dlt;
    k.ftw = reinterpret_cast<const double2 *>(L.ftw);
    k.cphi = L.cphi; k.ncol = L.ncol; k.nstep = L.nstep; k.tabsz = (int)L.tabsz;
    return k;
}

// kernel arguments of a launch on level lev (with its finer and coarser
// neighbours where they exist); cur[l] is level l's current buffer
static KArgs kargs(const vv_run *r, int lev, int mode, const int *cur)
{
    KArgs a;
    std::memset(&a, 0, sizeof(a));
    for (int p = 0; p < MG_MAXPROB; ++p) {
        const MgProblemDev &Q = r->mg.prob[p];
        KProb &P = a.P[p];
        P.sten = Q.sten;
        P.rphi = Q.rphi;
        P.n = Q.n;
        P.L = klev(Q.lev[lev], cur[lev]);
        if (lev > 0) P.F = klev(Q.lev[lev - 1], cur[lev - 1]);
        if (lev + 1 < Q.nlev) P.Cc = klev(Q.lev[lev + 1], cur[lev + 1]);
    }
    a.done = r->mg.done;
    a.cycles = r->mg.cycles;
    a.mode = mode;
    a.lev0 = (lev == 0);
    return a;
}

// The FP32 descriptors of the mixed-precision inner cycle: the smoothed
// levels' FP32 buffers and tables; the direct level appears only as the
// coarser neighbour whose FP32 copy u32 the prolongation reads.
static KLevT<float> klev32(const MgLevelDev &L, int cur)
{
    KLevT<float> k;
    std::memset(&k, 0, sizeof(k));
    k.u = cur ? L.u2_32 : L.u32;
    k.uo = cur ? L.u32 : L.u2_32;
    k.f = L.f32; k.lu = L.lu32; k.dlt = L.dlt32;
    k.cphi = L.cphi; k.ncol = L.ncol; k.nstep = L.nstep; k.tabsz = (int)L.tabsz;
    return k;
}

static KArgsT<float> kargs32(const vv_run *r, int lev, int mode, const int *cur)
{
    KArgsT<float> a;
    std::memset(&a, 0, sizeof(a));
    for (int p = 0; p < MG_MAXPROB; ++p) {
        const MgProblemDev &Q = r->mg.prob[p];
        KProbT<float> &P = a.P[p];
        P.sten = Q.sten32;
        P.rphi = Q.rphi32;
        P.n = Q.n;
        P.L = klev32(Q.lev[lev], cur[lev]);
        if (lev > 0) P.F = klev32(Q.lev[lev - 1], cur[lev - 1]);
        if (lev + 1 < Q.nlev) P.Cc = klev32(Q.lev[lev + 1], cur[lev + 1]);
    }
    a.done = r->mg.done;
    a.cycles = r->mg.cycles;
    a.mode = mode;
    a.lev0 = (lev == 0);
    return a;
}

// One V(nu1,nu2) cycle of both problems.  Levels 0..L-1 are smoothed by
// zebra sweeps (both colours per launch, buffer u -> uo, so a level's
// current buffer alternates), level L is solved exactly.  The restriction
// into a smoothed level is fused into its first sweep (ZB_RESTRICT), the
// prolongation of a correction into the first post-sweep of the finer level
// (ZB_PROLONG).  nu1 + nu2 is even, so every level is back in buffer 0 at
// the end of the cycle and all cycles enqueue the same launches.
static int enqueue_cycle(vv_run *r, int *check = nullptr)
{
    vv::MgBatchDev &B = r->mg;
    const int L = B.prob[0].nlev - 1;
    const int n = r->nr;
    const double ntot = (double)(B.prob[0].n + B.prob[1].n);
    const int zt = zebra_threads(n);
    const bool two = n > 1024;
    int cur[MG_MAXLEV] = {0};
    hipEvent_t pe;
    auto ncol = [&](int l) { return B.prob[0].lev[l].ncol; };
    // mixed precision: the smoothed levels run the FP32 correction equation
    const bool mixed = (r->cfg.precision == VV_PREC_MIXED) && L > 0;
    const double wb = mixed ? 4.0 : 8.0;   // bytes of a smoothed level's word
    auto sweep = [&](int l, int mode, int store_delta) {
        // algorithmic words per point N of the level: plain 2.5 (read f and
        // the old black lines, write u); restriction 6 (read the finer u and
        // f, write f and u); prolongation 3 (plain + the coarse correction
        // at the black lines)
        const double w = (mode == ZB_RESTRICT) ? 6.0 : (mode == ZB_PROLONG) ? 3.0 : 2.5;
        const double bytes = wb * w * ntot * ncol(l);
        prof_begin(r, &pe);
        // line groups per workgroup: one on small levels, enough to keep
        // about 512 workgroups per problem on large ones
        const int nb = ncol(l) / 2, G = two ? 1 : (nb + 511) / 512;
        const dim3 grid((unsigned)((nb + G - 1) / G), 2);
        if (mixed) {
            KArgsT<float> a = kargs32(r, l, mode, cur);
            a.store_delta = store_delta;
            if (two)
                hipLaunchKernelGGL((k_mg_zebra1<2, float>), grid, dim3(zt), 0, r->stream, a);
            else
                hipLaunchKernelGGL((k_mg_zebra<1, float>), grid, dim3(zt), 0, r->stream, a);
        } else {
            KArgs a = kargs(r, l, mode, cur);
            a.store_delta = store_delta;
            if (two)
                hipLaunchKernelGGL((k_mg_zebra1<2, double>), grid, dim3(zt), 0, r->stream, a);
            else
                hipLaunchKernelGGL((k_mg_zebra<1, double>), grid, dim3(zt), 0, r->stream, a);
        }
        prof_end(r, VV_K_ZEBRA, pe, bytes);
        cur[l] ^= 1;
    };
    MixPtr mp;
    std::memset(&mp, 0, sizeof(mp));
    const long long cnt0 = (long long)n * ncol(0);
    const int nbf = std::max(1, std::min(2048, nblk(cnt0, 256)));
    if (mixed) {
        // the defect of the FP64 iterate -> FP32 right-hand side, zero correction
        for (int p = 0; p < 2; ++p) {
            mp.f[p] = B.prob[p].lev[0].f32;
            mp.u[p] = B.prob[p].lev[0].u32;
        }
        prof_begin(r, &pe);
        hipLaunchKernelGGL(k_mg_defect, dim3(nbf, 2), dim3(256), 0, r->stream, kargs(r, 0, ZB_PLAIN, cur), mp);
        prof_end(r, VV_K_NORM, pe, (16.0 + 8.0) * ntot * ncol(0));
    }
    for (int l = 0; l < L; ++l)
        for (int s = 0; s < r->cfg.nu1; ++s)
            sweep(l, (l > 0 && s == 0) ? ZB_RESTRICT : ZB_PLAIN, (l == L - 1 && s == r->cfg.nu1 - 1) ? 1 : 0);
    {   // exact solve of the coarsest level; its right-hand side is the
        // restriction of the last pre-sweep's black-line changes
        const int C = ncol(L);
        const KArgs a = kargs(r, L, ZB_PLAIN, cur);
        if (L > 0) {
            prof_begin(r, &pe);
            const dim3 rg(nblk((long long)n * C, 256), 2);
            if (mixed)
                hipLaunchKernelGGL(k_mg_restrict<float>, rg, dim3(256), 0, r->stream, kargs32(r, L, ZB_PLAIN, cur),
                                   B.prob[0].lev[L].f, B.prob[1].lev[L].f);
            else
                hipLaunchKernelGGL(k_mg_restrict<double>, rg, dim3(256), 0, r->stream, a, B.prob[0].lev[L].f,
                                   B.prob[1].lev[L].f);
            prof_end(r, VV_K_RESTRICT, pe, (wb + 8.0) * ntot * C);
        }
        const int TM = C < DFT_TM ? C : DFT_TM;
        const dim3 gd(nblk(n, DFT_TN), 2, (C + TM - 1) / TM);
        // FFTs where the level has twiddles (even C, 8 <= C <= 8192), the
        // dense VALU transforms otherwise.  LP line pairs per workgroup, few
        // enough for about 256 workgroups (measured against LP = 2048/C:
        // 8243 -> 10027 steps/s at 129 x 65, 5386 -> 5991 at 513 x 257;
        // against the dense transforms 1902 -> 2896 steps/s at 1025 x 513;
        // matrix-core GEMMs lose to both, their k loop being latency-bound)
        const bool fft = B.prob[0].lev[L].ftw != nullptr;
        const int fbits = fft ? __builtin_ctz((unsigned)C) : 0;
        const int fodd = C >> fbits;              // odd factor of C (1: radix 2 only)
        const int LP = std::max(1, std::min(fft_pairs(C), n / 256));
        const dim3 fg((unsigned)((n + 2 * LP - 1) / (2 * LP)), 2);
        const size_t flds = (size_t)LP * C * sizeof(double2);
        // registers of the mixed-radix combination: LP C <= R 256
        const int fR = (fodd == 1) ? 0 : (LP * C <= 8 * 256) ? 8 : 32;
        if (fft && LP * C > 32 * 256) { set_error("FFT line longer than 8192 points"); return VV_EINVAL; }
        prof_begin(r, &pe);
        // projection (read the level, write the mode lines) and line solves
        // (read and write the mode lines)
        if (fft && fR == 0) hipLaunchKernelGGL(k_mg_fft_fwd<0>, fg, dim3(256), flds, r->stream, a, fbits, 1, LP);
        else if (fft && fR == 8) hipLaunchKernelGGL(k_mg_fft_fwd<8>, fg, dim3(256), flds, r->stream, a, fbits, fodd, LP);
        else if (fft) hipLaunchKernelGGL(k_mg_fft_fwd<32>, fg, dim3(256), flds, r->stream, a, fbits, fodd, LP);
        else hipLaunchKernelGGL(k_mg_dft<0>, gd, dim3(256), 0, r->stream, a);
        if (two)
            hipLaunchKernelGGL(k_mg_modes<2>, dim3(C / 2 + 1, 2), dim3(zt), 0, r->stream, a);
        else
            hipLaunchKernelGGL(k_mg_modes<1>, dim3(C / 2 + 1, 2), dim3(zt), 0, r->stream, a);
        prof_end(r, VV_K_DIRECT, pe, 32.0 * ntot * C);
        prof_begin(r, &pe);
        if (fft && fR == 0) hipLaunchKernelGGL(k_mg_fft_inv<0>, fg, dim3(256), flds, r->stream, a, fbits, 1, LP);
        else if (fft && fR == 8) hipLaunchKernelGGL(k_mg_fft_inv<8>, fg, dim3(256), flds, r->stream, a, fbits, fodd, LP);
        else if (fft) hipLaunchKernelGGL(k_mg_fft_inv<32>, fg, dim3(256), flds, r->stream, a, fbits, fodd, LP);
        else hipLaunchKernelGGL(k_mg_dft<1>, gd, dim3(256), 0, r->stream, a);
        if (mixed) {   // the correction's FP32 copy for the prolongation
            for (int p = 0; p < 2; ++p) {
                mp.src[p] = B.prob[p].lev[L].u;
                mp.u[p] = B.prob[p].lev[L].u32;
                mp.cnt[p] = (long long)B.prob[p].n * C;
            }
            hipLaunchKernelGGL(k_mg_to32, dim3(std::max(1, std::min(2048, nblk((long long)n * C, 256))), 2),
                               dim3(256), 0, r->stream, (const int *)B.done, mp);
        }
        prof_end(r, VV_K_PROLONG, pe, 24.0 * ntot * C);
    }
    for (int l = L - 1; l >= 0; --l)
        for (int s = 0; s < r->cfg.nu2; ++s) sweep(l, s == 0 ? ZB_PROLONG : ZB_PLAIN, 0);
    if (mixed) {   // u += e in FP64 (level 0 is back in its first buffer)
        for (int p = 0; p < 2; ++p) {
            mp.u[p] = B.prob[p].lev[0].u32;
            mp.dst[p] = B.prob[p].lev[0].u;
            mp.cnt[p] = (long long)B.prob[p].n * ncol(0);
        }
        prof_begin(r, &pe);
        hipLaunchKernelGGL(k_mg_correct, dim3(nbf, 2), dim3(256), 0, r->stream, (const int *)B.done, mp);
        prof_end(r, VV_K_NORM, pe, 20.0 * ntot * ncol(0));
    }
    const int nc0 = ncol(0);
    int nb = (int)((cnt0 + 255) / 256);    // one point per thread
    if (nb > MG_NORM_MAXB) nb = MG_NORM_MAXB;
    if (nb < 1) nb = 1;
    StopArgs sa;
    sa.done = B.done; sa.cycles = B.cycles; sa.fnorm = B.fnorm; sa.relres = B.relres; sa.hist = B.hist;
    sa.npart = B.npart; sa.fpart = B.fpart; sa.nbn = nb; sa.nbr = B.nbr;
    sa.maxc = r->cfg.max_cycles; sa.rtol = r->cfg.rtol;
    sa.check = check;
    prof_begin(r, &pe);
    hipLaunchKernelGGL(k_mg_norm, dim3(nb, 2), dim3(256), 0, r->stream, kargs(r, 0, ZB_PLAIN, cur), B.npart);
    hipLaunchKernelGGL(k_mg_stop, dim3(1), dim3(256), 0, r->stream, sa);
    prof_end(r, VV_K_NORM, pe, 16.0 * ntot * nc0);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

// The time loop's deferred solves, with no host synchronisation: one cycle
// of both problems, then a refinement cycle whose stop test carries the
// sticky check.  Every kernel of a cycle returns at once for a problem
// already done, so the second cycle costs empty launches when the first
// converged and refines on the device when it did not; only a problem still
// above rtol after both sets the check.
int mg_enqueue_deferred(vv_run *r)
{
    const int st = enqueue_cycle(r);
    if (st) return st;
    return enqueue_cycle(r, r->mg_flags + 4);
}

// Runs cycles until both active problems are done.  Cycles are queued in
// chunks sized from the previous solve, then the flags are read back.
int mg_solve(vv_run *r, int which, int *cycles, double *relres)
{
    int chunk = r->last_cycles[0] > r->last_cycles[1] ? r->last_cycles[0] : r->last_cycles[1];
    if (chunk < 1) chunk = 2;
    int queued = 0;
    struct { int flags[8]; unsigned long long bits[4]; double rel[2]; } h;
    static_assert(sizeof(h) == 80, "solver state block layout");
    const int *hflags = h.flags;
    const unsigned long long *hbits = h.bits;
    const double *hrel = h.rel;
    for (;;) {
        for (int c = 0; c < chunk; ++c) {
            int st = enqueue_cycle(r);
            if (st) return st;
        }
        queued += chunk;
        VV_HIP(hipMemcpyAsync(&h, r->mg_state, sizeof(h), hipMemcpyDeviceToHost, r->stream));
        VV_HIP(hipStreamSynchronize(r->stream));
        for (int p = 0; p < 2; ++p) {
            double fn;
            std::memcpy(&fn, &hbits[2 + p], sizeof(double));
            if (!std::isfinite(fn)) {
                set_error("non-finite Poisson right-hand side");
                return VV_ENONFINITE;
            }
        }
        if (hflags[0] && hflags[1]) break;
        if (queued >= r->cfg.max_cycles) break;
        chunk = 1;
    }
    int st = VV_OK;
    for (int p = 0; p < 2; ++p) {
        const bool active = (which == 5) || (which == 2 && p == 0) || (which == 3 && p == 1);
        if (!active) continue;
        r->last_cycles[p] = hflags[2 + p];
        r->last_relres[p] = hrel[p];
        if (!(hrel[p] <= r->cfg.rtol)) st = VV_NOT_CONVERGED;
    }
    if (cycles) {
        cycles[0] = hflags[2];
        cycles[1] = hflags[3];
    }
    if (relres) {
        relres[0] = hrel[0];
        relres[1] = hrel[1];
    }
    return st;
}

int launch_mg_scatter_out(vv_run *r, int which)
{
    const long long n = (long long)(r->nr + 2) * (r->np + 1);
    const int w = (which == 5) ? 3 : (which == 2 ? 1 : 2);
    hipEvent_t pe;
    prof_begin(r, &pe);
    hipLaunchKernelGGL(k_mg_out, dim3(nblk(n, 256)), dim3(256), 0, r->stream,
                       (const double *)r->mg.prob[0].lev[0].u, (const double *)r->mg.prob[1].lev[0].u,
                       r->vel, r->velp, r->uel, r->uelp, r->nr, r->np, r->ld, w);
    prof_end(r, VV_K_OUT, pe, 8.0 * 6.0 * (double)n);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

}  // namespace vv
