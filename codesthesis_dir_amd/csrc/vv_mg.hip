// vv_mg.hip -- multigrid Poisson solver for the velocity components.
//
// The converged discrete solution is fixed by the reference's residual
// operators (rb.f relax2 residual rb.f:2306-2357 for v, relax3 residual
// rb.f:2553-2587 for u): a periodic 3-point second difference along phi
// plus the r operator D_r, which for v carries the one-sided closure rows
// (-2, 5/3, -1/5)/dr^2 at both walls.  That closure is not preserved under
// r-coarsening (a point-smoother V-cycle with rediscretised coarse grids
// diverges on it), so the cycle here never coarsens r:
//   - smoother: zebra r-line Gauss-Seidel, each line solved exactly with
//     a banded LU factorised once per level on the host;
//   - coarsening: phi only, full weighting / linear interpolation, down to
//     a single column where the line solve is the exact coarse solve.
// Both velocity problems of a run are batched into every launch
// (blockIdx.y = problem).  Kernels return early once their problem has
// converged, so several cycles can be queued between host checks.

#include "vv_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace vv {

namespace {

constexpr int WG = 64;

__device__ __forceinline__ size_t mix(int j, int q, int n, int h)
{
    return (size_t)(q & 1) * n * h + (size_t)j * h + (q >> 1);
}

// residual f - (D_r u + cphi*rphi*(u_{q+1} - 2u + u_{q-1})) at (j, q)
__device__ __forceinline__ double resid_at(const MgProblemDev &P, const MgLevelDev &L, int j, int q)
{
    const int n = P.n, h = L.h, ncol = L.ncol;
    const double *u = L.u;
    const double *s = P.sten + 5 * j;
    double lu = 0.0;
#pragma unroll
    for (int o = -2; o <= 2; ++o) {
        const int jj = j + o;
        if (jj >= 0 && jj < n && s[o + 2] != 0.0) lu += s[o + 2] * u[mix(jj, q, n, h)];
    }
    const double uc = u[mix(j, q, n, h)];
    if (ncol > 1) {
        const int qm = (q == 0) ? ncol - 1 : q - 1;
        const int qp = (q == ncol - 1) ? 0 : q + 1;
        lu += L.cphi * P.rphi[j] * (u[mix(j, qp, n, h)] - 2.0 * uc + u[mix(j, qm, n, h)]);
    }
    return L.f[mix(j, q, n, h)] - lu;
}

__global__ void __launch_bounds__(WG) k_mg_zebra(const MgBatchDev *__restrict__ Bp, int lev, int color)
{
    const MgBatchDev &B = *Bp;
    const int p = blockIdx.y;
    if (p >= B.nprob || B.done[p]) return;
    const MgProblemDev &P = B.prob[p];
    if (lev >= P.nlev) return;
    const MgLevelDev &L = P.lev[lev];
    const int n = P.n, h = L.h, ncol = L.ncol;
    const int Q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool coarsest = (ncol == 1);
    if (coarsest) {
        if (color != 0 || Q != 0) return;
    } else if (Q >= h) {
        return;
    }
    const int q = coarsest ? 0 : 2 * Q + color;
    double *u = L.u;
    const double *f = L.f;
    const double *l1 = L.lu, *l2 = L.lu + n, *u1 = L.lu + 2 * n, *u2 = L.lu + 3 * n, *dinv = L.lu + 4 * n;
    const size_t self = (size_t)color * n * h + Q;
    size_t om = 0, op = 0;
    if (!coarsest) {
        const int qm = (q == 0) ? ncol - 1 : q - 1;
        const int qp = (q == ncol - 1) ? 0 : q + 1;
        om = (size_t)(qm & 1) * n * h + (qm >> 1);
        op = (size_t)(qp & 1) * n * h + (qp >> 1);
    }
    const double c = L.cphi;
    double y1 = 0.0, y2 = 0.0;
    for (int j = 0; j < n; ++j) {
        const size_t row = (size_t)j * h;
        double r = f[self + row];
        if (!coarsest) r -= c * P.rphi[j] * (u[om + row] + u[op + row]);
        const double y = r - l1[j] * y1 - l2[j] * y2;
        u[self + row] = y;
        y2 = y1;
        y1 = y;
    }
    double x1 = 0.0, x2 = 0.0;
    for (int j = n - 1; j >= 0; --j) {
        const size_t row = (size_t)j * h;
        const double x = (u[self + row] - u1[j] * x1 - u2[j] * x2) * dinv[j];
        u[self + row] = x;
        x2 = x1;
        x1 = x;
    }
}

// fine level lev -> coarse lev+1: full weighting of the residual along phi;
// the coarse correction starts from zero.
__global__ void __launch_bounds__(256) k_mg_restrict(const MgBatchDev *__restrict__ Bp, int lev)
{
    const MgBatchDev &B = *Bp;
    const int p = blockIdx.y;
    if (p >= B.nprob || B.done[p]) return;
    const MgProblemDev &P = B.prob[p];
    if (lev + 1 >= P.nlev) return;
    const MgLevelDev &F = P.lev[lev];
    const MgLevelDev &Cc = P.lev[lev + 1];
    const int n = P.n, nc = Cc.ncol;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * nc) return;
    const int j = (int)(t / nc), qc = (int)(t - (long long)j * nc);
    const int q0 = 2 * qc, q1 = 2 * qc + 1, q2 = (2 * qc + 2) % F.ncol;
    const double r0 = resid_at(P, F, j, q0);
    const double r1 = resid_at(P, F, j, q1);
    const double r2 = resid_at(P, F, j, q2);
    const size_t m = mix(j, qc, n, Cc.h);
    Cc.f[m] = 0.25 * r0 + 0.5 * r1 + 0.25 * r2;
    Cc.u[m] = 0.0;
}

// coarse lev+1 -> fine lev: linear interpolation along phi, added.
__global__ void __launch_bounds__(256) k_mg_prolong(const MgBatchDev *__restrict__ Bp, int lev)
{
    const MgBatchDev &B = *Bp;
    const int p = blockIdx.y;
    if (p >= B.nprob || B.done[p]) return;
    const MgProblemDev &P = B.prob[p];
    if (lev + 1 >= P.nlev) return;
    const MgLevelDev &F = P.lev[lev];
    const MgLevelDev &Cc = P.lev[lev + 1];
    const int n = P.n, nc = Cc.ncol;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * nc) return;
    const int j = (int)(t / nc), qc = (int)(t - (long long)j * nc);
    const int qcm = (qc == 0) ? nc - 1 : qc - 1;
    const double e = Cc.u[mix(j, qc, n, Cc.h)];
    const double em = Cc.u[mix(j, qcm, n, Cc.h)];
    F.u[mix(j, 2 * qc + 1, n, F.h)] += e;
    F.u[mix(j, 2 * qc, n, F.h)] += 0.5 * (e + em);
}

// max-norm of the level-0 residual
__global__ void __launch_bounds__(256) k_mg_norm(const MgBatchDev *__restrict__ Bp)
{
    const MgBatchDev &B = *Bp;
    const int p = blockIdx.y;
    if (p >= B.nprob || B.done[p]) return;
    const MgProblemDev &P = B.prob[p];
    const MgLevelDev &L = P.lev[0];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (t < (long long)P.n * L.ncol) {
        const int j = (int)(t / L.ncol), q = (int)(t - (long long)j * L.ncol);
        m = fabs(resid_at(P, L, j, q));
    }
    // wavefront reduction, then one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(m, off, 64);
        m = (o > m || o != o) ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(&B.res[p], (unsigned long long)__double_as_longlong(m));
}

__global__ void k_mg_check(const MgBatchDev *__restrict__ Bp, double rtol, int maxc)
{
    const MgBatchDev &B = *Bp;
    const int p = threadIdx.x;
    if (p >= B.nprob || B.done[p]) return;
    const double r = __longlong_as_double((long long)B.res[p]);
    const double fn = __longlong_as_double((long long)B.fnorm[p]);
    const double rel = r / (fn > 0.0 ? fn : 1.0);
    const int c = B.cycles[p] + 1;
    B.cycles[p] = c;
    B.relres[p] = rel;
    B.res[p] = 0ull;
    if (!(rel > rtol) || c >= maxc) B.done[p] = 1;    // NaN also stops
}

__global__ void k_mg_reset(const MgBatchDev *__restrict__ Bp, int which)
{
    const MgBatchDev &B = *Bp;
    const int p = threadIdx.x;
    if (p >= MG_MAXPROB) return;
    const bool active = (which == 5) || (which == 2 && p == 0) || (which == 3 && p == 1);
    B.done[p] = active ? 0 : 1;
    B.cycles[p] = 0;
    B.res[p] = 0ull;
    B.fnorm[p] = 0ull;
    B.relres[p] = 0.0;
}

// level 0 -> fields: psip <- psi, psi <- solution (rb.f:1894-1913)
__global__ void __launch_bounds__(256) k_mg_out(const MgBatchDev *__restrict__ Bp, double *vel, double *velp, double *uel,
                                                double *uelp, int nr, int np, int ld, int which)
{
    const MgBatchDev &B = *Bp;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int rows = nr + 2;
    if (t >= (long long)rows * (np + 1)) return;
    const int k = (int)(t / rows), i = (int)(t - (long long)k * rows);
    const size_t o = (size_t)k * ld + i;
    const int q = (k == 0) ? np - 1 : k - 1;
    const int h = np / 2 > 0 ? np / 2 : 1;
    if (which & 1) {
        const double old = vel[o];
        velp[o] = old;
        double nv = 0.0;
        if (i >= 1 && i <= nr) nv = B.prob[0].lev[0].u[mix(i - 1, q, nr, h)];
        vel[o] = nv;
    }
    if ((which & 2) && i <= nr) {
        const double old = uel[o];
        uelp[o] = old;
        double nu = 0.0;
        if (i >= 1 && i <= nr - 1) nu = B.prob[1].lev[0].u[mix(i - 1, q, nr - 1, h)];
        uel[o] = nu;
    }
}

inline int nblk(long long n, int wg) { return (int)((n + wg - 1) / wg); }

}  // namespace

// banded LU without pivoting (Doolittle, bandwidth 2)
BandLU make_band_lu(int n, const std::vector<double> &s2, const std::vector<double> &s1,
                    const std::vector<double> &d, const std::vector<double> &p1,
                    const std::vector<double> &p2)
{
    BandLU f;
    f.n = n;
    f.l1.assign(n, 0.0); f.l2.assign(n, 0.0); f.u1.assign(n, 0.0); f.u2.assign(n, 0.0);
    f.dinv.assign(n, 0.0);
    std::vector<double> dd(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double l2 = 0.0, l1 = 0.0;
        if (i >= 2) l2 = s2[i] / dd[i - 2];
        if (i >= 1) l1 = (s1[i] - (i >= 2 ? l2 * f.u1[i - 2] : 0.0)) / dd[i - 1];
        double di = d[i] - (i >= 1 ? l1 * f.u1[i - 1] : 0.0) - (i >= 2 ? l2 * f.u2[i - 2] : 0.0);
        f.u1[i] = (i + 1 < n) ? p1[i] - (i >= 1 ? l1 * f.u2[i - 1] : 0.0) : 0.0;
        f.u2[i] = (i + 2 < n) ? p2[i] : 0.0;
        f.l1[i] = l1;
        f.l2[i] = l2;
        dd[i] = di;
        f.dinv[i] = 1.0 / di;
    }
    return f;
}

int mg_build(vv_run *r)
{
    const int nr = r->nr, np = r->np;
    const double hhri = 1.0 / (r->dr * r->dr);
    int nlev = 0;
    for (int c = np; c >= 1; c /= 2) {
        ++nlev;
        if (c == 1) break;
    }
    if (nlev > MG_MAXLEV) { set_error("too many multigrid levels"); return VV_EINVAL; }
    // sizes: per problem per level u,f (n*ncol each) + lu (5n) ; sten+rphi (6n)
    const int nvec[2] = {nr, nr - 1};
    size_t total = 0;
    for (int p = 0; p < 2; ++p) {
        const int n = nvec[p];
        total += 6 * (size_t)n;
        for (int l = 0, c = np; l < nlev; ++l, c /= 2) total += 2 * (size_t)n * c + 5 * (size_t)n;
    }
    VV_HIP(hipMalloc(&r->mg_pool, total * sizeof(double)));
    VV_HIP(hipMemsetAsync(r->mg_pool, 0, total * sizeof(double), r->stream));
    std::vector<double> host(total, 0.0);
    size_t off = 0;
    r->mg.nprob = 2;
    for (int p = 0; p < 2; ++p) {
        const int n = nvec[p];
        MgProblemDev &P = r->mg.prob[p];
        P.n = n;
        P.nlev = nlev;
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
        double *sten_h = host.data() + off;
        for (int j = 0; j < n; ++j) {
            sten_h[5 * j + 0] = s2[j]; sten_h[5 * j + 1] = s1[j]; sten_h[5 * j + 2] = dg[j];
            sten_h[5 * j + 3] = p1[j]; sten_h[5 * j + 4] = p2[j];
        }
        P.sten = r->mg_pool + off;
        off += 5 * (size_t)n;
        for (int j = 0; j < n; ++j) host[off + j] = rphi[j];
        P.rphi = r->mg_pool + off;
        off += n;
        double hp = r->dp;
        for (int l = 0, c = np; l < nlev; ++l, c /= 2, hp *= 2.0) {
            MgLevelDev &L = P.lev[l];
            L.ncol = c;
            L.h = c / 2 > 0 ? c / 2 : 1;
            L.cphi = 1.0 / (hp * hp);
            L.u = r->mg_pool + off; off += (size_t)n * c;
            L.f = r->mg_pool + off; off += (size_t)n * c;
            std::vector<double> d2(dg);
            if (c > 1)
                for (int j = 0; j < n; ++j) d2[j] = dg[j] - 2.0 * L.cphi * rphi[j];
            BandLU lu = make_band_lu(n, s2, s1, d2, p1, p2);
            double *t = host.data() + off;
            for (int j = 0; j < n; ++j) {
                t[j] = lu.l1[j]; t[n + j] = lu.l2[j]; t[2 * n + j] = lu.u1[j];
                t[3 * n + j] = lu.u2[j]; t[4 * n + j] = lu.dinv[j];
            }
            L.lu = r->mg_pool + off;
            off += 5 * (size_t)n;
        }
    }
    VV_HIP(hipMemcpyAsync(r->mg_pool, host.data(), total * sizeof(double), hipMemcpyHostToDevice,
                          r->stream));
    r->mg.done = r->mg_flags;
    r->mg.cycles = r->mg_flags + 2;
    r->mg.res = r->mg_bits;
    r->mg.fnorm = r->mg_bits + 2;
    r->mg.relres = r->mg_relres;
    VV_HIP(hipMalloc(&r->mg_dev, sizeof(vv::MgBatchDev)));
    VV_HIP(hipMemcpyAsync(r->mg_dev, &r->mg, sizeof(vv::MgBatchDev), hipMemcpyHostToDevice, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    r->nlev_v = r->nlev_u = nlev;
    return VV_OK;
}

// algorithmic bytes of one zebra half-sweep of level l over both problems:
// read f and the other colour, write the own colour (the coarsest level:
// read f, write u)
static double zebra_bytes(const vv::MgBatchDev &B, int l)
{
    double b = 0.0;
    for (int p = 0; p < B.nprob; ++p) {
        const int n = B.prob[p].n, h = B.prob[p].lev[l].h, nc = B.prob[p].lev[l].ncol;
        b += (nc == 1) ? 16.0 * n : 24.0 * (double)n * h;
    }
    return b;
}

static int enqueue_cycle(vv_run *r)
{
    vv::MgBatchDev &B = r->mg;
    const vv::MgBatchDev *Bd = r->mg_dev;
    const int nlev = B.prob[0].nlev;
    const int n = r->nr;
    const int ntot = B.prob[0].n + B.prob[1].n;
    const double rtol = r->cfg.rtol;
    hipEvent_t pe;
    auto zebra = [&](int l, int c) {
        const int h = B.prob[0].lev[l].h;
        const bool coarse = B.prob[0].lev[l].ncol == 1;
        prof_begin(r, &pe);
        hipLaunchKernelGGL(k_mg_zebra, dim3(coarse ? 1 : nblk(h, WG), 2), dim3(WG), 0, r->stream, Bd, l, c);
        prof_end(r, VV_K_ZEBRA, pe, zebra_bytes(B, l));
    };
    for (int l = 0; l < nlev - 1; ++l) {
        for (int s = 0; s < r->cfg.nu1; ++s)
            for (int c = 0; c < 2; ++c) zebra(l, c);
        const int ncf = B.prob[0].lev[l].ncol, ncc = B.prob[0].lev[l + 1].ncol;
        const long long cnt = (long long)n * ncc;
        prof_begin(r, &pe);
        hipLaunchKernelGGL(k_mg_restrict, dim3(nblk(cnt, 256), 2), dim3(256), 0, r->stream, Bd, l);
        prof_end(r, VV_K_RESTRICT, pe, 8.0 * ntot * (2.0 * ncf + 2.0 * ncc));
    }
    zebra(nlev - 1, 0);
    for (int l = nlev - 2; l >= 0; --l) {
        const int ncf = B.prob[0].lev[l].ncol, ncc = B.prob[0].lev[l + 1].ncol;
        const long long cnt = (long long)n * ncc;
        prof_begin(r, &pe);
        hipLaunchKernelGGL(k_mg_prolong, dim3(nblk(cnt, 256), 2), dim3(256), 0, r->stream, Bd, l);
        prof_end(r, VV_K_PROLONG, pe, 8.0 * ntot * (1.0 * ncc + 2.0 * ncf));
        for (int s = 0; s < r->cfg.nu2; ++s)
            for (int c = 0; c < 2; ++c) zebra(l, c);
    }
    const int nc0 = B.prob[0].lev[0].ncol;
    const long long cnt0 = (long long)n * nc0;
    prof_begin(r, &pe);
    hipLaunchKernelGGL(k_mg_norm, dim3(nblk(cnt0, 256), 2), dim3(256), 0, r->stream, Bd);
    prof_end(r, VV_K_NORM, pe, 16.0 * ntot * nc0);
    prof_begin(r, &pe);
    hipLaunchKernelGGL(k_mg_check, dim3(1), dim3(64), 0, r->stream, Bd, rtol, r->cfg.max_cycles);
    prof_end(r, VV_K_CHECK, pe, 0.0);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int mg_reset(vv_run *r, int which)
{
    hipLaunchKernelGGL(k_mg_reset, dim3(1), dim3(64), 0, r->stream, r->mg_dev, which);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

// Runs cycles until both active problems are done.  Cycles are queued in
// chunks sized from the previous solve, then the flags are read back.
int mg_solve(vv_run *r, int which, int *cycles, double *relres)
{
    int chunk = r->last_cycles[0] > r->last_cycles[1] ? r->last_cycles[0] : r->last_cycles[1];
    if (chunk < 1) chunk = 2;
    int queued = 0;
    int hflags[4];
    unsigned long long hbits[4];
    double hrel[2];
    for (;;) {
        for (int c = 0; c < chunk; ++c) {
            int st = enqueue_cycle(r);
            if (st) return st;
        }
        queued += chunk;
        VV_HIP(hipMemcpyAsync(hflags, r->mg_flags, 4 * sizeof(int), hipMemcpyDeviceToHost, r->stream));
        VV_HIP(hipMemcpyAsync(hbits, r->mg_bits, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              r->stream));
        VV_HIP(hipMemcpyAsync(hrel, r->mg_relres, 2 * sizeof(double), hipMemcpyDeviceToHost, r->stream));
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
    hipLaunchKernelGGL(k_mg_out, dim3(nblk(n, 256)), dim3(256), 0, r->stream, r->mg_dev, r->vel, r->velp,
                       r->uel, r->uelp, r->nr, r->np, r->ld, w);
    prof_end(r, VV_K_OUT, pe, 8.0 * 6.0 * (double)n);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

}  // namespace vv
