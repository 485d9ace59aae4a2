// vvm.hip -- the meridional (r,z) plane with swirl on stretched nodes, on
// gfx950 (include/vvmerid.h).
//
// Explicit part: the compact advection derivatives as batched line solves
// (lines along r and along z staged in LDS with their row tables, one
// thread per line), with the row tables of the uneven nodes and their elimination
// factors formed once on the host; then one thread per node for the
// DuFort-Frankel step, the right-hand sides and the walls.  Every
// expression is evaluated in the operation order of the oracle (the library
// is compiled with -ffp-contract=off), so the explicit results are bit for
// bit the oracle's.
//
// Poisson: the operators L = D_r + D_z (three-point Fornberg rows on the
// uneven nodes; D_r carries -1/r^2 for u) are separable.  D_z is
// diagonalised once at create (through its symmetric similar form); a solve is a forward GEMM into its
// eigenbasis (FP64 matrix cores), one tridiagonal system along r per z
// eigenvalue (parallel cyclic reduction in LDS, one workgroup per system),
// and the inverse GEMM straight into the velocity field.  The two velocity
// solves of a step run batched in the same launches.  A whole step is
// captured once as a HIP graph and replayed.

#include "vv_eig.h"
#include "vv_gemm.h"
#include "vv_reduce.h"

#include "../../include/vvmerid.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace vvm {

namespace {

constexpr int WG = 256;
constexpr int MAXB = 1024;   // partials of the reductions

// compact tables of one axis of n intervals: the forward rows (a, b, c of
// the right-hand side, Thomas factor) 4 per node | the backward rows (beta,
// b', 1/b', 0) 4 per node | alpha of row 1, beta of row n-1, the end
// weights e0[5], en[5], 4 spare
inline size_t ctab_size(int n) { return 8 * (size_t)(n + 1) + 16; }
constexpr int LDS_WORDS = 8192;   // 64 KB of LDS per workgroup

struct MGeo {
    int nr, nz, ld;
    double dt, Ra, Re, Pr, RRP;
    const double *rn;          // r nodes
    const double *wr, *wz;     // 6 per node: first-derivative weights (i-1, i, i+1), then second
    const double *cr, *cz;     // compact tables
    double va0, va1, vb1, vb0; // wall vorticity -w_r at r0 (nodes 1, 2) and rmax (nr-2, nr-1)
    double ve0, ve1, vg1, vg0; // wall vorticity u_z at the bottom (1, 2) and the top (nz-2, nz-1)
    double t0, t1, t2;         // one-sided th_z at the bottom (nodes 0, 1, 2)
    double s2, s1, s0;         // and at the top (nz-2, nz-1, nz)
};

#define M2(a, i, j) (a)[(unsigned)((j) * g.ld + (i))]

inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// x / b for a positive or negative pivot b with rb = RN(1/b): Markstein's
// correction of q0 = x rb by the exact remainder, the correctly rounded
// quotient in three dependent operations (the IEEE division of the oracle).
__device__ __forceinline__ double qdiv(double x, double b, double rb)
{
    const double q0 = x * rb;
    const double t = __builtin_fma(b, q0, -x);
    return __builtin_fma(-t, rb, q0);
}

// ------------------------------------------------------- compact lines
struct DF3 {
    const double *f[3];   // w, v, th
    double *d[6];         // w_r, w_z, v_r, v_z, th_r, th_z
};

// One line of n intervals, nodes L[0..n] in LDS, in two passes with the
// matching half of the table in LDS (tab: 4 words per row, misc: the end
// weights).  The forward pass leaves the eliminated right-hand sides in
// L[1..n-1]; the backward pass the derivative.  A group's nodes are read
// ahead of the rows that overwrite them.
__device__ __forceinline__ void line_fwd(double *L, int n, const double *tab, const double *misc)
{
    const double *e0 = misc + 2, *en = misc + 7;
    const double al1 = misc[0], bn1 = misc[1];
    const double d0 = e0[0] * L[0] + e0[1] * L[1] + e0[2] * L[2] + e0[3] * L[3] + e0[4] * L[4];
    const double dn = en[0] * L[n - 4] + en[1] * L[n - 3] + en[2] * L[n - 2] + en[3] * L[n - 1] + en[4] * L[n];
    double fa = L[0], fb = L[1], rp = 0.0;
    // the table and the nodes share the LDS with the rows being written: a
    // group's reads are all issued before its first write
    auto row = [&](int ii, const double *q, double fc) {
        double r = q[0] * fa + q[1] * fb + q[2] * fc;
        if (ii == 1) r = r - al1 * d0;
        if (ii == n - 1) r = r - bn1 * dn;
        rp = (ii == 1) ? r : r - q[3] * rp;
        L[ii] = rp;
        fa = fb;
        fb = fc;
    };
    constexpr int U = 8;
    int i = 1;
    for (; i + U - 1 <= n - 1; i += U) {
        double nx[U], qq[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            nx[u] = L[i + 1 + u];
#pragma unroll
            for (int k = 0; k < 4; ++k) qq[u][k] = tab[4 * (i + u) + k];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) row(i + u, qq[u], nx[u]);
    }
    for (; i <= n - 1; ++i) {
        double qq[4];
        for (int k = 0; k < 4; ++k) qq[k] = tab[4 * i + k];
        row(i, qq, L[i + 1]);
    }
}

__device__ __forceinline__ void line_bwd(double *L, int n, const double *tab)
{
    double y = qdiv(L[n - 1], tab[4 * (n - 1) + 1], tab[4 * (n - 1) + 2]);
    L[n - 1] = y;
    constexpr int U = 8;
    int i = n - 2;
    for (; i - U + 1 >= 1; i -= U) {
        double rr[U], qq[U][3];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            rr[u] = L[i - u];
#pragma unroll
            for (int k = 0; k < 3; ++k) qq[u][k] = tab[4 * (i - u) + k];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            y = qdiv(rr[u] - qq[u][0] * y, qq[u][1], qq[u][2]);
            L[i - u] = y;
        }
    }
    for (; i >= 1; --i) {
        const double *q = tab + 4 * i;
        y = qdiv(L[i] - q[0] * y, q[1], q[2]);
        L[i] = y;
    }
}

// dst[0..n) <- src[0..n) through a 64-thread workgroup, 16 loads in flight
// per thread before their LDS stores
__device__ __forceinline__ void lds_fill(double *dst, const double *__restrict__ src, int n, int tid)
{
    constexpr int B = 16;
    for (int e0 = 0; e0 < n; e0 += B * 64) {
        double v[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int e = e0 + u * 64 + tid;
            if (e < n) v[u] = src[e];
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int e = e0 + u * 64 + tid;
            if (e < n) dst[e] = v[u];
        }
    }
}

// Workgroups [0, nbr) take nlr lines along r (rows j, contiguous), the rest
// nlz lines along z (columns i, loaded row by row and stored transposed);
// each line sits in LDS with an odd pitch, one thread runs it, and the
// derivative goes back the way the nodes came.
__global__ void __launch_bounds__(64) k_m_deriv(MGeo g, DF3 F, int nbr, int nlr, int nlz)
{
    extern __shared__ double sh[];
    const int tid = threadIdx.x;
    const bool rdir = (int)blockIdx.x < nbr;
    const int n = rdir ? g.nr : g.nz;
    const double *__restrict__ T = rdir ? g.cr : g.cz;
    const int tw = 4 * (n + 1), sp = (n + 1) | 1;
    double *tab = sh, *misc = sh + tw, *lines = sh + tw + 16;
    int q, o0, nlines;
    if (rdir) {
        const int tiles = (g.nz - 1 + nlr - 1) / nlr;
        q = blockIdx.x / tiles;
        o0 = 1 + (blockIdx.x - q * tiles) * nlr;     // first row j
        nlines = min(nlr, g.nz - o0);
    } else {
        const int b = blockIdx.x - nbr, tiles = (g.nr - 1 + nlz - 1) / nlz;
        q = b / tiles;
        o0 = 1 + (b - q * tiles) * nlz;              // first column i
        nlines = min(nlz, g.nr - o0);
    }
    const double *__restrict__ f = F.f[q];
    double *__restrict__ d = F.d[2 * q + (rdir ? 0 : 1)];
    // element e of the tile is (major, minor) = (e / mn, e % mn): r tiles
    // (line, node), z tiles (node, line); walked in steps of 64 without
    // divisions
    {
        const int mn = rdir ? n + 1 : nlines, total = nlines * (n + 1);
        const int q64 = 64 / mn, r64 = 64 - q64 * mn;
        int ma = tid / mn, mi = tid - ma * mn;
        for (int e0 = 0; e0 < total; e0 += 16 * 64) {
            double v[16];
            int li[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                li[u] = -1;
                if (e0 + u * 64 + tid < total) {
                    v[u] = rdir ? f[(unsigned)((o0 + ma) * g.ld + mi)] : f[(unsigned)(ma * g.ld + o0 + mi)];
                    li[u] = rdir ? ma * sp + mi : mi * sp + ma;
                }
                ma += q64;
                mi += r64;
                if (mi >= mn) { mi -= mn; ++ma; }
            }
#pragma unroll
            for (int u = 0; u < 16; ++u)
                if (li[u] >= 0) lines[li[u]] = v[u];
        }
    }
    lds_fill(tab, T, tw, tid);
    if (tid < 16) misc[tid] = T[2 * tw + tid];
    __syncthreads();
    if (tid < nlines) line_fwd(lines + tid * sp, n, tab, misc);
    __syncthreads();
    lds_fill(tab, T + tw, tw, tid);
    __syncthreads();
    if (tid < nlines) line_bwd(lines + tid * sp, n, tab);
    __syncthreads();
    // interior nodes back: r tiles (line, node 1..n-1), z tiles (node
    // 1..n-1, line), walked as the loads
    const int mn = rdir ? n - 1 : nlines, tot = nlines * (n - 1);
    const int q64 = 64 / mn, r64 = 64 - q64 * mn;
    int ma = tid / mn, mi = tid - ma * mn;
    for (int e = tid; e < tot; e += 64) {
        if (rdir) d[(unsigned)((o0 + ma) * g.ld + 1 + mi)] = lines[ma * sp + 1 + mi];
        else d[(unsigned)((1 + ma) * g.ld + o0 + mi)] = lines[mi * sp + 1 + ma];
        ma += q64;
        mi += r64;
        if (mi >= mn) { mi -= mn; ++ma; }
    }
}

// ------------------------------------------------------- explicit step
struct EF {
    const double *w, *wp, *v, *vp, *th, *thp, *u, *ww;
    const double *d[6];
    double *wn, *vn, *thn;
};

// rotann2.f:281-306 (vorticity), 342-345 (velocity), 370-385 (temperature)
// in velocity form: f_new = (AA f_p + OFF(f) + sources - advection) / A with
// A, AA = 1/(2 P dt) -/+ D/2, D the centre coefficient of the diffusion
// operator, OFF the rest of it.  Wall nodes: wn takes the current w.
__global__ void __launch_bounds__(WG) k_m_explicit(MGeo g, EF F)
{
    const int t = blockIdx.x * WG + threadIdx.x;
    const int ni = g.nr + 1;
    if (t >= ni * (g.nz + 1)) return;
    const int j = t / ni, i = t - j * ni;
    if (i == 0 || i == g.nr || j == 0 || j == g.nz) {
        F.wn[t] = F.w[t];
        return;
    }
    const double *r1 = g.wr + 6 * i, *r2 = r1 + 3, *z1 = g.wz + 6 * j, *z2 = z1 + 3;
    const double r = g.rn[i];
    const double dt = g.dt, Pr = g.Pr, Ra = g.Ra, Re = g.Re, RRP = g.RRP;
    const double uu = F.u[t], wv = F.ww[t];
    const unsigned ld = (unsigned)g.ld;
#define OFF(f) (r2[0] * f[t - 1] + r2[2] * f[t + 1] + z2[0] * f[t - ld] + z2[2] * f[t + ld] + \
                (r1[0] * f[t - 1] + r1[1] * f[t] + r1[2] * f[t + 1]) / r)
    const double Dc = r2[1] + z2[1];
    const double *v = F.v, *th = F.th, *w = F.w;
    const double vz = z1[0] * v[t - ld] + z1[1] * v[t] + z1[2] * v[t + ld];
    const double thz = z1[0] * th[t - ld] + z1[1] * th[t] + z1[2] * th[t + ld];
    const double vv = v[t], tt = th[t];
    {
        const double D = Dc - 1.0 / (r * r);
        const double A = 1.0 / (2.0 * Pr * dt) - D / 2.0, AA = 1.0 / (2.0 * Pr * dt) + D / 2.0;
        const double adv = (uu * F.d[0][t] + wv * F.d[1][t] - uu * w[t] / r) / Pr;
        const double src = 2.0 / Pr * vv * vz / r + 2.0 * Re * vz - Ra * r * thz - 2.0 * RRP * (vv * thz + tt * vz);
        F.wn[t] = (AA * F.wp[t] + OFF(w) + src - adv) / A;
    }
    {
        const double D = Dc - 1.0 / (r * r);
        const double A = 1.0 / (2.0 * Pr * dt) - D / 2.0, AA = 1.0 / (2.0 * Pr * dt) + D / 2.0;
        const double adv = (uu * F.d[2][t] + wv * F.d[3][t] + uu * vv / r) / Pr;
        const double src = -2.0 * Re * uu + 2.0 * RRP * tt * uu;
        F.vn[t] = (AA * F.vp[t] + OFF(v) + src - adv) / A;
    }
    {
        const double A = 1.0 / (2.0 * dt) - Dc / 2.0, AA = 1.0 / (2.0 * dt) + Dc / 2.0;
        const double adv = uu * F.d[4][t] + wv * F.d[5][t];
        F.thn[t] = (AA * F.thp[t] + OFF(th) - adv) / A;
    }
#undef OFF
}

// fu = w_z, fw = -w_r - w/r at the inner nodes, from wn
__global__ void __launch_bounds__(WG) k_m_rhs(MGeo g, const double *__restrict__ w, double *fu, double *fw)
{
    const int t = blockIdx.x * WG + threadIdx.x;
    const int ni = g.nr + 1;
    if (t >= ni * (g.nz + 1)) return;
    const int j = t / ni, i = t - j * ni;
    if (i == 0 || i == g.nr || j == 0 || j == g.nz) return;
    const double *r1 = g.wr + 6 * i, *z1 = g.wz + 6 * j;
    const unsigned ld = (unsigned)g.ld;
    const double wz = z1[0] * w[t - ld] + z1[1] * w[t] + z1[2] * w[t + ld];
    const double wr = r1[0] * w[t - 1] + r1[1] * w[t] + r1[2] * w[t + 1];
    fu[t] = wz;
    fw[t] = -wr - w[t] / g.rn[i];
}

// ------------------------------------------------------- boundaries
struct UF {
    double *w, *wp, *wn, *v, *vp, *vn, *th, *thp, *thn;
    const double *u, *ww;
};

// rotann2.f:497-540 in velocity form: the wall vorticity from the new
// velocities (-w_r at the cylinders, u_z at the lids, zero at the corners),
// v = 0 on the walls, th = 0 / 1 at the inner / outer cylinder, insulated
// lids; then wp <- w <- wn (and likewise v, th), one thread per node.
__global__ void __launch_bounds__(WG) k_m_update(MGeo g, UF F)
{
    const int t = blockIdx.x * WG + threadIdx.x;
    const int ni = g.nr + 1, nr = g.nr, nz = g.nz;
    if (t >= ni * (nz + 1)) return;
    const int j = t / ni, i = t - j * ni;
    double wv, vv, tv;
    if (i == 0 || i == nr || j == 0 || j == nz) {
        const double *ww = F.ww, *u = F.u, *thn = F.thn;
        if (j >= 1 && j <= nz - 1)
            wv = (i == 0) ? -(g.va0 * M2(ww, 1, j) + g.va1 * M2(ww, 2, j))
                          : -(g.vb1 * M2(ww, nr - 2, j) + g.vb0 * M2(ww, nr - 1, j));
        else if (i == 0 || i == nr)
            wv = 0.0;
        else if (j == 0)
            wv = g.ve0 * M2(u, i, 1) + g.ve1 * M2(u, i, 2);
        else
            wv = g.vg1 * M2(u, i, nz - 2) + g.vg0 * M2(u, i, nz - 1);
        vv = 0.0;
        if (i == 0)
            tv = 0.0;
        else if (i == nr)
            tv = 1.0;
        else if (j == 0)
            tv = -(g.t1 * M2(thn, i, 1) + g.t2 * M2(thn, i, 2)) / g.t0;
        else
            tv = -(g.s1 * M2(thn, i, nz - 1) + g.s2 * M2(thn, i, nz - 2)) / g.s0;
        F.wn[t] = wv;
        F.vn[t] = vv;
        F.thn[t] = tv;
    } else {
        wv = F.wn[t];
        vv = F.vn[t];
        tv = F.thn[t];
    }
    F.wp[t] = F.w[t];
    F.w[t] = wv;
    F.vp[t] = F.v[t];
    F.v[t] = vv;
    F.thp[t] = F.th[t];
    F.th[t] = tv;
}

// ------------------------------------------------------- heat, energy, residual
// rotann2.f:468-495 in velocity form: the one-sided three-point dth/dr at
// each cylinder, trapezoid-summed over z in the order j = 0..nz (one
// thread sums what the workgroup formed).
__global__ void __launch_bounds__(WG) k_m_heat(MGeo g, const double *__restrict__ th, const double *__restrict__ hz,
                                               const double *__restrict__ hw, double *out)
{
    extern __shared__ double sh[];
    const int nz = g.nz, nr = g.nr;
    for (int j = threadIdx.x; j <= nz; j += WG) {
        sh[j] = hw[0] * M2(th, 0, j) + hw[1] * M2(th, 1, j) + hw[2] * M2(th, 2, j);
        sh[nz + 1 + j] = hw[3] * M2(th, nr - 2, j) + hw[4] * M2(th, nr - 1, j) + hw[5] * M2(th, nr, j);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int j = 0; j <= nz; ++j) {
            s0 = s0 + hz[j] * sh[j];
            s1 = s1 + hz[j] * sh[nz + 1 + j];
        }
        out[0] = s0;
        out[1] = s1;
    }
}

__global__ void __launch_bounds__(WG) k_m_energy(MGeo g, const double *u, const double *ww, const double *v,
                                                 const double *hz, const double *hr, double *part)
{
    __shared__ double sh[16];
    const int ni = g.nr + 1, n = ni * (g.nz + 1);
    double s = 0.0;
    for (int t = blockIdx.x * WG + threadIdx.x; t < n; t += gridDim.x * WG) {
        const int j = t / ni, i = t - j * ni;
        s += hz[j] * hr[i] * g.rn[i] * (u[t] * u[t] + ww[t] * ww[t] + v[t] * v[t]);
    }
    s = vv::block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = 0.5 * s;
}

// max |L psi - f| and max |f| over the inner nodes, per workgroup
__global__ void __launch_bounds__(WG) k_m_resid(MGeo g, int which, const double *__restrict__ psi,
                                                const double *__restrict__ f, double *part)
{
    __shared__ double sh[16];
    const int ni = g.nr + 1, n = ni * (g.nz + 1);
    double er = 0.0, fm = 0.0;
    const unsigned ld = (unsigned)g.ld;
    for (int t = blockIdx.x * WG + threadIdx.x; t < n; t += gridDim.x * WG) {
        const int j = t / ni, i = t - j * ni;
        if (i == 0 || i == g.nr || j == 0 || j == g.nz) continue;
        const double *r1 = g.wr + 6 * i, *r2 = r1 + 3, *z2 = g.wz + 6 * j + 3;
        const double r = g.rn[i];
        const double fm1 = (i > 1) ? psi[t - 1] : 0.0, fp1 = (i < g.nr - 1) ? psi[t + 1] : 0.0;
        const double gm = (j > 1) ? psi[t - ld] : 0.0, gp = (j < g.nz - 1) ? psi[t + ld] : 0.0;
        const double f0 = psi[t];
        double L = r2[0] * fm1 + r2[1] * f0 + r2[2] * fp1 + (r1[0] * fm1 + r1[1] * f0 + r1[2] * fp1) / r +
                   z2[0] * gm + z2[1] * f0 + z2[2] * gp;
        if (which == 1) L = L - f0 / (r * r);
        er = vv::nanmax(er, fabs(L - f[t]));
        fm = vv::nanmax(fm, fabs(f[t]));
    }
    er = vv::block_max(er, sh);
    fm = vv::block_max(fm, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = er;
        part[2 * blockIdx.x + 1] = fm;
    }
}

// ------------------------------------------------------- Poisson
// G[q][l][i] = sum_j Vinv[l][j] f_q(i+1, j+1)
struct ZFwd {
    const double *Vi;
    int M, N, K;
    const double *f0, *f1;
    double *G;
    int ld;
    __device__ double a(int, int m, int k) const { return Vi[(size_t)m * K + k]; }
    __device__ double b(int q, int k, int n) const { return (q ? f1 : f0)[(size_t)(k + 1) * ld + (n + 1)]; }
    __device__ void c(int q, int m, int n, double v) const { G[((size_t)q * M + m) * N + n] = v; }
};
// psi_q(i+1, j+1) = sum_l V[j][l] X[q][l][i]
struct ZInv {
    const double *V;
    int M, N, K;
    const double *X;
    double *p0, *p1;
    int ld;
    __device__ double a(int, int m, int k) const { return V[(size_t)m * K + k]; }
    __device__ double b(int q, int k, int n) const { return X[((size_t)q * K + k) * N + n]; }
    __device__ void c(int q, int m, int n, double v) const { (q ? p1 : p0)[(size_t)(m + 1) * ld + (n + 1)] = v; }
};

struct RS {
    const double *rt[2];   // per system kind: sub | diagonal | super, N each
    double *psi[2];        // their fields (walls cleared by mode 0's workgroup)
};

// (D_r + lam_l) x = G[q][l] along r by parallel cyclic reduction: each of
// the ceil(log2 N) levels eliminates the neighbours at distance s from every
// row at once, and after the last every row stands alone.  One workgroup per
// (l, q); the rows double-buffered in LDS.
__global__ void __launch_bounds__(WG) k_m_rsolve(MGeo g, RS S, const double *__restrict__ lam,
                                                 const double *__restrict__ G, double *__restrict__ X, int M, int N)
{
    extern __shared__ double sh[];
    const int l = blockIdx.x, q = blockIdx.y;
    double *a = sh, *b = sh + N, *c = sh + 2 * N, *d = sh + 3 * N;
    double *a2 = sh + 4 * N, *b2 = sh + 5 * N, *c2 = sh + 6 * N, *d2 = sh + 7 * N;
    const double *rt = S.rt[q];
    const double lm = lam[l];
    const double *Gq = G + ((size_t)q * M + l) * N;
    for (int i = threadIdx.x; i < N; i += WG) {
        a[i] = (i > 0) ? rt[i] : 0.0;
        b[i] = rt[N + i] + lm;
        c[i] = (i < N - 1) ? rt[2 * N + i] : 0.0;
        d[i] = Gq[i];
    }
    if (l == 0) {   // walls of the field
        double *p = S.psi[q];
        const int ni = g.nr + 1;
        for (int e = threadIdx.x; e < 2 * ni + 2 * (g.nz - 1); e += WG) {
            if (e < ni) p[e] = 0.0;
            else if (e < 2 * ni) p[(size_t)g.nz * g.ld + (e - ni)] = 0.0;
            else {
                const int k = e - 2 * ni, jj = 1 + (k >> 1);
                p[(size_t)jj * g.ld + ((k & 1) ? g.nr : 0)] = 0.0;
            }
        }
    }
    __syncthreads();
    for (int s = 1; s < N; s *= 2) {
        for (int i = threadIdx.x; i < N; i += WG) {
            const int im = i - s, ip = i + s;
            const double k1 = (im >= 0) ? a[i] / b[im] : 0.0;
            const double k2 = (ip < N) ? c[i] / b[ip] : 0.0;
            a2[i] = (im >= 0) ? -a[im] * k1 : 0.0;
            c2[i] = (ip < N) ? -c[ip] * k2 : 0.0;
            b2[i] = b[i] - ((im >= 0) ? c[im] * k1 : 0.0) - ((ip < N) ? a[ip] * k2 : 0.0);
            d2[i] = d[i] - ((im >= 0) ? d[im] * k1 : 0.0) - ((ip < N) ? d[ip] * k2 : 0.0);
        }
        __syncthreads();
        double *t;
        t = a; a = a2; a2 = t;
        t = b; b = b2; b2 = t;
        t = c; c = c2; c2 = t;
        t = d; d = d2; d2 = t;
    }
    double *Xq = X + ((size_t)q * M + l) * N;
    for (int i = threadIdx.x; i < N; i += WG) Xq[i] = d[i] / b[i];
}

// ------------------------------------------------------- host tables
// wfdif.f:59-85 (Fornberg): c[k*n + q], the weight of node q of x[0..n-1]
// for the k-th derivative at xx, k = 0..m
void fornberg(const double *x, int n, double xx, int m, double *c)
{
    std::memset(c, 0, sizeof(double) * (size_t)(m + 1) * n);
    c[0] = 1.0;
    double c1 = 1.0, c4 = x[0] - xx;
    for (int i = 1; i < n; ++i) {
        const int mn = std::min(i, m);
        double c2 = 1.0;
        const double c5 = c4;
        c4 = x[i] - xx;
        for (int j = 0; j < i; ++j) {
            const double c3 = x[i] - x[j];
            c2 = c2 * c3;
            if (j == i - 1) {
                for (int k = mn; k >= 1; --k) c[k * n + i] = c1 * (k * c[(k - 1) * n + i - 1] - c5 * c[k * n + i - 1]) / c2;
                c[i] = -c1 * c5 * c[i - 1] / c2;
            }
            for (int k = mn; k >= 1; --k) c[k * n + j] = (c4 * c[k * n + j] - k * c[(k - 1) * n + j]) / c3;
            c[j] = c4 * c[j] / c3;
        }
        c1 = c2;
    }
}

// The compact row at inner node i of x: alpha f'(i-1) + f'(i) + beta
// f'(i+1) = a f(i-1) + b f(i) + c f(i+1), fourth order on the uneven
// spacing (the Taylor conditions of orders 0..4 by Gaussian elimination
// with partial pivoting); out = (alpha, beta, a, b, c).
void compact_row(const double *x, int i, double *out)
{
    const double hm = x[i] - x[i - 1], hp = x[i + 1] - x[i];
    double M[5][6] = {
        {0.0, 0.0, 1.0, 1.0, 1.0, 0.0},
        {-1.0, -1.0, -hm, 0.0, hp, 1.0},
        {hm, -hp, hm * hm / 2.0, 0.0, hp * hp / 2.0, 0.0},
        {-hm * hm / 2.0, -hp * hp / 2.0, -hm * hm * hm / 6.0, 0.0, hp * hp * hp / 6.0, 0.0},
        {hm * hm * hm / 6.0, -hp * hp * hp / 6.0, hm * hm * hm * hm / 24.0, 0.0, hp * hp * hp * hp / 24.0, 0.0}};
    for (int col = 0; col < 5; ++col) {
        int piv = col;
        for (int r = col + 1; r < 5; ++r)
            if (std::fabs(M[r][col]) > std::fabs(M[piv][col])) piv = r;
        if (piv != col)
            for (int q = 0; q < 6; ++q) std::swap(M[col][q], M[piv][q]);
        for (int r = col + 1; r < 5; ++r) {
            const double f = M[r][col] / M[col][col];
            for (int q = col; q < 6; ++q) M[r][q] = M[r][q] - f * M[col][q];
        }
    }
    for (int r = 4; r >= 0; --r) {
        double s = M[r][5];
        for (int q = r + 1; q < 5; ++q) s = s - M[r][q] * out[q];
        out[r] = s / M[r][r];
    }
}

// the tables of a line of n intervals (layout of ctab_size): the compact
// rows, the Thomas elimination of the rows alpha_i, 1, beta_i (i = 1..n-1)
// and the five-point one-sided end derivatives
void compact_tables(int n, const double *x, double *T)
{
    std::fill(T, T + ctab_size(n), 0.0);
    std::vector<double> R(5 * (size_t)(n + 1), 0.0), fact(n + 1, 0.0), bp(n + 1, 0.0);
    for (int i = 1; i <= n - 1; ++i) compact_row(x, i, &R[5 * i]);
    bp[1] = 1.0;
    for (int i = 2; i <= n - 1; ++i) {
        fact[i] = R[5 * i + 0] / bp[i - 1];
        bp[i] = 1.0 - fact[i] * R[5 * (i - 1) + 1];
    }
    double *fw = T, *bw = T + 4 * (n + 1), *misc = T + 8 * (n + 1);
    for (int i = 1; i <= n - 1; ++i) {
        fw[4 * i + 0] = R[5 * i + 2];
        fw[4 * i + 1] = R[5 * i + 3];
        fw[4 * i + 2] = R[5 * i + 4];
        fw[4 * i + 3] = fact[i];
        bw[4 * i + 0] = R[5 * i + 1];
        bw[4 * i + 1] = bp[i];
        bw[4 * i + 2] = 1.0 / bp[i];
    }
    misc[0] = R[5 * 1 + 0];
    misc[1] = R[5 * (n - 1) + 1];
    double c[10];
    fornberg(x, 5, x[0], 1, c);
    for (int q = 0; q < 5; ++q) misc[2 + q] = c[5 + q];
    fornberg(x + n - 4, 5, x[n], 1, c);
    for (int q = 0; q < 5; ++q) misc[7 + q] = c[5 + q];
}

// three-point weights at the inner nodes: first derivative, then second
void three_point(int n, const double *x, std::vector<double> &w)
{
    w.assign(6 * (size_t)(n + 1), 0.0);
    double c[9];
    for (int i = 1; i <= n - 1; ++i) {
        fornberg(x + i - 1, 3, x[i], 2, c);
        for (int q = 0; q < 3; ++q) {
            w[6 * i + q] = c[3 + q];
            w[6 * i + 3 + q] = c[6 + q];
        }
    }
}

}  // namespace
}  // namespace vvm

using namespace vvm;

struct vvm_run {
    std::recursive_mutex mu;
    vvm_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    int nr = 0, nz = 0, ld = 0;
    size_t nf = 0;
    double *f[VVM_NFIELD] = {nullptr};
    double *dv[6] = {nullptr};
    double *geo = nullptr;   // r nodes | wr | wz | cr | cz | hz | hr | heat weights
    MGeo g;
    const double *hz = nullptr, *hr = nullptr, *hw = nullptr;
    int nlr = 1, nlz = 1, nbr = 0, nbz = 0;
    size_t dlds = 0;
    // Poisson
    double *V = nullptr, *Vi = nullptr, *lam = nullptr, *rt = nullptr;   // rt: [2][3][nr-1]
    double *G = nullptr, *X = nullptr;                                   // [2][nz-1][nr-1]
    double *part = nullptr;                                              // reductions, then results
    // graphs of one and of ten steps
    bool use_graphs = true;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    // profiling
    bool prof = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    long long prof_n[VVM_K_COUNT] = {0};
    double prof_ms[VVM_K_COUNT] = {0}, prof_work[VVM_K_COUNT] = {0};
};

namespace vvm {
namespace {

struct Prof {
    vvm_run *r;
    int id;
    double work;
    Prof(vvm_run *rr, int k, double w) : r(rr), id(k), work(w)
    {
        if (r->prof) (void)hipEventRecord(r->ev[0], r->stream);
    }
    ~Prof()
    {
        if (!r->prof) return;
        (void)hipEventRecord(r->ev[1], r->stream);
        (void)hipEventSynchronize(r->ev[1]);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r->ev[0], r->ev[1]);
        r->prof_n[id] += 1;
        r->prof_ms[id] += ms;
        r->prof_work[id] += work;
    }
};

double npts(const vvm_run *r) { return (double)(r->nr + 1) * (r->nz + 1); }

int launch_derivatives(vvm_run *r)
{
    DF3 F;
    F.f[0] = r->f[VVM_W];
    F.f[1] = r->f[VVM_V];
    F.f[2] = r->f[VVM_TH];
    for (int d = 0; d < 6; ++d) F.d[d] = r->dv[d];
    Prof P(r, VVM_K_DERIV, 9.0 * 8.0 * npts(r));
    hipLaunchKernelGGL(k_m_deriv, dim3(r->nbr + r->nbz), dim3(64), r->dlds, r->stream, r->g, F, r->nbr, r->nlr,
                       r->nlz);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_explicit(vvm_run *r)
{
    EF F;
    F.w = r->f[VVM_W]; F.wp = r->f[VVM_WP]; F.v = r->f[VVM_V]; F.vp = r->f[VVM_VP];
    F.th = r->f[VVM_TH]; F.thp = r->f[VVM_THP]; F.u = r->f[VVM_U]; F.ww = r->f[VVM_WW];
    for (int d = 0; d < 6; ++d) F.d[d] = r->dv[d];
    F.wn = r->f[VVM_WN]; F.vn = r->f[VVM_VN]; F.thn = r->f[VVM_THN];
    Prof P(r, VVM_K_EXPLICIT, 17.0 * 8.0 * npts(r));
    hipLaunchKernelGGL(k_m_explicit, dim3(nblk((long long)r->nf, WG)), dim3(WG), 0, r->stream, r->g, F);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_rhs(vvm_run *r)
{
    Prof P(r, VVM_K_RHS, 3.0 * 8.0 * npts(r));
    hipLaunchKernelGGL(k_m_rhs, dim3(nblk((long long)r->nf, WG)), dim3(WG), 0, r->stream, r->g,
                       (const double *)r->f[VVM_WN], r->f[VVM_FU], r->f[VVM_FW]);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

// which = 1, 2, or 3 for both solves batched
int launch_poisson(vvm_run *r, int which)
{
    const int M = r->nz - 1, N = r->nr - 1;
    const int batch = (which == 3) ? 2 : 1;
    const int k0 = (which == 2) ? 1 : 0;   // the first system kind
    double *src[2] = {r->f[VVM_FU], r->f[VVM_FW]};
    double *dst[2] = {r->f[VVM_U], r->f[VVM_WW]};
    const double flops = 2.0 * M * (double)N * M * batch;
    {
        ZFwd op{r->Vi, M, N, M, src[k0], src[1], r->G, r->ld};
        Prof P(r, VVM_K_ZFWD, flops);
        vv::gemm_f64_ksplit<true, false>(op, batch, r->stream);
        VV_HIP(hipGetLastError());
    }
    {
        RS S;
        S.rt[0] = r->rt + (size_t)k0 * 3 * N;
        S.rt[1] = r->rt + (size_t)3 * N;
        S.psi[0] = dst[k0];
        S.psi[1] = dst[1];
        Prof P(r, VVM_K_RSOLVE, 2.0 * 8.0 * M * (double)N * batch);
        hipLaunchKernelGGL(k_m_rsolve, dim3(M, batch), dim3(WG), 8 * sizeof(double) * (size_t)N, r->stream, r->g,
                           S, (const double *)r->lam, (const double *)r->G, r->X, M, N);
        VV_HIP(hipGetLastError());
    }
    {
        ZInv op{r->V, M, N, M, r->X, dst[k0], dst[1], r->ld};
        Prof P(r, VVM_K_ZINV, flops);
        vv::gemm_f64_ksplit<true, false>(op, batch, r->stream);
        VV_HIP(hipGetLastError());
    }
    return VV_OK;
}

int launch_update(vvm_run *r)
{
    UF F;
    F.w = r->f[VVM_W]; F.wp = r->f[VVM_WP]; F.wn = r->f[VVM_WN];
    F.v = r->f[VVM_V]; F.vp = r->f[VVM_VP]; F.vn = r->f[VVM_VN];
    F.th = r->f[VVM_TH]; F.thp = r->f[VVM_THP]; F.thn = r->f[VVM_THN];
    F.u = r->f[VVM_U]; F.ww = r->f[VVM_WW];
    Prof P(r, VVM_K_UPDATE, 12.0 * 8.0 * npts(r));
    hipLaunchKernelGGL(k_m_update, dim3(nblk((long long)r->nf, WG)), dim3(WG), 0, r->stream, r->g, F);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_step(vvm_run *r)
{
    int st;
    if ((st = launch_derivatives(r))) return st;
    if ((st = launch_explicit(r))) return st;
    if ((st = launch_rhs(r))) return st;
    if ((st = launch_poisson(r, 3))) return st;
    return launch_update(r);
}

int capture(vvm_run *r, int k, int steps)
{
    VV_HIP(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
    int st = VV_OK;
    for (int s = 0; s < steps && st == VV_OK; ++s) st = launch_step(r);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(r->stream, &gr);
    if (st) {
        if (gr) (void)hipGraphDestroy(gr);
        return st;
    }
    VV_HIP(e);
    r->graph[k] = gr;
    VV_HIP(hipGraphInstantiate(&r->gexec[k], gr, nullptr, nullptr, 0));
    return VV_OK;
}

int setup(vvm_run *r, const double *rn, const double *zn)
{
    const int nr = r->nr, nz = r->nz;
    std::vector<double> wr, wz;
    three_point(nr, rn, wr);
    three_point(nz, zn, wz);
    std::vector<double> cr(ctab_size(nr)), cz(ctab_size(nz));
    compact_tables(nr, rn, cr.data());
    compact_tables(nz, zn, cz.data());
    std::vector<double> hz(nz + 1), hr(nr + 1);
    for (int j = 0; j <= nz; ++j)
        hz[j] = 0.5 * ((j > 0 ? zn[j] - zn[j - 1] : 0.0) + (j < nz ? zn[j + 1] - zn[j] : 0.0));
    for (int i = 0; i <= nr; ++i)
        hr[i] = 0.5 * ((i > 0 ? rn[i] - rn[i - 1] : 0.0) + (i < nr ? rn[i + 1] - rn[i] : 0.0));
    double c[9], hw[6];
    fornberg(rn, 3, rn[0], 1, c);
    hw[0] = c[3]; hw[1] = c[4]; hw[2] = c[5];
    r->g.va0 = c[4]; r->g.va1 = c[5];
    fornberg(rn + nr - 2, 3, rn[nr], 1, c);
    hw[3] = c[3]; hw[4] = c[4]; hw[5] = c[5];
    r->g.vb1 = c[3]; r->g.vb0 = c[4];
    fornberg(zn, 3, zn[0], 1, c);
    r->g.ve0 = c[4]; r->g.ve1 = c[5];
    r->g.t0 = c[3]; r->g.t1 = c[4]; r->g.t2 = c[5];
    fornberg(zn + nz - 2, 3, zn[nz], 1, c);
    r->g.vg1 = c[3]; r->g.vg0 = c[4];
    r->g.s2 = c[3]; r->g.s1 = c[4]; r->g.s0 = c[5];
    // one device block of tables
    std::vector<double> all;
    auto put = [&](const double *p, size_t n) {
        const size_t o = all.size();
        all.insert(all.end(), p, p + n);
        return o;
    };
    const size_t o_r = put(rn, nr + 1), o_wr = put(wr.data(), wr.size()), o_wz = put(wz.data(), wz.size());
    const size_t o_cr = put(cr.data(), cr.size()), o_cz = put(cz.data(), cz.size());
    const size_t o_hz = put(hz.data(), hz.size()), o_hr = put(hr.data(), hr.size()), o_hw = put(hw, 6);
    VV_HIP(hipMalloc(&r->geo, all.size() * sizeof(double)));
    VV_HIP(hipMemcpy(r->geo, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice));
    MGeo &g = r->g;
    g.nr = nr; g.nz = nz; g.ld = r->ld;
    g.dt = r->cfg.dt; g.Ra = r->cfg.Ra; g.Re = r->cfg.Re; g.Pr = r->cfg.Pr;
    g.RRP = (r->cfg.Re != 0.0) ? r->cfg.Ra / r->cfg.Re / r->cfg.Pr : 0.0;
    g.rn = r->geo + o_r; g.wr = r->geo + o_wr; g.wz = r->geo + o_wz;
    g.cr = r->geo + o_cr; g.cz = r->geo + o_cz;
    r->hz = r->geo + o_hz; r->hr = r->geo + o_hr; r->hw = r->geo + o_hw;
    // derivative launch shape: per direction the half table (4 words a node),
    // 16 words of end weights and as many lines as fit in 64 KB of LDS
    auto lines_fit = [](int n) { return (LDS_WORDS - 4 * (n + 1) - 16) / ((n + 1) | 1); };
    r->nlr = std::max(1, std::min(64, lines_fit(nr)));
    r->nlz = std::max(1, std::min(16, lines_fit(nz)));
    r->nbr = 3 * ((nz - 1 + r->nlr - 1) / r->nlr);
    r->nbz = 3 * ((nr - 1 + r->nlz - 1) / r->nlz);
    r->dlds = sizeof(double) * std::max(4 * (size_t)(nr + 1) + 16 + (size_t)r->nlr * ((nr + 1) | 1),
                                        4 * (size_t)(nz + 1) + 16 + (size_t)r->nlz * ((nz + 1) | 1));
    // Poisson: z eigenbasis, r rows of both operators
    const int M = nz - 1, N = nr - 1;
    std::vector<double> za(M), zb(M), zc(M);
    for (int j = 1; j <= nz - 1; ++j) {
        const double *z2 = &wz[6 * j + 3];
        za[j - 1] = z2[0];
        zb[j - 1] = z2[1];
        zc[j - 1] = z2[2];
    }
    std::vector<double> lam, V, Vi;
    if (!vv::tridiag_eigen(za, zb, zc, lam, V, Vi)) {
        vv::set_error("vvm_create: the eigen-decomposition of the z rows failed");
        return VV_EINVAL;
    }
    std::vector<double> rt(6 * (size_t)N);
    for (int k = 0; k < 2; ++k)
        for (int i = 1; i <= nr - 1; ++i) {
            const double *r1 = &wr[6 * i], *r2 = r1 + 3;
            const double rr = rn[i];
            double *o = &rt[(size_t)k * 3 * N];
            o[i - 1] = r2[0] + r1[0] / rr;
            o[N + i - 1] = r2[1] + r1[1] / rr - (k == 0 ? 1.0 / (rr * rr) : 0.0);
            o[2 * N + i - 1] = r2[2] + r1[2] / rr;
        }
    VV_HIP(hipMalloc(&r->V, V.size() * sizeof(double)));
    VV_HIP(hipMemcpy(r->V, V.data(), V.size() * sizeof(double), hipMemcpyHostToDevice));
    VV_HIP(hipMalloc(&r->Vi, Vi.size() * sizeof(double)));
    VV_HIP(hipMemcpy(r->Vi, Vi.data(), Vi.size() * sizeof(double), hipMemcpyHostToDevice));
    VV_HIP(hipMalloc(&r->lam, lam.size() * sizeof(double)));
    VV_HIP(hipMemcpy(r->lam, lam.data(), lam.size() * sizeof(double), hipMemcpyHostToDevice));
    VV_HIP(hipMalloc(&r->rt, rt.size() * sizeof(double)));
    VV_HIP(hipMemcpy(r->rt, rt.data(), rt.size() * sizeof(double), hipMemcpyHostToDevice));
    VV_HIP(hipMalloc(&r->G, 2 * (size_t)M * N * sizeof(double)));
    VV_HIP(hipMalloc(&r->X, 2 * (size_t)M * N * sizeof(double)));
    VV_HIP(hipMalloc(&r->part, (2 * MAXB + 8) * sizeof(double)));
    return VV_OK;
}

int read_partials(vvm_run *r, int nb, int stride, std::vector<double> &h)
{
    h.resize((size_t)nb * stride);
    VV_HIP(hipMemcpyAsync(h.data(), r->part, h.size() * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

}  // namespace
}  // namespace vvm

extern "C" {

int vvm_create(const vvm_config *cfg, const double *rn, const double *zn, int device, vvm_run **out)
{
    using vv::set_error;
    if (!cfg || !rn || !zn || !out) { set_error("vvm_create: null argument"); return VV_EINVAL; }
    *out = nullptr;
    if (cfg->nr < 8 || cfg->nz < 8 || cfg->nr > 1024 || cfg->nz > 1024) {
        set_error("vvm_create: need 8 <= nr, nz <= 1024");
        return VV_EINVAL;
    }
    if (!(cfg->dt > 0) || !(cfg->Pr > 0) || !std::isfinite(cfg->Ra) || !std::isfinite(cfg->Re) ||
        !std::isfinite(cfg->dt) || !std::isfinite(cfg->Pr)) {
        set_error("vvm_create: bad physical parameters");
        return VV_EINVAL;
    }
    if (!(rn[0] > 0) || !std::isfinite(rn[0]) || !std::isfinite(zn[0])) {
        set_error("vvm_create: need 0 < r[0] and finite nodes");
        return VV_EINVAL;
    }
    for (int i = 0; i < cfg->nr; ++i)
        if (!(rn[i + 1] > rn[i]) || !std::isfinite(rn[i + 1])) {
            set_error("vvm_create: r nodes must increase strictly");
            return VV_EINVAL;
        }
    for (int j = 0; j < cfg->nz; ++j)
        if (!(zn[j + 1] > zn[j]) || !std::isfinite(zn[j + 1])) {
            set_error("vvm_create: z nodes must increase strictly");
            return VV_EINVAL;
        }
    VV_HIP(hipSetDevice(device));
    vvm_run *r = new vvm_run();
    r->cfg = *cfg;
    r->device = device;
    r->nr = cfg->nr;
    r->nz = cfg->nz;
    r->ld = cfg->nr + 1;
    r->nf = (size_t)(cfg->nr + 1) * (cfg->nz + 1);
    auto fail = [&](int s) { vvm_destroy(r); return s; };
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) return fail(VV_EHIP);
    for (int q = 0; q < VVM_NFIELD; ++q) {
        if (hipMalloc(&r->f[q], r->nf * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
        if (hipMemsetAsync(r->f[q], 0, r->nf * sizeof(double), r->stream) != hipSuccess) return fail(VV_EHIP);
    }
    for (int d = 0; d < 6; ++d) {
        if (hipMalloc(&r->dv[d], r->nf * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
        if (hipMemsetAsync(r->dv[d], 0, r->nf * sizeof(double), r->stream) != hipSuccess) return fail(VV_EHIP);
    }
    if (hipEventCreate(&r->ev[0]) != hipSuccess || hipEventCreate(&r->ev[1]) != hipSuccess) return fail(VV_EHIP);
    {
        const int st = setup(r, rn, zn);
        if (st) return fail(st);
    }
    if (hipStreamSynchronize(r->stream) != hipSuccess) return fail(VV_EHIP);
    *out = r;
    return VV_OK;
}

int vvm_destroy(vvm_run *r)
{
    if (!r) return VV_OK;
    {
        std::lock_guard<std::recursive_mutex> lk(r->mu);
        (void)hipSetDevice(r->device);
        if (r->stream) (void)hipStreamSynchronize(r->stream);
        for (int k = 0; k < 2; ++k) {
            if (r->gexec[k]) (void)hipGraphExecDestroy(r->gexec[k]);
            if (r->graph[k]) (void)hipGraphDestroy(r->graph[k]);
        }
        for (double *p : r->f)
            if (p) (void)hipFree(p);
        for (double *p : r->dv)
            if (p) (void)hipFree(p);
        for (double *p : {r->geo, r->V, r->Vi, r->lam, r->rt, r->G, r->X, r->part})
            if (p) (void)hipFree(p);
        for (hipEvent_t e : r->ev)
            if (e) (void)hipEventDestroy(e);
        if (r->stream) (void)hipStreamDestroy(r->stream);
    }
    delete r;
    return VV_OK;
}

static int vvm_copy(vvm_run *r, double *dev, double *host, bool to_dev)
{
    VV_HIP(hipSetDevice(r->device));
    if (to_dev)
        VV_HIP(hipMemcpyAsync(dev, host, r->nf * sizeof(double), hipMemcpyHostToDevice, r->stream));
    else
        VV_HIP(hipMemcpyAsync(host, dev, r->nf * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vvm_put(vvm_run *r, int field, const double *host)
{
    if (!r || !host || field < 0 || field >= VVM_NFIELD) { vv::set_error("vvm_put: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    return vvm_copy(r, r->f[field], const_cast<double *>(host), true);
}

int vvm_get(vvm_run *r, int field, double *host)
{
    if (!r || !host || field < 0 || field >= VVM_NFIELD) { vv::set_error("vvm_get: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    return vvm_copy(r, r->f[field], host, false);
}

int vvm_get_derivative(vvm_run *r, int d, double *host)
{
    if (!r || !host || d < 0 || d >= 6) { vv::set_error("vvm_get_derivative: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    return vvm_copy(r, r->dv[d], host, false);
}

#define VVM_CALL(expr)                                       \
    do {                                                     \
        if (!r) { vv::set_error("null run"); return VV_EINVAL; } \
        std::lock_guard<std::recursive_mutex> lk(r->mu);     \
        VV_HIP(hipSetDevice(r->device));                     \
        const int st_ = (expr);                              \
        if (st_) return st_;                                 \
        VV_HIP(hipStreamSynchronize(r->stream));             \
        return VV_OK;                                        \
    } while (0)

int vvm_derivatives(vvm_run *r) { VVM_CALL(launch_derivatives(r)); }
int vvm_explicit(vvm_run *r) { VVM_CALL(launch_explicit(r)); }
int vvm_rhs(vvm_run *r) { VVM_CALL(launch_rhs(r)); }
int vvm_update(vvm_run *r) { VVM_CALL(launch_update(r)); }

int vvm_poisson(vvm_run *r, int which)
{
    if (which < 1 || which > 2) { vv::set_error("vvm_poisson: which must be 1 or 2"); return VV_EINVAL; }
    VVM_CALL(launch_poisson(r, which));
}

int vvm_step(vvm_run *r, int n)
{
    if (!r || n < 0) { vv::set_error("vvm_step: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    if (r->prof || !r->use_graphs) {
        for (int s = 0; s < n; ++s) {
            const int st = launch_step(r);
            if (st) return st;
        }
    } else {
        int left = n;
        if (left >= 10 && !r->gexec[1]) {
            const int st = capture(r, 1, 10);
            if (st) return st;
        }
        for (; left >= 10; left -= 10) VV_HIP(hipGraphLaunch(r->gexec[1], r->stream));
        if (left > 0 && !r->gexec[0]) {
            const int st = capture(r, 0, 1);
            if (st) return st;
        }
        for (; left > 0; --left) VV_HIP(hipGraphLaunch(r->gexec[0], r->stream));
    }
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vvm_graphs(vvm_run *r, int enable)
{
    if (!r) { vv::set_error("null run"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    r->use_graphs = enable != 0;
    return VV_OK;
}

int vvm_residual(vvm_run *r, int which, double *relres)
{
    if (!r || !relres || which < 1 || which > 2) { vv::set_error("vvm_residual: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    const int nb = std::min(MAXB, std::max(1, nblk((long long)r->nf, WG)));
    hipLaunchKernelGGL(k_m_resid, dim3(nb), dim3(WG), 0, r->stream, r->g, which,
                       (const double *)r->f[which == 1 ? VVM_U : VVM_WW],
                       (const double *)r->f[which == 1 ? VVM_FU : VVM_FW], r->part);
    VV_HIP(hipGetLastError());
    std::vector<double> h;
    const int st = read_partials(r, nb, 2, h);
    if (st) return st;
    double er = 0.0, fm = 0.0;
    for (int b = 0; b < nb; ++b) {
        if (!(h[2 * b] <= er)) er = h[2 * b];
        if (!(h[2 * b + 1] <= fm)) fm = h[2 * b + 1];
    }
    *relres = (fm > 0.0) ? er / fm : er;
    return VV_OK;
}

int vvm_heat(vvm_run *r, double out[2])
{
    if (!r || !out) { vv::set_error("vvm_heat: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    hipLaunchKernelGGL(k_m_heat, dim3(1), dim3(WG), 2 * (size_t)(r->nz + 1) * sizeof(double), r->stream, r->g,
                       (const double *)r->f[VVM_TH], r->hz, r->hw, r->part + 2 * MAXB);
    VV_HIP(hipGetLastError());
    VV_HIP(hipMemcpyAsync(out, r->part + 2 * MAXB, 2 * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vvm_energy(vvm_run *r, double *ke)
{
    if (!r || !ke) { vv::set_error("vvm_energy: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    const int nb = std::min(MAXB, std::max(1, nblk((long long)r->nf, WG)));
    hipLaunchKernelGGL(k_m_energy, dim3(nb), dim3(WG), 0, r->stream, r->g, (const double *)r->f[VVM_U],
                       (const double *)r->f[VVM_WW], (const double *)r->f[VVM_V], r->hz, r->hr, r->part);
    VV_HIP(hipGetLastError());
    std::vector<double> h;
    const int st = read_partials(r, nb, 1, h);
    if (st) return st;
    double s = 0.0;
    for (double x : h) s += x;
    *ke = s;
    return VV_OK;
}

int vvm_profile(vvm_run *r, int enable)
{
    if (!r) { vv::set_error("null run"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    if (enable) {
        for (int k = 0; k < VVM_K_COUNT; ++k) {
            r->prof_n[k] = 0;
            r->prof_ms[k] = 0.0;
            r->prof_work[k] = 0.0;
        }
    }
    r->prof = enable != 0;
    return VV_OK;
}

int vvm_profile_read(vvm_run *r, int kernel, long long *launches, double *ms, double *work)
{
    if (!r || kernel < 0 || kernel >= VVM_K_COUNT) { vv::set_error("vvm_profile_read: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    if (launches) *launches = r->prof_n[kernel];
    if (ms) *ms = r->prof_ms[kernel];
    if (work) *work = r->prof_work[kernel];
    return VV_OK;
}

}  // extern "C"
