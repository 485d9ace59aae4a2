// vv_explicit.hip -- explicit part of the velocity-vorticity time step on
// gfx950: compact-scheme derivatives (batched constant-coefficient Thomas
// and cyclic solves), DuFort-Frankel/leapfrog stencils with the Robert
// filter, Poisson right-hand sides, wall update, wall heat flux.
//
// Arithmetic follows the reference expressions operation by operation
// (rb.f = velvor2dperiodicrbconv.f) and the library is compiled with
// -ffp-contract=off, so these kernels reproduce the reference bit for bit.
// Field layout: F[k*ld + i], i = 0..nr+1 across the walls, k = 0..np.

#include "vv_internal.h"
#include "vv_reduce.h"

#include <algorithm>
#include <cmath>

namespace vv {

namespace {

constexpr int WG = 256;

__device__ __forceinline__ size_t ix(int i, int k, int ld) { return (size_t)k * ld + i; }

// x / b for b > 0, correctly rounded, from rb = 1/b correctly rounded
// (computed on the host): q0 = x*rb is within one ulp of x/b, the remainder
// is exact in one FMA, and q0 + rem*rb rounds to x/b (Markstein's theorem
// for division with a fused multiply-add; no overflow or underflow occurs for
// these operands).  The remainder is formed as t = b*q0 - x and applied as
// -t: when it is zero, t = +0 and q0 + (-0)*rb keeps q0 and the sign of a
// zero quotient, so no select is needed.  Three dependent operations replace
// the IEEE division sequence on the recurrence's critical path, with the same
// result.  The elimination tables are checked for b > 0 when they are built.
__device__ __forceinline__ double div_rb(double x, double b, double rb)
{
    const double q0 = x * rb;
    const double t = __builtin_fma(b, q0, -x);
    return __builtin_fma(-t, rb, q0);
}

// single-precision literals of rb.f:660-668 promoted to double
__device__ __forceinline__ double f23() { return (double)(2.f / 3.f); }
__device__ __forceinline__ double f16() { return (double)(1.f / 6.f); }

struct CompactArgs {
    const double *C, *th;
    double *dCr, *dCh, *dthrC, *dthhC, *dthrT, *dthhT;
    const double *tab_r1, *tab_rT, *tab_p;
    int nr, np, ld, mask;
    double dri, dpi;
};

// Right-hand sides of the compact systems.  FAM 0: Cr (rb.f:810-824),
// FAM 1: thr at vorticity points (half derivative, rb.f:854-875), FAM 2: thr
// at temperature points with one-sided closures (rb.f:1244-1262); row i of
// column k.
template <int FAM>
__device__ __forceinline__ double rhs_r(const CompactArgs &a, int i, int k)
{
    const int nr = a.nr, ld = a.ld, np = a.np;
    const double dri = a.dri;
    const double *C = a.C, *th = a.th;
    const int ka = (k == np) ? 1 : k + 1;
    if (FAM == 0) {
        if (i == 1)
            return 3.0 * (C[ix(2, k, ld)] - C[ix(0, k, ld)]) * dri
                 - (-1.5 * C[ix(0, k, ld)] + 2.0 * C[ix(1, k, ld)] - 0.5 * C[ix(2, k, ld)]) * dri;
        if (i == nr - 1)
            return 3.0 * (C[ix(nr, k, ld)] - C[ix(nr - 2, k, ld)]) * dri
                 - (0.5 * C[ix(nr - 2, k, ld)] - 2.0 * C[ix(nr - 1, k, ld)] + 1.5 * C[ix(nr, k, ld)]) * dri;
        return 3.0 * (C[ix(i + 1, k, ld)] - C[ix(i - 1, k, ld)]) * dri;
    } else if (FAM == 1) {
        if (i == 1)
            return 1.5 * (th[ix(2, k, ld)] + th[ix(2, ka, ld)] - th[ix(1, k, ld)] - th[ix(1, ka, ld)]) * dri
                 - (-f23() * (th[ix(0, k, ld)] + th[ix(0, ka, ld)])
                    + 0.5 * (th[ix(1, k, ld)] + th[ix(1, ka, ld)])
                    + f16() * (th[ix(2, k, ld)] + th[ix(2, ka, ld)])) * dri;
        if (i == nr - 1)
            return 1.5 * (th[ix(nr, k, ld)] + th[ix(nr, ka, ld)] - th[ix(nr - 1, k, ld)] - th[ix(nr - 1, ka, ld)]) * dri
                 - (-f16() * (th[ix(nr - 1, k, ld)] + th[ix(nr - 1, ka, ld)])
                    - 0.5 * (th[ix(nr, k, ld)] + th[ix(nr, ka, ld)])
                    + f23() * (th[ix(nr + 1, k, ld)] + th[ix(nr + 1, ka, ld)])) * dri;
        return 1.5 * (th[ix(i + 1, k, ld)] + th[ix(i + 1, ka, ld)] - th[ix(i, k, ld)] - th[ix(i, ka, ld)]) * dri;
    } else {
        if (i == 1)
            return dri * (0.533333333 * th[ix(0, k, ld)] - 3.6 * th[ix(1, k, ld)]
                          + 2.966666667 * th[ix(2, k, ld)] + 0.1 * th[ix(4, k, ld)]);
        if (i == nr)
            return dri * (-0.1 * th[ix(nr - 3, k, ld)] - 2.966666667 * th[ix(nr - 1, k, ld)]
                          + 3.6 * th[ix(nr, k, ld)] - 0.533333333 * th[ix(nr + 1, k, ld)]);
        return 3.0 * (th[ix(i + 1, k, ld)] - th[ix(i - 1, k, ld)]) * dri;
    }
}

// Right-hand side of entry j (1-based, column k = j-1) of the cyclic systems
// on np+1 points: FAM 3 Ch (rb.f:836-842), FAM 4 thh at vorticity points
// (rb.f:884-891), FAM 5 thh at temperature points (rb.f:1276-1282); row i.
template <int FAM>
__device__ __forceinline__ double rhs_p(const CompactArgs &a, int i, int j)
{
    const int np = a.np, ld = a.ld;
    const double dpi = a.dpi;
    const double *X = (FAM == 3) ? a.C : a.th;
    const int k = j - 1;
    if (FAM == 4) {
        const int ka = (k == np) ? 1 : k + 1;
        return 1.5 * (X[ix(i + 1, ka, ld)] + X[ix(i, ka, ld)] - X[ix(i + 1, k, ld)] - X[ix(i, k, ld)]) * dpi;
    }
    if (k == np || k == 0) return 3.0 * (X[ix(i, 1, ld)] - X[ix(i, np - 1, ld)]) * dpi;
    return 3.0 * (X[ix(i, k + 1, ld)] - X[ix(i, k - 1, ld)]) * dpi;
}

// Batched Thomas along r, the reference tdma (rb.f:3046-3064) with its
// elimination factors from TdmaTab (fact | bp | c): the workgroup computes
// the right-hand sides of its T columns into LDS with coalesced loads, one
// thread per column then runs the recurrences in the reference's order, and
// the workgroup writes the solutions back.  LDS row stride nP is odd.
template <int FAM, bool TL>
__device__ void tdma_tile(const CompactArgs &a, double *R, double *Tab, int T)
{
    const int nr = a.nr, ld = a.ld, np = a.np;
    const int n = (FAM == 2) ? nr : nr - 1;
    const int nP = n | 1;
    const double *tab = (FAM == 2) ? a.tab_rT : a.tab_r1;
    double *out = (FAM == 0) ? a.dCr : (FAM == 1) ? a.dthrC : a.dthrT;
    const int k0 = blockIdx.x * T;
    if (k0 > np) return;
    const int tc = (np + 1 - k0 < T) ? np + 1 - k0 : T;
    // the elimination factors go to LDS with the right-hand sides (when they
    // fit), so the recurrences below read nothing from global memory
    if (TL) {
#pragma unroll 8
        for (int idx = threadIdx.x; idx < 4 * n; idx += blockDim.x) Tab[idx] = tab[idx];
    }
    const double *tb = TL ? Tab : tab;
    const double *fact = tb, *bp = tb + n, *cc = tb + 2 * n, *rbp = tb + 3 * n;
#pragma unroll 8
    for (int idx = threadIdx.x; idx < tc * n; idx += blockDim.x) {
        const int c = idx / n, j = idx - c * n;
        R[c * nP + j] = rhs_r<FAM>(a, j + 1, k0 + c);
    }
    __syncthreads();
    if ((int)threadIdx.x < tc) {
        // the recurrences in blocks of U: the loads of a block are issued
        // before its dependent chain, so only the arithmetic is serial
        constexpr int U = 8;
        double *r = R + threadIdx.x * nP;
        double rp = r[0];
        int j = 1;
        for (; j + U <= n; j += U) {
            double fr[U], rr[U];
#pragma unroll
            for (int u = 0; u < U; ++u) { fr[u] = fact[j + u]; rr[u] = r[j + u]; }
#pragma unroll
            for (int u = 0; u < U; ++u) { rp = rr[u] - fr[u] * rp; r[j + u] = rp; }
        }
        for (; j < n; ++j) {
            rp = r[j] - fact[j] * rp;
            r[j] = rp;
        }
        double y = div_rb(rp, bp[n - 1], rbp[n - 1]);
        r[n - 1] = y;
        j = n - 2;
        for (; j - U + 1 >= 0; j -= U) {
            double cr[U], br[U], ir[U], rr[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                cr[u] = cc[j - u]; br[u] = bp[j - u]; ir[u] = rbp[j - u]; rr[u] = r[j - u];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) { y = div_rb(rr[u] - cr[u] * y, br[u], ir[u]); r[j - u] = y; }
        }
        for (; j >= 0; --j) {
            y = div_rb(r[j] - cc[j] * y, bp[j], rbp[j]);
            r[j] = y;
        }
    }
    __syncthreads();
#pragma unroll 8
    for (int idx = threadIdx.x; idx < tc * n; idx += blockDim.x) {
        const int c = idx / n, j = idx - c * n;
        out[ix(j + 1, k0 + c, ld)] = R[c * nP + j];
    }
}

// Batched cyclic solve along phi, the reference ptdma (rb.f:3066-3116) on
// n = np+1 points with the factors of PtdmaTab (1-based, stride n+1:
// f1 | f2 | a | b | c | fA fB fC).  LDS holds entry j of row c at
// (j-1)*T + c, so the per-row recurrences of consecutive threads touch
// consecutive words.
template <int FAM, bool TL>
__device__ void ptdma_tile(const CompactArgs &a, double *R, double *Tab, int T)
{
    const int np = a.np, ld = a.ld;
    const int n = np + 1;
    const int s = n + 1;
    const int rows = (FAM == 5) ? a.nr : a.nr - 1;   // rows 1..rows
    const int i0 = 1 + blockIdx.x * T;
    if (i0 > rows) return;
    const int tc = (rows + 1 - i0 < T) ? rows + 1 - i0 : T;
    double *out = (FAM == 3) ? a.dCh : (FAM == 4) ? a.dthhC : a.dthhT;
    // elimination factors to LDS when they fit (see tdma_tile)
    if (TL) {
#pragma unroll 8
        for (int idx = threadIdx.x; idx < 6 * s + 3; idx += blockDim.x) Tab[idx] = a.tab_p[idx];
    }
    const double *tb = TL ? Tab : a.tab_p;
#pragma unroll 8
    for (int idx = threadIdx.x; idx < tc * n; idx += blockDim.x) {
        const int j = idx / tc, c = idx - j * tc;
        R[j * T + c] = rhs_p<FAM>(a, i0 + c, j + 1);
    }
    __syncthreads();
    if ((int)threadIdx.x < tc) {
        const double *f1 = tb, *f2 = tb + s, *aa = tb + 2 * s;
        const double *bb = tb + 3 * s, *cc = tb + 4 * s;
        const double fA = tb[5 * s], fB = tb[5 * s + 1], fC = tb[5 * s + 2];
        const double *rbb = tb + 5 * s + 3;
        const int c = threadIdx.x;
#define R_(j) R[((j) - 1) * T + c]
        constexpr int U = 8;   // blocked as in tdma_tile
        double rn = R_(n);
        double rm = R_(1);
        int j = 2;
        for (; j + U - 1 <= n - 2; j += U) {
            double g1[U], g2[U], rr[U];
#pragma unroll
            for (int u = 0; u < U; ++u) { g1[u] = f1[j + u]; g2[u] = f2[j + u]; rr[u] = R_(j + u); }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const double r = rr[u] - g1[u] * rm;
                rn = rn - g2[u] * rm;
                R_(j + u) = r;
                rm = r;
            }
        }
        for (; j <= n - 2; ++j) {
            const double r = R_(j) - f1[j] * rm;
            rn = rn - f2[j] * rm;
            R_(j) = r;
            rm = r;
        }
        const double rn1 = R_(n - 1) - fA * R_(n - 2);
        rn = rn - fB * R_(n - 2);
        rn = rn - fC * rn1;
        const double yn = div_rb(rn, bb[n], rbb[n]);
        double y = div_rb(rn1 - cc[n - 1] * yn, bb[n - 1], rbb[n - 1]);
        R_(n) = yn;
        R_(n - 1) = y;
        j = n - 2;
        for (; j - U + 1 >= 1; j -= U) {
            double cr[U], ar[U], br[U], ir[U], rr[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                cr[u] = cc[j - u]; ar[u] = aa[j - u]; br[u] = bb[j - u]; ir[u] = rbb[j - u]; rr[u] = R_(j - u);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                y = div_rb(rr[u] - cr[u] * y - ar[u] * yn, br[u], ir[u]);
                R_(j - u) = y;
            }
        }
        for (; j >= 1; --j) {
            y = div_rb(R_(j) - cc[j] * y - aa[j] * yn, bb[j], rbb[j]);
            R_(j) = y;
        }
#undef R_
    }
    __syncthreads();
#pragma unroll 8
    for (int idx = threadIdx.x; idx < tc * n; idx += blockDim.x) {
        const int j = idx / tc, c = idx - j * tc;
        out[ix(i0 + c, j, ld)] = R[j * T + c];
    }
}

struct CompactTiles {
    int Tr, Tp;    // columns per workgroup (r families), rows per workgroup (phi families)
    int rwords;    // LDS words of the right-hand-side tile; the tables follow
    int tab_lds;   // 1: the elimination tables are staged in LDS
};

__global__ void __launch_bounds__(WG) k_compact(CompactArgs a, CompactTiles t)
{
    extern __shared__ double R[];
    const int fam = blockIdx.y;
    if (!(a.mask & (1 << fam))) return;
    // the table's place is a template parameter, so that the recurrences'
    // table reads compile to LDS reads (a pointer chosen at run time between
    // LDS and global memory would make them generic loads, whose every wait
    // covers the global-memory counter as well)
    if (t.tab_lds) {
        double *Tab = R + t.rwords;
        switch (fam) {
        case 0: tdma_tile<0, true>(a, R, Tab, t.Tr); break;
        case 1: tdma_tile<1, true>(a, R, Tab, t.Tr); break;
        case 2: tdma_tile<2, true>(a, R, Tab, t.Tr); break;
        case 3: ptdma_tile<3, true>(a, R, Tab, t.Tp); break;
        case 4: ptdma_tile<4, true>(a, R, Tab, t.Tp); break;
        case 5: ptdma_tile<5, true>(a, R, Tab, t.Tp); break;
        }
    } else {
        switch (fam) {
        case 0: tdma_tile<0, false>(a, R, nullptr, t.Tr); break;
        case 1: tdma_tile<1, false>(a, R, nullptr, t.Tr); break;
        case 2: tdma_tile<2, false>(a, R, nullptr, t.Tr); break;
        case 3: ptdma_tile<3, false>(a, R, nullptr, t.Tp); break;
        case 4: ptdma_tile<4, false>(a, R, nullptr, t.Tp); break;
        case 5: ptdma_tile<5, false>(a, R, nullptr, t.Tp); break;
        }
    }
}

// ---------------------------------------------------------------------------
// The same six line families with the recurrences evaluated in parallel
// (vv_config.lines = VV_LINES_SCAN): one wavefront per line, each lane a
// chunk of m consecutive rows.  Every recurrence of the reference's
// elimination is affine in its predecessor (y_j = r_j - f_j y_{j-1}; the
// back substitution x_j = p_j + q_j x_{j+1}), so a chunk composes its rows
// into one affine map, an inclusive scan of those maps across the 64 lanes
// (6 shuffle steps) gives every chunk its entry value, and the chunk then
// runs its rows from there.  The same elimination factors and pivots as the
// sequential form; only the association of the products changes, so the
// results agree to rounding, not bit for bit.  SM: the rows a lane holds
// (at most 32: lines of up to 2048 rows), a template parameter so that the
// chunk stays in registers.
constexpr int SM_MAX = 32;

// composes (a, b) after the maps of the lanes below (inclusive scan up)
__device__ __forceinline__ void scan_up(double &a, double &b)
{
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double pa = __shfl_up(a, d, 64), pb = __shfl_up(b, d, 64);
        if (lane >= d) {
            a = a + b * pa;
            b = b * pb;
        }
    }
}
// composes (a, b) before the maps of the lanes above (inclusive scan down)
__device__ __forceinline__ void scan_down(double &a, double &b)
{
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double na = __shfl_down(a, d, 64), nb = __shfl_down(b, d, 64);
        if (lane + d < 64) {
            a = a + b * na;
            b = b * nb;
        }
    }
}
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Thomas along r (rows 0..n-1 of column k), the tables of tdma_tile
template <int FAM, int SM>
__device__ void tdma_scan(const CompactArgs &a, int k)
{
    const int nr = a.nr, ld = a.ld;
    const int n = (FAM == 2) ? nr : nr - 1;
    const double *tab = (FAM == 2) ? a.tab_rT : a.tab_r1;
    double *out = (FAM == 0) ? a.dCr : (FAM == 1) ? a.dthrC : a.dthrT;
    const double *fact = tab, *bp = tab + n, *cc = tab + 2 * n, *rbp = tab + 3 * n;
    const int lane = (int)(threadIdx.x & 63), m = (n + 63) / 64, j0 = lane * m;
    double v[SM];
    // the forward elimination's map of this chunk: y_out = fa + fb y_in
    double fa = 0.0, fb = 1.0;
#pragma unroll
    for (int u = 0; u < SM; ++u) {
        const int j = j0 + u;
        if (u < m && j < n) {
            v[u] = rhs_r<FAM>(a, j + 1, k);
            if (j == 0) { fa = v[u]; fb = 0.0; }
            else { fa = v[u] - fact[j] * fa; fb = -fact[j] * fb; }
        }
    }
    scan_up(fa, fb);
    double y = __shfl_up(fa, 1, 64);   // the row before the chunk (row 0 resets it)
#pragma unroll
    for (int u = 0; u < SM; ++u) {
        const int j = j0 + u;
        if (u < m && j < n) {
            y = (j == 0) ? v[u] : v[u] - fact[j] * y;
            v[u] = y;
        }
    }
    // back substitution: x_j = p_j + q_j x_{j+1}, the last row q = 0
    double ba = 0.0, bb = 1.0;
#pragma unroll
    for (int u = SM - 1; u >= 0; --u) {
        const int j = j0 + u;
        if (u < m && j < n) {
            const double p = v[u] * rbp[j], q = (j == n - 1) ? 0.0 : -cc[j] * rbp[j];
            ba = p + q * ba;
            bb = q * bb;
        }
    }
    scan_down(ba, bb);
    double x = __shfl_down(ba, 1, 64);   // the row after the chunk (the last row ends it)
#pragma unroll
    for (int u = SM - 1; u >= 0; --u) {
        const int j = j0 + u;
        if (u < m && j < n) {
            const double p = v[u] * rbp[j], q = (j == n - 1) ? 0.0 : -cc[j] * rbp[j];
            x = p + q * x;
            out[ix(j + 1, k, ld)] = x;
        }
    }
}

// cyclic solve along phi (entries 1..n of row i), the tables of ptdma_tile
template <int FAM, int SM>
__device__ void ptdma_scan(const CompactArgs &a, int i)
{
    const int np = a.np, ld = a.ld;
    const int n = np + 1, s = n + 1;
    double *out = (FAM == 3) ? a.dCh : (FAM == 4) ? a.dthhC : a.dthhT;
    const double *tb = a.tab_p;
    const double *f1 = tb, *f2 = tb + s, *aa = tb + 2 * s, *bbt = tb + 3 * s, *cc = tb + 4 * s;
    const double fA = tb[5 * s], fB = tb[5 * s + 1], fC = tb[5 * s + 2];
    const double *rbb = tb + 5 * s + 3;
    // entries 1..n-2 in chunks; n-1 and n are closed afterwards
    const int ne = n - 2, lane = (int)(threadIdx.x & 63), m = (ne + 63) / 64, j0 = 1 + lane * m;
    double v[SM];
    double fa = 0.0, fb = 1.0;
#pragma unroll
    for (int u = 0; u < SM; ++u) {
        const int j = j0 + u;
        if (u < m && j <= ne) {
            v[u] = rhs_p<FAM>(a, i, j);
            if (j == 1) { fa = v[u]; fb = 0.0; }
            else { fa = v[u] - f1[j] * fa; fb = -f1[j] * fb; }
        }
    }
    scan_up(fa, fb);
    double rm = __shfl_up(fa, 1, 64);
    double sum = 0.0;   // this chunk's part of sum_j f2_j rm_{j-1}
#pragma unroll
    for (int u = 0; u < SM; ++u) {
        const int j = j0 + u;
        if (u < m && j <= ne) {
            if (j == 1) rm = v[u];
            else {
                sum += f2[j] * rm;
                rm = v[u] - f1[j] * rm;
            }
            v[u] = rm;
        }
    }
    sum = wave_sum(sum);
    // rm_{n-2}: the last lane's chunk that holds entry n-2
    const int lastlane = (ne - 1) / m;
    const double rm2 = __shfl(rm, lastlane, 64);
    const double Rn = rhs_p<FAM>(a, i, n), Rn1 = rhs_p<FAM>(a, i, n - 1);
    const double rn1 = Rn1 - fA * rm2;
    double rn = Rn - sum;
    rn = rn - fB * rm2;
    rn = rn - fC * rn1;
    const double yn = div_rb(rn, bbt[n], rbb[n]);
    const double yn1 = div_rb(rn1 - cc[n - 1] * yn, bbt[n - 1], rbb[n - 1]);
    // back substitution: y_j = p_j + q_j y_{j+1}, p_j = (rm_j - a_j yn)/b_j
    double ba = 0.0, bb = 1.0;
#pragma unroll
    for (int u = SM - 1; u >= 0; --u) {
        const int j = j0 + u;
        if (u < m && j <= ne) {
            const double p = (v[u] - aa[j] * yn) * rbb[j], q = -cc[j] * rbb[j];
            ba = p + q * ba;
            bb = q * bb;
        }
    }
    scan_down(ba, bb);
    // entry after the chunk: the next lane's first, or y_{n-1} past the last
    double y = __shfl_down(ba, 1, 64), yb = __shfl_down(bb, 1, 64);
    if (j0 + m > ne) { y = yn1; yb = 0.0; }
    else y = y + yb * yn1;   // the next lanes' maps end at entry n-1
#pragma unroll
    for (int u = SM - 1; u >= 0; --u) {
        const int j = j0 + u;
        if (u < m && j <= ne) {
            const double p = (v[u] - aa[j] * yn) * rbb[j], q = -cc[j] * rbb[j];
            y = p + q * y;
            out[ix(i, j - 1, ld)] = y;
        }
    }
    if (lane == 0) {
        out[ix(i, n - 2, ld)] = yn1;
        out[ix(i, n - 1, ld)] = yn;
    }
}

template <int SM>
__global__ void __launch_bounds__(WG) k_compact_scan(CompactArgs a)
{
    const int fam = blockIdx.y;
    if (!(a.mask & (1 << fam))) return;
    const int line = blockIdx.x * (WG / 64) + (int)(threadIdx.x >> 6);
    if (fam < 3) {
        if (line > a.np) return;   // whole waves leave together: the shuffles see full waves
        switch (fam) {
        case 0: tdma_scan<0, SM>(a, line); break;
        case 1: tdma_scan<1, SM>(a, line); break;
        default: tdma_scan<2, SM>(a, line); break;
        }
    } else {
        const int rows = (fam == 5) ? a.nr : a.nr - 1;
        if (line >= rows) return;
        switch (fam) {
        case 3: ptdma_scan<3, SM>(a, 1 + line); break;
        case 4: ptdma_scan<4, SM>(a, 1 + line); break;
        default: ptdma_scan<5, SM>(a, 1 + line); break;
        }
    }
}

struct DFArgs {
    const double *C, *Cp, *th, *thp, *uel, *vel;
    const double *dCr, *dCh, *dthrC, *dthhC, *dthrT, *dthhT;
    double *Cn, *Cf, *thn, *thf;
    const double *rC, *rT;
    int nr, np, ld, first, do_c, do_t, geom;
    double dri, dpi, dt, Pr, Ra, robert, Re;
};

// DuFort-Frankel vorticity (rb.f:900-982, first step rb.f:699-775) and
// temperature (rb.f:1292-1377, first step rb.f:1124-1204).  The Robert
// filtered current level goes to Cf/thf (the reference filters in place,
// rb.f:1041-1045 and 1435-1440, after all neighbours were read).
// Annulus (geom 1, velvor2dnewsch3.4.f): r = r0+i*dr at vorticity rows
// (ann.f:918) and r0+(i-1/2)*dr at temperature rows (ann.f:1331); RHS1
// gains Cra/r and thra/r, the leapfrog term uses Cp, and RHS2 gains the
// centrifugal buoyancy 2Ra(-thr u - thh v/r)/Re/Pr (ann.f:981-987, 1383).
__global__ void __launch_bounds__(WG) k_df(DFArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nr = a.nr, np = a.np, ld = a.ld;
    const int rows = nr + 2;
    if (t >= rows * (np + 1)) return;
    const int k = t / rows, i = t - k * rows;
    const int km = (k == 0) ? np - 1 : k - 1;
    const int ka = (k == np) ? 1 : k + 1;
    const double dri = a.dri, dpi = a.dpi, dt = a.dt, Ptl = a.Pr;
    const size_t o = ix(i, k, ld);
    const bool ann = (a.geom == 1);
    if (a.do_c) {
        const double Cc = a.C[o];
        if (i >= 1 && i <= nr - 1) {
            const double r = ann ? a.rC[i] : 1.0;
            const double ri = 1.0 / r;
            const double u = (a.uel[ix(i, ka, ld)] + a.uel[o]) * 0.5;
            const double v = (a.vel[o] + a.vel[ix(i + 1, k, ld)]) * 0.5;
            const double Chha = (a.C[ix(i, ka, ld)] + a.C[ix(i, km, ld)]) * (dpi * dpi);
            const double Ch = a.dCh[o];
            const double Crra = (a.C[ix(i - 1, k, ld)] + a.C[ix(i + 1, k, ld)]) * (dri * dri);
            const double cnr = -1.0 * (dri * dri);
            const double Cr = a.dCr[o];
            const double thh = a.dthhC[o];
            const double cnh = (-1.0 * (dpi * dpi)) * (ri * ri);
            double RHS1, cf1, RHS2;
            if (ann) {
                const double Cra = (a.C[ix(i + 1, k, ld)] - a.C[ix(i - 1, k, ld)]) * 0.5 * dri;
                const double thr = a.dthrC[o];
                if (a.first) {
                    RHS1 = Crra + Cra * ri + Chha * (ri * ri) + Cc / (Ptl * dt);
                    cf1 = 1.0 / (Ptl * dt) - cnr - cnh;
                } else {
                    RHS1 = Crra + Cra * ri + Chha * (ri * ri) + 0.5 * a.Cp[o] / (Ptl * dt);
                    cf1 = 0.5 / (Ptl * dt) - cnr - cnh;
                }
                RHS2 = -(a.Ra * thh) + 2.0 * a.Ra * (-(thr * u) - thh * v * ri) / a.Re / Ptl;
            } else {
                if (a.first) {
                    RHS1 = Crra + Chha * (ri * ri) + Cc / (Ptl * dt);
                    cf1 = 1.0 / (Ptl * dt) - cnr - cnh;
                } else {
                    RHS1 = Crra + Chha * (ri * ri) + 0.5 * Cc / (Ptl * dt);
                    cf1 = 0.5 / (Ptl * dt) - cnr - cnh;
                }
                RHS2 = -(a.Ra * thh);
            }
            const double ULT = (u * Cr + v * Ch * ri) / Ptl;
            const double cf2 = cnr + cnh;
            const double cn = (cf2 * a.Cp[o] - ULT - RHS2 + RHS1) / cf1;
            a.Cn[o] = cn;
            a.Cf[o] = a.first ? Cc : Cc + a.robert * (cn - 2.0 * Cc + a.Cp[o]);
        } else {
            a.Cn[o] = 0.0;
            a.Cf[o] = Cc;
        }
    }
    if (a.do_t) {
        const double T = a.th[o];
        if (i >= 1 && i <= nr) {
            const double r = ann ? a.rT[i] : 1.0;
            const double ri = 1.0 / r;
            const double u = (a.uel[o] + a.uel[ix(i - 1, k, ld)]) * 0.5;
            const double v = (a.vel[o] + a.vel[ix(i, km, ld)]) * 0.5;
            double thrra, cnr, thra;
            if (i == 1) {
                thrra = (3.2 * a.th[ix(0, k, ld)] + 2.0 * a.th[ix(2, k, ld)] - 0.2 * a.th[ix(3, k, ld)]) * (dri * dri);
                thra = (-4.0 * a.th[ix(0, k, ld)] / 3.0 + a.th[ix(2, k, ld)] / 3.0) * dri;
                cnr = -(2.5 * (dri * dri)) + 0.5 * dri * ri;
            } else if (i == nr) {
                thrra = (3.2 * a.th[ix(nr + 1, k, ld)] + 2.0 * a.th[ix(nr - 1, k, ld)] - 0.2 * a.th[ix(nr - 2, k, ld)]) * (dri * dri);
                thra = (-a.th[ix(nr - 1, k, ld)] / 3.0 + 4.0 * a.th[ix(nr + 1, k, ld)] / 3.0) * dri;
                cnr = -(2.5 * (dri * dri)) - 0.5 * dri * ri;
            } else {
                thrra = (a.th[ix(i - 1, k, ld)] + a.th[ix(i + 1, k, ld)]) * (dri * dri);
                thra = (a.th[ix(i + 1, k, ld)] - a.th[ix(i - 1, k, ld)]) * 0.5 * dri;
                cnr = -1.0 * (dri * dri);
            }
            const double thr = a.dthrT[o];
            const double thhha = (a.th[ix(i, ka, ld)] + a.th[ix(i, km, ld)]) * (dpi * dpi);
            const double thh = a.dthhT[o];
            const double cnp = -(1.0 * (dpi * dpi)) * (ri * ri);
            double RHS1, cf1;
            if (a.first) {
                RHS1 = ann ? thrra + thra * ri + thhha * (ri * ri) + T / dt
                           : thrra + thhha * (ri * ri) + T / dt;
                cf1 = 1.0 / dt - cnr - cnp;
            } else {
                RHS1 = ann ? thrra + thra * ri + thhha * (ri * ri) + 0.5 * a.thp[o] / dt
                           : thrra + thhha * (ri * ri) + 0.5 * a.thp[o] / dt;
                cf1 = 0.5 / dt - cnr - cnp;
            }
            const double ULT = u * thr + v * thh * ri;
            const double cf2 = cnr + cnp;
            const double tn = (cf2 * a.thp[o] - ULT + RHS1) / cf1;
            a.thn[o] = tn;
            a.thf[o] = a.first ? T : T + a.robert * (tn - 2.0 * T + a.thp[o]);
        } else {
            a.thf[o] = T;
        }
    }
}

// Wall heat flux, rb.f:1452-1500: the per-column terms are computed in
// parallel, then accumulated in the reference's order by one thread.
__global__ void __launch_bounds__(WG) k_heat(const double *th, int nr, int np, int ld,
                                             double dri, double dp, double *dst)
{
    __shared__ double xb[WG + 1], xt[WG + 1];
    double t1 = 0.0, t2 = 0.0;
    double first_b = 0.0, first_t = 0.0;
    for (int base = 1; base <= np; base += WG) {
        const int k = base + (int)threadIdx.x;
        if (k <= np) {
            xb[threadIdx.x] = 0.5 * dp * (-3.35238095 * th[ix(0, k, ld)] + 4.375 * th[ix(1, k, ld)]
                                          - 1.45833333 * th[ix(2, k, ld)] + 0.525 * th[ix(3, k, ld)]
                                          - 0.08928571 * th[ix(4, k, ld)]) * dri;
            xt[threadIdx.x] = 0.5 * dp * (-3.35238095 * th[ix(nr + 1, k, ld)] + 4.375 * th[ix(nr, k, ld)]
                                          - 1.45833333 * th[ix(nr - 1, k, ld)] + 0.525 * th[ix(nr - 2, k, ld)]
                                          - 0.08928571 * th[ix(nr - 3, k, ld)]) * dri;
        }
        if (threadIdx.x == 0) {   // term of column k+1 for the last k of this chunk
            const int kn = base + WG;
            if (kn <= np) {
                xb[WG] = 0.5 * dp * (-3.35238095 * th[ix(0, kn, ld)] + 4.375 * th[ix(1, kn, ld)]
                                     - 1.45833333 * th[ix(2, kn, ld)] + 0.525 * th[ix(3, kn, ld)]
                                     - 0.08928571 * th[ix(4, kn, ld)]) * dri;
                xt[WG] = 0.5 * dp * (-3.35238095 * th[ix(nr + 1, kn, ld)] + 4.375 * th[ix(nr, kn, ld)]
                                     - 1.45833333 * th[ix(nr - 1, kn, ld)] + 0.525 * th[ix(nr - 2, kn, ld)]
                                     - 0.08928571 * th[ix(nr - 3, kn, ld)]) * dri;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (base == 1) { first_b = xb[0]; first_t = xt[0]; }
            const int cnt = (np - base + 1 < WG) ? np - base + 1 : WG;
            // the neighbour of column np is column 1 (periodic); slot cnt of
            // the last chunk is free for it
            if (base + cnt - 1 == np) { xb[cnt] = first_b; xt[cnt] = first_t; }
            // the reference's order, t1 = (t1 + x_k) + x_{k+1}; blocks of U
            // terms are read from LDS before their dependent additions
            constexpr int U = 16;
            int j = 0;
            for (; j + U <= cnt; j += U) {
                double b[U + 1], t[U + 1];
#pragma unroll
                for (int u = 0; u <= U; ++u) { b[u] = xb[j + u]; t[u] = xt[j + u]; }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    t1 = t1 + b[u] + b[u + 1];
                    t2 = t2 + t[u] + t[u + 1];
                }
            }
            for (; j < cnt; ++j) {
                t1 = t1 + xb[j] + xb[j + 1];
                t2 = t2 + xt[j] + xt[j + 1];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { dst[0] = t1; dst[1] = t2; }
}

struct RhsArgs {
    const double *Cn, *vel, *velp, *uel, *uelp;
    double *vrh, *urh;
    double *fv, *uv, *fu, *uu;     // multigrid level 0
    double *fpart;                 // [2][gridDim.x] per-workgroup max |f|
    // solver state reset for a new solve (workgroup 0): problems active per
    // `which` (2: v, 3: u, 5: both)
    int *done, *cycles;
    unsigned long long *fnorm;
    double *relres;
    const double *rC, *rT;
    int nr, np, ld, which, from_fields, geom, solve;
    double dr, dp;
};

// multigrid level 0 is column-major: row j of column q at q*n + j
__device__ __forceinline__ size_t mgix(int j, int q, int n) { return (size_t)q * n + j; }

// rhsv (rb.f:1629-1676), rhsu (rb.f:1679-1721), the extrapolated initial
// guess of poisson2/3 (rb.f:1827-1835) scattered into the multigrid level 0,
// and per-workgroup maxima of |rhs| (combined by the solver's stop test).
__device__ __forceinline__ void rhs_point(const RhsArgs &a, int t, double &mv, double &mu)
{
    const int nr = a.nr, np = a.np, ld = a.ld;
    const int rows = nr + 2;
    const int k = t / rows, i = t - k * rows;
    const size_t o = ix(i, k, ld);
    if ((a.which & 1) && i >= 1 && i <= nr) {
        double v;
        if (a.from_fields) {
            v = a.vrh[o];
        } else {
            if (a.geom == 1) {          // ann.f:1663-1670
                const double r = a.rT[i];
                const double Cr = (a.Cn[o] - a.Cn[ix(i - 1, k, ld)]) / a.dr;
                if (i == 1) v = a.Cn[ix(1, k, ld)] / a.dr + a.Cn[ix(1, k, ld)] / r;
                else if (i == nr) v = -a.Cn[ix(nr - 1, k, ld)] / a.dr + a.Cn[ix(nr - 1, k, ld)] / r;
                else v = Cr + (a.Cn[o] + a.Cn[ix(i - 1, k, ld)]) / r;
            } else {
                if (i == 1) v = a.Cn[ix(1, k, ld)] / a.dr;
                else if (i == nr) v = -a.Cn[ix(nr - 1, k, ld)] / a.dr;
                else v = (a.Cn[o] - a.Cn[ix(i - 1, k, ld)]) / a.dr;
            }
            a.vrh[o] = v;
        }
        if (k >= 1) {
            const size_t m = mgix(i - 1, k - 1, nr);
            a.fv[m] = v;
            a.uv[m] = 2.0 * a.vel[o] - a.velp[o];
            mv = nanmax(mv, fabs(v));
        }
    }
    if ((a.which & 2) && i >= 1 && i <= nr - 1) {
        double u;
        if (a.from_fields) {
            u = a.urh[o];
        } else {
            const int km = (k == 0) ? np - 1 : k - 1;
            const double Ch = (a.Cn[o] - a.Cn[ix(i, km, ld)]) / a.dp;
            u = -Ch / (a.geom == 1 ? a.rC[i] : 1.0);   // ann.f:1715-1719
            a.urh[o] = u;
        }
        if (k >= 1) {
            const size_t m = mgix(i - 1, k - 1, nr - 1);
            a.fu[m] = u;
            a.uu[m] = 2.0 * a.uel[o] - a.uelp[o];
            mu = nanmax(mu, fabs(u));
        }
    }
}

__global__ void __launch_bounds__(WG) k_rhs_guess(RhsArgs a)
{
    __shared__ double sh[WG / 64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < MG_MAXPROB) {
        const bool active = (a.solve == 5) || (a.solve == 2 && t == 0) || (a.solve == 3 && t == 1);
        a.done[t] = active ? 0 : 1;
        a.cycles[t] = 0;
        a.fnorm[t] = 0ull;
        a.relres[t] = 0.0;
    }
    double mv = 0.0, mu = 0.0;
    if (t < (a.nr + 2) * (a.np + 1)) rhs_point(a, t, mv, mu);
    mv = block_max(mv, sh);
    mu = block_max(mu, sh);
    if (threadIdx.x == 0) {
        a.fpart[blockIdx.x] = mv;
        a.fpart[gridDim.x + blockIdx.x] = mu;
    }
}

struct UpdArgs {
    double *Cn, *thn;
    const double *vel, *uel;
    const double *rT;
    int geom;
    double *mT2, *mT, *mU2, *mV2;
    double *upart;      // [3][gridDim.x]: max CFL u, max CFL v, clamp count
    // from_mg: first move the solver's level-0 solutions u0v (nr rows) and
    // u0u (nr-1 rows) into vel/uel with velp/uelp <- old values
    double *velw, *uelw, *velp, *uelp;
    const double *u0v, *u0u;
    int from_mg;
    int nr, np, ld, kt, stats;
    double dr, dp, dt;
};

// update (rb.f:1557-1626) on the new level; the rotation Cp<-C<-Cn,
// thp<-th<-thn is a pointer swap on the host.  Running means
// (rb.f:409-417) and CFL maxima (rb.f:421-458) are fused here.
__device__ __forceinline__ void update_point(const UpdArgs &a, int t, double &cu, double &cv, double &ncl)
{
    const int nr = a.nr, ld = a.ld;
    const int rows = nr + 2;
    const int k = t / rows, i = t - k * rows;
    const size_t o = ix(i, k, ld);
    double U, V;
    if (a.from_mg) {
        // poisson2/3 output (rb.f:1894-1913): psip <- psi, psi <- solution,
        // walls 0; column k maps to solver column k-1 (k = 0 is k = np)
        const int q = (k == 0) ? a.np - 1 : k - 1;
        const double *sv = a.u0v + (size_t)q * nr;
        const double *su = a.u0u + (size_t)q * (nr - 1);
        const double vo = a.velw[o];
        a.velp[o] = vo;
        V = (i >= 1 && i <= nr) ? sv[i - 1] : 0.0;
        a.velw[o] = V;
        if (i <= nr) {
            const double uo = a.uelw[o];
            a.uelp[o] = uo;
            U = (i >= 1 && i <= nr - 1) ? su[i - 1] : 0.0;
            a.uelw[o] = U;
        } else {
            U = a.uelw[o];
        }
        if (i == 0) a.Cn[o] = (3.0 * sv[0] - sv[1] / 3.0) / a.dr;
        if (i == nr) a.Cn[o] = (sv[nr - 2] / 3.0 - 3.0 * sv[nr - 1]) / a.dr;
    } else {
        U = a.uel[o];
        V = a.vel[o];
        if (i == 0) a.Cn[o] = (3.0 * a.vel[ix(1, k, ld)] - a.vel[ix(2, k, ld)] / 3.0) / a.dr;
        if (i == nr) a.Cn[o] = (a.vel[ix(nr - 1, k, ld)] / 3.0 - 3.0 * a.vel[ix(nr, k, ld)]) / a.dr;
    }
    double T = a.thn[o];
    if (i == 0) T = 0.0;
    if (i == nr + 1) T = 1.0;
    int clamped = 0;
    if (T > 1.0) { T = 0.99; ++clamped; }
    if (T < 0.0) { T = 0.01; ++clamped; }
    a.thn[o] = T;
    ncl += (double)clamped;
    if (a.stats) {
        if (a.kt >= 2) {
            const double tkti = 1.0 / (double)(a.kt - 1);
            const double km2 = (double)(a.kt - 2);
            a.mT2[o] = tkti * a.mT2[o] * km2 + tkti * (T * T);
            a.mT[o] = tkti * a.mT[o] * km2 + tkti * T;
            a.mU2[o] = tkti * a.mU2[o] * km2 + tkti * (U * U);
            a.mV2[o] = tkti * a.mV2[o] * km2 + tkti * (V * V);
        }
        if (i >= 1 && i <= nr) cu = nanmax(cu, fabs(U * a.dt / a.dr));
        if (i >= 1 && i <= nr - 1) cv = nanmax(cv, fabs(V * a.dt / ((a.geom == 1 ? a.rT[i] : 1.0) * a.dp)));
    }
}

__global__ void __launch_bounds__(WG) k_update(UpdArgs a)
{
    __shared__ double sh[WG / 64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double cu = 0.0, cv = 0.0, ncl = 0.0;
    if (t < (a.nr + 2) * (a.np + 1)) update_point(a, t, cu, cv, ncl);
    cu = block_max(cu, sh);
    cv = block_max(cv, sh);
    ncl = block_sum(ncl, sh);
    if (threadIdx.x == 0) {
        a.upart[blockIdx.x] = cu;
        a.upart[gridDim.x + blockIdx.x] = cv;
        a.upart[2 * gridDim.x + blockIdx.x] = ncl;
    }
}

inline int nblk(long long n, int wg) { return (int)((n + wg - 1) / wg); }

// Kinetic energy 1/2 dr dp sum_k [sum_{i=1..nr} wT vel^2 + sum_{i=1..nr-1}
// wC uel^2] over the columns k = 0..np-1 (no counterpart in the reference;
// the weights are the row radii in the annulus, 1 in the box: wr = 1 / 0
// selects).  Workgroup b sums the columns b, b + gridDim.x, ... (rows across
// the threads), reduced by wavefront shuffles; k_energy_sum adds the
// partials.  The order is fixed, so the result repeats bit for bit.
__global__ void __launch_bounds__(WG) k_energy(const double *uel, const double *vel, const double *rC,
                                               const double *rT, double wr, int nr, int np, int ld, double *part)
{
    __shared__ double sh[WG / 64];
    double s = 0.0;
    for (int k = blockIdx.x; k < np; k += gridDim.x) {
        const double *u = uel + (size_t)k * ld, *v = vel + (size_t)k * ld;
        for (int i = 1 + (int)threadIdx.x; i <= nr; i += WG) {
            const double vv = v[i];
            s += (wr != 0.0 ? rT[i] : 1.0) * vv * vv;
            if (i <= nr - 1) {
                const double uu = u[i];
                s += (wr != 0.0 ? rC[i] : 1.0) * uu * uu;
            }
        }
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(WG) k_energy_sum(const double *part, int nb, double scale, double *out)
{
    __shared__ double sh[WG / 64];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += WG) s += part[b];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) *out = scale * s;
}

}  // namespace

static int compact_launch(vv_run *r, int mask)
{
    CompactArgs a;
    a.C = r->C; a.th = r->th;
    a.dCr = r->dCr; a.dCh = r->dCh; a.dthrC = r->dthrC; a.dthhC = r->dthhC;
    a.dthrT = r->dthrT; a.dthhT = r->dthhT;
    a.tab_r1 = r->tab_r1; a.tab_rT = r->tab_rT; a.tab_p = r->tab_p;
    a.nr = r->nr; a.np = r->np; a.ld = r->ld; a.mask = mask;
    a.dri = 1.0 / r->dr; a.dpi = 1.0 / r->dp;
    // tiles sized for 32 KB of LDS per workgroup
    const int budget = 32768 / 8;
    const int nP = r->nr | 1;
    CompactTiles t;
    t.Tr = budget / nP;
    if (t.Tr > 64) t.Tr = 64;
    if (t.Tr < 1) t.Tr = 1;
    t.Tp = budget / (r->np + 1);
    if (t.Tp > 64) t.Tp = 64;
    if (t.Tp < 1) t.Tp = 1;
    // a workgroup holds at least one whole line: Tp = 1 needs np+1 words,
    // Tr = 1 nP words, either of which can exceed the budget
    t.rwords = std::max(budget, std::max(t.Tr * nP, t.Tp * (r->np + 1)));
    // + the elimination tables: 4n words (r families, n <= nr), 6(np+2)+3 (phi)
    const int twords = (4 * r->nr > 6 * (r->np + 2) + 3) ? 4 * r->nr : 6 * (r->np + 2) + 3;
    // dynamic LDS is kept within 64 KB; larger tables stay in global memory
    t.tab_lds = ((size_t)(t.rwords + twords) * 8 <= 65536) ? 1 : 0;
    const size_t lds = (size_t)(t.rwords + (t.tab_lds ? twords : 0)) * 8;
    if (lds > 160 * 1024) { set_error("compact-scheme line longer than the LDS of a CU holds"); return VV_EINVAL; }
    const int gr = nblk(r->np + 1, t.Tr), gp = nblk(r->nr, t.Tp);
    dim3 grid(gr > gp ? gr : gp, 6);
    // the scan form: a wavefront per line, lines of up to 64 SM_MAX rows
    const bool scan = r->cfg.lines == VV_LINES_SCAN && r->nr <= 64 * SM_MAX && r->np - 1 <= 64 * SM_MAX;
    const int mrow = std::max((r->nr + 63) / 64, (r->np - 1 + 63) / 64);   // rows per lane
    hipEvent_t pe_;
    prof_begin(r, &pe_);
    if (scan) {
        const int lines = (r->np + 1 > r->nr) ? r->np + 1 : r->nr;
        const dim3 g(nblk(lines, WG / 64), 6);
        if (mrow <= 4) hipLaunchKernelGGL(k_compact_scan<4>, g, dim3(WG), 0, r->stream, a);
        else if (mrow <= 8) hipLaunchKernelGGL(k_compact_scan<8>, g, dim3(WG), 0, r->stream, a);
        else if (mrow <= 16) hipLaunchKernelGGL(k_compact_scan<16>, g, dim3(WG), 0, r->stream, a);
        else hipLaunchKernelGGL(k_compact_scan<32>, g, dim3(WG), 0, r->stream, a);
    } else {
        hipLaunchKernelGGL(k_compact, grid, dim3(WG), lds, r->stream, a, t);
    }
    // read C and th, write six derivative fields
    prof_end(r, VV_K_COMPACT, pe_, 8.0 * 8.0 * (double)(r->nr + 2) * (r->np + 1));
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int field_blocks(const vv_run *r) { return nblk((long long)(r->nr + 2) * (r->np + 1), WG); }

int launch_compact(vv_run *r) { return compact_launch(r, 0x3f); }
int launch_compact_mask(vv_run *r, int mask) { return compact_launch(r, mask); }

int launch_df(vv_run *r, int first, int do_c, int do_t)
{
    DFArgs a;
    a.C = r->C; a.Cp = r->Cp; a.th = r->th; a.thp = r->thp; a.uel = r->uel; a.vel = r->vel;
    a.dCr = r->dCr; a.dCh = r->dCh; a.dthrC = r->dthrC; a.dthhC = r->dthhC; a.dthrT = r->dthrT;
    a.dthhT = r->dthhT;
    a.Cn = r->Cn; a.Cf = r->Cf; a.thn = r->thn; a.thf = r->thf;
    a.rC = r->rC; a.rT = r->rT;
    a.nr = r->nr; a.np = r->np; a.ld = r->ld; a.first = first; a.do_c = do_c; a.do_t = do_t;
    a.geom = r->cfg.geom;
    a.dri = 1.0 / r->dr; a.dpi = 1.0 / r->dp; a.dt = r->cfg.dt; a.Pr = r->cfg.Pr; a.Ra = r->cfg.Ra;
    a.robert = r->cfg.robert; a.Re = r->cfg.Re;
    const long long n = (long long)(r->nr + 2) * (r->np + 1);
    hipEvent_t pe_;
    prof_begin(r, &pe_);
    hipLaunchKernelGGL(k_df, dim3(nblk(n, WG)), dim3(WG), 0, r->stream, a);
    prof_end(r, VV_K_DF, pe_, 8.0 * 15.0 * (double)n);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_heat(vv_run *r, double *dst2)
{
    hipEvent_t pe_;
    prof_begin(r, &pe_);
    hipLaunchKernelGGL(k_heat, dim3(1), dim3(WG), 0, r->stream, (const double *)r->thf, r->nr,
                       r->np, r->ld, 1.0 / r->dr, r->dp, dst2);
    prof_end(r, VV_K_HEAT, pe_, 8.0 * 10.0 * (r->np + 1));
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_energy(vv_run *r)
{
    const int nb = r->np < KE_MAXB ? r->np : KE_MAXB;
    const double w = (r->cfg.geom == VV_GEOM_ANNULUS) ? 1.0 : 0.0;
    hipLaunchKernelGGL(k_energy, dim3(nb), dim3(WG), 0, r->stream, (const double *)r->uel, (const double *)r->vel,
                       (const double *)r->rC, (const double *)r->rT, w, r->nr, r->np, r->ld, r->kepart);
    hipLaunchKernelGGL(k_energy_sum, dim3(1), dim3(WG), 0, r->stream, (const double *)r->kepart, nb,
                       0.5 * r->dr * r->dp, r->kepart + KE_MAXB);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_heat_from(vv_run *r, const double *th, double *dst2)
{
    hipLaunchKernelGGL(k_heat, dim3(1), dim3(WG), 0, r->stream, th, r->nr, r->np, r->ld,
                       1.0 / r->dr, r->dp, dst2);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

static int launch_rhs_common(vv_run *r, int which, int from_fields, int solve)
{
    RhsArgs a;
    a.from_fields = from_fields;
    a.solve = solve;
    a.done = r->mg.done; a.cycles = r->mg.cycles; a.fnorm = r->mg.fnorm; a.relres = r->mg.relres;
    a.Cn = r->Cn; a.vel = r->vel; a.velp = r->velp; a.uel = r->uel; a.uelp = r->uelp;
    a.vrh = r->vrh; a.urh = r->urh;
    a.fv = r->mg.prob[0].lev[0].f; a.uv = r->mg.prob[0].lev[0].u;
    a.fu = r->mg.prob[1].lev[0].f; a.uu = r->mg.prob[1].lev[0].u;
    a.fpart = r->fpart;
    a.nr = r->nr; a.np = r->np; a.ld = r->ld; a.which = which;
    a.dr = r->dr; a.dp = r->dp;
    a.rC = r->rC; a.rT = r->rT; a.geom = r->cfg.geom;
    const long long n = (long long)(r->nr + 2) * (r->np + 1);
    hipEvent_t pe_;
    prof_begin(r, &pe_);
    hipLaunchKernelGGL(k_rhs_guess, dim3(nblk(n, WG)), dim3(WG), 0, r->stream, a);
    prof_end(r, VV_K_RHS, pe_, 8.0 * 11.0 * (double)n);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_rhs_guess(vv_run *r, int which, int solve) { return launch_rhs_common(r, which, 0, solve); }
int launch_rhs_fields(vv_run *r, int which, int solve) { return launch_rhs_common(r, which, 1, solve); }

int launch_update(vv_run *r, int kt, int from_mg)
{
    UpdArgs a;
    a.Cn = r->Cn; a.thn = r->thn; a.vel = r->vel; a.uel = r->uel;
    a.velw = r->vel; a.uelw = r->uel; a.velp = r->velp; a.uelp = r->uelp;
    a.u0v = r->mg.prob[0].lev[0].u; a.u0u = r->mg.prob[1].lev[0].u;
    a.from_mg = from_mg;
    a.rT = r->rT; a.geom = r->cfg.geom;
    a.mT2 = r->mT2; a.mT = r->mT; a.mU2 = r->mU2; a.mV2 = r->mV2;
    a.upart = r->upart;
    a.nr = r->nr; a.np = r->np; a.ld = r->ld; a.kt = kt; a.stats = r->cfg.stats;
    a.dr = r->dr; a.dp = r->dp; a.dt = r->cfg.dt;
    const long long n = (long long)(r->nr + 2) * (r->np + 1);
    hipEvent_t pe_;
    prof_begin(r, &pe_);
    hipLaunchKernelGGL(k_update, dim3(nblk(n, WG)), dim3(WG), 0, r->stream, a);
    // + solution scatter: read the solutions and old fields, write both levels
    prof_end(r, VV_K_UPDATE, pe_, 8.0 * ((r->cfg.stats ? 13.0 : 5.0) + (from_mg ? 6.0 : 0.0)) * (double)n);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

}  // namespace vv
