// vvt.hip -- the 3D horizontal-convection tank step on gfx950
// (include/vvtank.h).
//
// Explicit part: the compact advection derivatives as batched line solves
// (one thread per line, the elimination tables of the constant rows computed
// once on the host in the order of the reference's tdma), then the A, B, C
// and temperature updates in one pass (one thread per (i, j) column marching
// along y), one thread per point for the right-hand sides and the
// boundaries.  Every expression is evaluated in the operation order the
// oracle restates (the library is compiled with -ffp-contract=off), so the
// explicit results are bit-for-bit the oracle's.
//
// Poisson: each operator L = D_y + D_z + D_x (Cartesian, constant rows) is
// solved exactly: its y rows have cosine modes (u, w: mirrored at the
// stress-free side walls) or sine modes (v: zero there), applied by FFTs in
// LDS when np is a power of two (k_t_yfft_*) and as FP64 matrix-core GEMMs
// otherwise; the z rows are diagonalised once (the eigenbasis of
// rot3dmg.f's z rows, shared with the annulus solver); what remains per
// (y mode, z eigenvalue) is one banded system along x, eliminated by one
// thread (Thomas, the closure rows' third coefficient folded in).

#include "vv3_internal.h"
#include "vv_fft.h"
#include "vv_gemm.h"
#include "vv_reduce.h"
#include "vv_ytrans.h"
#include "vv_zsplit.h"

#include "../../include/vvtank.h"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace vvt {

namespace {

constexpr int WG = 256;
constexpr int KE_MAXB = 1024;

struct TGeo {
    int nr, nz, np, ldi, ldj;
    double dt, dr, dz, dp, Re, Ptl, Gr;
};

#define X(a, i, j, k) (a)[(unsigned)(((k) * g.ldj + (j)) * g.ldi + (i))]

inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// ------------------------------------------------------- compact lines
// One derivative job: field f along `axis` (0: x, 1: y, 2: z), end rows of
// `type` (0: nodes equally spaced, velvor2dperiodicrbconv.f:613-620; 1: walls
// half a spacing away, its 1085-1097), m unknowns, the lines' other two
// indices in [1, n1] x [1, n2] (x lines (j, k), y lines (i, j), z lines
// (i, k)), tables fact | bp | c | 1/bp (1-based, m+1 each), nl lines per
// workgroup (x lines, staged in LDS).
struct DJob {
    const double *f;
    double *d;
    const double *tab;
    int axis, type, m, n1, n2, nl;
    double dri;
};
struct DJobs {
    DJob j[12];
};

// x / b for a positive pivot b with rb = RN(1/b): q0 = x rb, t = b q0 - x
// (exact, one fma), q = q0 - t rb: the correctly rounded quotient
// (Markstein's correction), i.e. the IEEE quotient the reference's division
// gives, in three dependent operations.  -t keeps the sign of a zero
// quotient as IEEE division gives it.
__device__ __forceinline__ double qdiv(double x, double b, double rb)
{
    const double q0 = x * rb;
    const double t = __builtin_fma(b, q0, -x);
    return __builtin_fma(-t, rb, q0);
}

__device__ __forceinline__ double line_rhs(int type, int m, int i, double dri, double fm3, double fm1, double f0,
                                           double fp1, double f2, double f4)
{
    // f0 = f[i], fm1 = f[i-1], fp1 = f[i+1]; fm3 = f[m-3] and f4 = f[4] for
    // the end rows of type 1, f2 = f[2] for row 1 of type 0
    if (type == 0) {
        if (i == 1) return 3.0 * (f2 - fm1) * dri - (-1.5 * fm1 + 2.0 * f0 - 0.5 * f2) * dri;
        if (i == m) return 3.0 * (fp1 - fm1) * dri - (0.5 * fm1 - 2.0 * f0 + 1.5 * fp1) * dri;
        return 3.0 * (fp1 - fm1) * dri;
    }
    if (i == 1) return dri * (0.533333333 * fm1 - 3.6 * f0 + 2.966666667 * fp1 + 0.1 * f4);
    if (i == m) return dri * (-0.1 * fm3 - 2.966666667 * fm1 + 3.6 * f0 - 0.533333333 * fp1);
    return 3.0 * (fp1 - fm1) * dri;
}

// Lines along y and z: one thread per line, consecutive threads on
// consecutive i (coalesced).  The lines (256-514 rows) are too long to keep
// their eliminated right-hand sides r' in registers, so the forward sweep
// keeps only every CK-th one (a checkpoint, in ckp: 1/CK of the line), and
// the back substitution works segment by segment from the last: it reads
// the segment's nodes again, recomputes its r' from the checkpoint before it
// in registers -- the same operations on the same values, so the same bits
// -- and substitutes back through them.  Per line that is two reads of the
// nodes and one write of the derivative, against one read, a write and a
// read of r' and the final write when r' went through d.  The forward rows
// are processed in groups of U with their loads issued ahead of the chain.
constexpr int CK = 32;

// TLDS: the job's tables (fact | b' | c | 1/b', 4 (m+1) words) are copied to
// LDS first; lines longer than DYZ_LDS_MAXM read them from global memory
constexpr int DYZ_LDS_MAXM = 2047;

template <bool TLDS>
__global__ void __launch_bounds__(WG) k_t_deriv_yz(TGeo g, DJobs jobs, int j0, double *__restrict__ ckp, int nseg_max,
                                                   int nls)
{
    extern __shared__ double tabl[];
    const DJob J = jobs.j[j0 + blockIdx.y];
    if constexpr (TLDS) {
        for (int e = threadIdx.x; e < 4 * (J.m + 1); e += WG) tabl[e] = J.tab[e];
        __syncthreads();
    }
    const int t = blockIdx.x * WG + threadIdx.x;
    if (t >= J.n1 * J.n2) return;
    const int q = t / J.n1, p = t - q * J.n1;
    const int P1 = p + 1, Q1 = q + 1;
    unsigned base, s;
    if (J.axis == 1) { base = (unsigned)(Q1 * g.ldi + P1); s = (unsigned)(g.ldi * g.ldj); }
    else { base = (unsigned)(Q1 * g.ldj * g.ldi + P1); s = (unsigned)g.ldi; }
    const double *__restrict__ f = J.f;
    double *__restrict__ d = J.d;
    double *__restrict__ cp = ckp + (size_t)blockIdx.y * nseg_max * nls + t;   // segment sg at cp[sg * nls]
    const int m = J.m, type = J.type;
    const double dri = J.dri;
    const double *T = TLDS ? tabl : J.tab;
    const double *fact = T, *bp = T + (m + 1), *cc = T + 2 * (m + 1), *rbp = T + 3 * (m + 1);
    const double f4 = (type == 1) ? f[base + 4u * s] : 0.0;
    const double fm3 = (type == 1) ? f[base + (unsigned)(m - 3) * s] : 0.0;
    const double f2 = f[base + 2u * s];
    double fa = f[base], fb = f[base + s];   // f[i-1], f[i]
    double rp = 0.0;
    constexpr int U = 8;
    int i = 1;
    for (; i + U - 1 <= m; i += U) {
        double nx[U];   // f[i+1 .. i+U]
#pragma unroll
        for (int u = 0; u < U; ++u) nx[u] = f[base + (unsigned)(i + 1 + u) * s];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int ii = i + u;
            const double r = line_rhs(type, m, ii, dri, fm3, fa, fb, nx[u], f2, f4);
            rp = (ii == 1) ? r : r - fact[ii] * rp;
            if (ii % CK == 0) cp[(size_t)(ii / CK - 1) * nls] = rp;
            fa = fb;
            fb = nx[u];
        }
    }
    for (; i <= m; ++i) {
        const double fc = f[base + (unsigned)(i + 1) * s];
        const double r = line_rhs(type, m, i, dri, fm3, fa, fb, fc, f2, f4);
        rp = (i == 1) ? r : r - fact[i] * rp;
        if (i % CK == 0) cp[(size_t)(i / CK - 1) * nls] = rp;
        fa = fb;
        fb = fc;
    }
    double y = 0.0;
    for (int sg = (m + CK - 1) / CK - 1; sg >= 0; --sg) {
        const int r0 = sg * CK + 1, r1 = min(r0 + CK - 1, m);
        double rq = (sg == 0) ? 0.0 : cp[(size_t)(sg - 1) * nls];
        double fv[CK + 2];   // nodes r0-1 .. r0+CK
#pragma unroll
        for (int u = 0; u < CK + 2; ++u) {
            const int row = r0 - 1 + u;
            fv[u] = (row <= r1 + 1) ? f[base + (unsigned)row * s] : 0.0;
        }
        double rr[CK];
#pragma unroll
        for (int u = 0; u < CK; ++u) {
            const int ii = r0 + u;
            rr[u] = 0.0;
            if (ii <= r1) {
                const double r = line_rhs(type, m, ii, dri, fm3, fv[u], fv[u + 1], fv[u + 2], f2, f4);
                rq = (ii == 1) ? r : r - fact[ii] * rq;
                rr[u] = rq;
            }
        }
#pragma unroll
        for (int u = CK - 1; u >= 0; --u) {
            const int ii = r0 + u;
            if (ii <= r1) {
                y = (ii == m) ? qdiv(rr[u], bp[m], rbp[m]) : qdiv(rr[u] - cc[ii] * y, bp[ii], rbp[ii]);
                d[base + (unsigned)ii * s] = y;
            }
        }
    }
}

// Lines along x (contiguous in memory): a workgroup of one wave takes 64
// lines (consecutive j of one plane k), thread l line l.  The lines pass
// through LDS in chunks of CX rows (nodes r0-1 .. r0+CX, odd pitch, loaded
// coalesced along the lines): the forward sweep runs chunk by chunk and
// keeps one checkpoint of the eliminated right-hand side per chunk (ckp, as
// the y/z lines); the back substitution takes the chunks from the last,
// reloads one, recomputes its r' from the checkpoint before it in registers
// (the same operations, the same bits), substitutes back into the tile and
// stores the chunk's rows coalesced.
constexpr int CX = 20, XSP = CX + 3;   // rows per chunk; tile pitch (odd)

__global__ void __launch_bounds__(64) k_t_deriv_x(TGeo g, DJobs jobs, double *__restrict__ ckp, int nseg_max, int nls)
{
    __shared__ double sh[64 * XSP];
    __shared__ double tb[4][CX];   // the chunk's table rows: fact, b', c, 1/b'
    const DJob J = jobs.j[3 * blockIdx.y];
    const int m = J.m, type = J.type;
    const int tiles = (J.n1 + 63) / 64;
    const int q = blockIdx.x / tiles, p0 = (blockIdx.x - q * tiles) * 64;
    if (q >= J.n2) return;
    const int nlines = min(64, J.n1 - p0);
    const int tid = threadIdx.x;
    // line l: (j, k) = (p0 + l + 1, q + 1), nodes contiguous from node 0
    const unsigned base0 = (unsigned)(((q + 1) * g.ldj + (p0 + 1)) * g.ldi);
    const double *__restrict__ f = J.f;
    double *__restrict__ d = J.d;
    const bool act = tid < nlines;
    const unsigned mybase = base0 + (unsigned)(tid * g.ldi);
    double *__restrict__ cp = ckp + (size_t)blockIdx.y * nseg_max * nls + (size_t)q * J.n1 + p0 + tid;
    const double dri = J.dri;
    const double *fact = J.tab, *bp = J.tab + (m + 1), *cc = J.tab + 2 * (m + 1), *rbp = J.tab + 3 * (m + 1);
    const double f4 = (act && type == 1) ? f[mybase + 4u] : 0.0;
    const double fm3 = (act && type == 1) ? f[mybase + (unsigned)(m - 3)] : 0.0;
    const double f2 = act ? f[mybase + 2u] : 0.0;
    const int nseg = (m + CX - 1) / CX;
    double *L = sh + tid * XSP;
    // nodes r0-1 .. r1+1 of chunk sg into the tile (coalesced: consecutive
    // threads on consecutive nodes of a line)
    auto load = [&](int sg) {
        const int r0 = sg * CX + 1, r1 = min(r0 + CX - 1, m), len = r1 - r0 + 3;
        const int total = nlines * len;
        for (int e0 = 0; e0 < total; e0 += 8 * 64) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + u * 64 + tid;
                if (e < total) {
                    const int l = e / len, o = e - l * len;
                    v[u] = f[base0 + (unsigned)(l * g.ldi + r0 - 1 + o)];
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + u * 64 + tid;
                if (e < total) {
                    const int l = e / len, o = e - l * len;
                    sh[l * XSP + o] = v[u];
                }
            }
        }
        if (tid < CX) {   // table rows r0 .. r0+CX-1 (only rows <= r1 are read)
            const int ii = min(r0 + tid, m);
            tb[0][tid] = fact[ii];
            tb[1][tid] = bp[ii];
            tb[2][tid] = cc[ii];
            tb[3][tid] = rbp[ii];
        }
        __syncthreads();
    };
    double rp = 0.0;
    for (int sg = 0; sg < nseg; ++sg) {
        load(sg);
        const int r0 = sg * CX + 1, r1 = min(r0 + CX - 1, m);
        if (act) {
            double fv[CX + 2];
#pragma unroll
            for (int u = 0; u < CX + 2; ++u) fv[u] = L[u];
#pragma unroll
            for (int u = 0; u < CX; ++u) {
                const int ii = r0 + u;
                if (ii <= r1) {
                    const double r = line_rhs(type, m, ii, dri, fm3, fv[u], fv[u + 1], fv[u + 2], f2, f4);
                    rp = (ii == 1) ? r : r - tb[0][u] * rp;
                }
            }
            if (sg + 1 < nseg) cp[(size_t)sg * nls] = rp;
        }
        __syncthreads();
    }
    double y = 0.0;
    for (int sg = nseg - 1; sg >= 0; --sg) {
        load(sg);
        const int r0 = sg * CX + 1, r1 = min(r0 + CX - 1, m);
        if (act) {
            double rq = (sg == 0) ? 0.0 : cp[(size_t)(sg - 1) * nls];
            double fv[CX + 2], rr[CX];
#pragma unroll
            for (int u = 0; u < CX + 2; ++u) fv[u] = L[u];
#pragma unroll
            for (int u = 0; u < CX; ++u) {
                const int ii = r0 + u;
                rr[u] = 0.0;
                if (ii <= r1) {
                    const double r = line_rhs(type, m, ii, dri, fm3, fv[u], fv[u + 1], fv[u + 2], f2, f4);
                    rq = (ii == 1) ? r : r - tb[0][u] * rq;
                    rr[u] = rq;
                }
            }
#pragma unroll
            for (int u = CX - 1; u >= 0; --u) {
                const int ii = r0 + u;
                if (ii <= r1) {
                    y = (ii == m) ? qdiv(rr[u], tb[1][u], tb[3][u])
                                  : qdiv(rr[u] - tb[2][u] * y, tb[1][u], tb[3][u]);
                    L[u + 1] = y;   // row ii sits at tile position ii - r0 + 1
                }
            }
        }
        __syncthreads();
        const int len = r1 - r0 + 1, total = nlines * len;
        for (int e = tid; e < total; e += 64) {
            const int l = e / len, o = e - l * len;
            d[base0 + (unsigned)(l * g.ldi + r0 + o)] = sh[l * XSP + 1 + o];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------- explicit step
struct VF {
    const double *A, *B, *C, *Ap, *Bp, *Cp, *th, *thp, *uel, *vel, *wel;
    const double *dv[12];
    double *An, *Bn, *Cn, *thn;
};

// The explicit updates, fused: Avorticity (rot3dmg.f:359-465; i = 1..nr,
// j = 1..nz-1, k = 1..np-1), Bvorticity (495-590; i = 1..nr-1, j = 1..nz-1,
// k = 1..np), Cvorticity (619-725; i = 1..nr-1, j = 1..nz, k = 1..np-1) and
// temperature (754-871; i = 1..nr, j = 1..nz, k = 1..np), in Cartesian form.
// MASK 1 runs the three vorticity updates, 2 the temperature, 3 all four:
// they read only the current and previous time levels, so one pass reads
// the fields the four share once.
//
// One thread per (i, j) column marches over KC consecutive y planes.  The
// values a column needs at several planes (its own node and the few
// neighbours the staggered averages take across y) stay in registers as the
// march moves on: each is loaded once per plane, not at k-1, k and k+1.
// Every expression is the routine's own, in its order (same bits).
//
// Tiles go to the XCDs in contiguous ranges (the hardware deals linear
// workgroup ids to the 8 XCDs in turn): the row blocks of a chunk of planes
// run together on one XCD, so the rows j-1 and j+1 a block reads were just
// fetched into that XCD's L2 by its neighbours.
constexpr int KC = 16;

__device__ __forceinline__ void xcd_tile(unsigned nbx, unsigned nby, int &bx, int &by)
{
    const unsigned T = nbx * nby, L = blockIdx.y * nbx + blockIdx.x;
    const unsigned x = L & 7u, q = L >> 3, base = T >> 3, rem = T & 7u;
    const unsigned W = x * base + (x < rem ? x : rem) + q;
    by = (int)(W / nbx);
    bx = (int)(W - (unsigned)by * nbx);
}

// The divisors the updates share, with their correctly rounded reciprocals
// (computed on the host: uniform, they stay in scalar registers); x / b is
// evaluated as qdiv(x, b, 1/b), the same quotient.  c25r and c25z are the
// closure rows' -2.5 / dr^2 and -2.5 / dz^2.
struct XDiv {
    double dr2, rdr2, dz2, rdz2, dp2, rdp2, pdt, rpdt, ptl2, rptl2, rdr, rdz, rdp, rptl, rdt, c25r, c25z;
};

XDiv xdiv(const TGeo &g)
{
    XDiv d;
    d.dr2 = g.dr * g.dr; d.dz2 = g.dz * g.dz; d.dp2 = g.dp * g.dp;
    d.pdt = g.Ptl * g.dt; d.ptl2 = 2.0 * g.Ptl;
    d.rdr2 = 1.0 / d.dr2; d.rdz2 = 1.0 / d.dz2; d.rdp2 = 1.0 / d.dp2; d.rpdt = 1.0 / d.pdt; d.rptl2 = 1.0 / d.ptl2;
    d.rdr = 1.0 / g.dr; d.rdz = 1.0 / g.dz; d.rdp = 1.0 / g.dp; d.rptl = 1.0 / g.Ptl; d.rdt = 1.0 / g.dt;
    d.c25r = -2.5 / d.dr2; d.c25z = -2.5 / d.dz2;
    return d;
}

template <int MASK>
__global__ void __launch_bounds__(WG) k_t_explicit(TGeo g, VF F, XDiv q)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    int bx, kc;
    xcd_tile(gridDim.x, gridDim.y, bx, kc);
    const int t = bx * WG + threadIdx.x;
    if (t >= nr * nz) return;
    const int jj = t / nr, i = 1 + (t - jj * nr), j = 1 + jj;
    const int k0 = 1 + kc * KC, k1 = min(np, k0 + KC - 1);
    const double dr = g.dr, dz = g.dz, dp = g.dp, dt = g.dt, Ptl = g.Ptl, Re = g.Re, Gr = g.Gr;
    const double dr2 = q.dr2, dz2 = q.dz2, dp2 = q.dp2, pdt = q.pdt, ptl2 = q.ptl2;
    const double rdr2 = q.rdr2, rdz2 = q.rdz2, rdp2 = q.rdp2, rpdt = q.rpdt, rptl2 = q.rptl2;
    const double rdr = q.rdr, rdz = q.rdz, rdp = q.rdp, rptl = q.rptl, rdt = q.rdt;
    const double *__restrict__ A = F.A, *__restrict__ B = F.B, *__restrict__ C = F.C, *__restrict__ th = F.th;
    const double *__restrict__ uel = F.uel, *__restrict__ vel = F.vel, *__restrict__ wel = F.wel;
    const bool doA = (MASK & 1) && j <= nz - 1;
    const bool doB = (MASK & 1) && i <= nr - 1 && j <= nz - 1;
    const bool doC = (MASK & 1) && i <= nr - 1;
    const unsigned P = (unsigned)(g.ldi * g.ldj), li = (unsigned)g.ldi;
    const unsigned o = (unsigned)j * li + (unsigned)i;   // (i, j) within a plane
    // in-plane offsets of the neighbours
    const unsigned oim = o - 1u, oip = o + 1u, ojm = o - li, ojp = o + li;
    const unsigned oimjp = ojp - 1u, oipjp = ojp + 1u, oipjm = ojm + 1u;

    // registers carried along y: <field><position><plane>, plane M = k-1,
    // K = k, N = k+1; positions 0 = (i, j), and the neighbours named
    unsigned b = (unsigned)(k0 - 1) * P;   // plane k0-1
    double aM = A[b + o], apM = A[b + oip];
    double bM = B[b + o], cM = C[b + o], cjM = C[b + ojp], tM = th[b + o];
    double vM00 = vel[b + o], vM01 = vel[b + ojp], vM10 = vel[b + oip], vM11 = vel[b + oipjp];
    b += P;   // plane k0
    double aK = A[b + o], bK = B[b + o], bmK = B[b + oim], bjK = B[b + ojm];
    double cK = C[b + o], tK = th[b + o], tjK = th[b + ojp];
    double uK00 = uel[b + o], uK01 = uel[b + ojp], uK10 = uel[b + oim], uK11 = uel[b + oimjp];
    double wK00 = wel[b + o], wK10 = wel[b + oip], wK0m = wel[b + ojm], wK1m = wel[b + oipjm];

    for (int k = k0; k <= k1; ++k) {
        const unsigned bk = (unsigned)k * P, bn = bk + P;
        // the planes the march brings in
        const double aN = A[bn + o], apK = A[bk + oip];
        const double bN = B[bn + o], bmN = B[bn + oim], bjN = B[bn + ojm];
        const double cN = C[bn + o], cjK = C[bk + ojp];
        const double tN = th[bn + o], tjN = th[bn + ojp];
        const double uN00 = uel[bn + o], uN01 = uel[bn + ojp], uN10 = uel[bn + oim], uN11 = uel[bn + oimjp];
        const double vK00 = vel[bk + o], vK01 = vel[bk + ojp], vK10 = vel[bk + oip], vK11 = vel[bk + oipjp];
        const double wN00 = wel[bn + o], wN10 = wel[bn + oip], wN0m = wel[bn + ojm], wN1m = wel[bn + oipjm];

        if (doA && k <= np - 1) {
            const double u = (uK00 + uK01 + uN00 + uN01 + uK10 + uK11 + uN10 + uN11) * 0.125;
            const double v = (vK01 + vK00) * 0.5;
            const double w = (wN00 + wK00) * 0.5;
            const double BB = (bK + bmK + bN + bmN) * 0.25;
            const double CC = (cK + cjK + C[bk + oim] + C[bk + oimjp]) * 0.25;
            double Arra, cnr;
            if (i == 1) {
                Arra = qdiv(3.2 * A[bk + oim] + 2.0 * apK - 0.2 * A[bk + o + 2u], dr2, rdr2);
                cnr = q.c25r;
            } else if (i == nr) {
                Arra = qdiv(3.2 * apK + 2.0 * A[bk + oim] - 0.2 * A[bk + o - 2u], dr2, rdr2);
                cnr = q.c25r;
            } else {
                Arra = qdiv(A[bk + oim] + apK, dr2, rdr2);
                cnr = -rdr2;
            }
            const double uelr = qdiv((uK00 + uK01 + uN00 + uN01) * 0.25 - (uK10 + uK11 + uN10 + uN11) * 0.25, dr, rdr);
            const double Azza = qdiv(A[bk + ojm] + A[bk + ojp], dz2, rdz2);
            const double uelz = qdiv((uK01 + uN01 + uK11 + uN11) * 0.25 - (uK00 + uN00 + uK10 + uN10) * 0.25, dz, rdz);
            const double thh = qdiv((tN + tjN - tK - tjK) * 0.5, dp, rdp);
            const double Ahha = qdiv(aN + aM, dp2, rdp2);
            const double uelh = qdiv((uN00 + uN01 + uN10 + uN11) * 0.25 - (uK00 + uK01 + uK10 + uK11) * 0.25, dp, rdp);
            const double cnz = -rdz2;
            const double cnh = -rdp2;
            const double Ax = F.dv[0][bk + o], Ay = F.dv[1][bk + o], Az = F.dv[2][bk + o];
            const double RHS1 = Arra + Ahha + Azza + qdiv(aK, pdt, rpdt);
            const double ULT = qdiv(u * Ax + v * Ay + w * Az - uelh * BB - uelz * CC, Ptl, rptl);
            const double RHS2 = -2.0 * Re * uelz - Gr * Ptl * thh;
            const double cf1 = rpdt - qdiv(uelr, ptl2, rptl2) - cnr - cnz - cnh;
            const double cf2 = qdiv(uelr, ptl2, rptl2) + cnr + cnz + cnh;
            F.An[bk + o] = (cf2 * F.Ap[bk + o] - ULT - RHS2 + RHS1) / cf1;
        }
        if (doB) {
            const double u = (uK00 + uK01) * 0.5;
            const double v = (vM00 + vM01 + vK00 + vK01 + vM10 + vM11 + vK10 + vK11) * 0.125;
            const double w = (wK00 + wK10) * 0.5;
            const double AA = (aK + apK + aM + apM) * 0.25;
            const double CC = (cK + cjK + cM + cjM) * 0.25;
            const double Brra = qdiv(bmK + B[bk + oip], dr2, rdr2);
            const double velr = qdiv((vM10 + vM11 + vK10 + vK11) * 0.25 - (vM00 + vM01 + vK00 + vK01) * 0.25, dr, rdr);
            const double Bzza = qdiv(bjK + B[bk + ojp], dz2, rdz2);
            const double velz = qdiv((vM01 + vM11 + vK01 + vK11) * 0.25 - (vM00 + vM10 + vK00 + vK10) * 0.25, dz, rdz);
            const double thr = qdiv((th[bk + oipjp] + th[bk + oip] - tjK - tK) * 0.5, dr, rdr);
            const double Bhha = qdiv(bN + bM, dp2, rdp2);
            const double velh = qdiv((vK00 + vK10 + vK01 + vK11) * 0.25 - (vM00 + vM10 + vM01 + vM11) * 0.25, dp, rdp);
            const double cnr = -rdr2;
            const double cnz = -rdz2;
            const double cnh = -rdp2;
            const double Bx = F.dv[3][bk + o], By = F.dv[4][bk + o], Bz = F.dv[5][bk + o];
            const double RHS1 = Brra + Bhha + Bzza + qdiv(bK, pdt, rpdt);
            const double ULT = qdiv(u * Bx + v * By + w * Bz - velr * AA - velz * CC, Ptl, rptl);
            const double RHS2 = -2.0 * Re * velz + Gr * Ptl * thr;
            const double cf1 = rpdt - qdiv(velh, ptl2, rptl2) - cnr - cnz - cnh;
            const double cf2 = qdiv(velh, ptl2, rptl2) + cnr + cnz + cnh;
            F.Bn[bk + o] = (cf2 * F.Bp[bk + o] - ULT - RHS2 + RHS1) / cf1;
        }
        if (doC && k <= np - 1) {
            const double u = (uN00 + uK00) * 0.5;
            const double v = (vK00 + vK10) * 0.5;
            const double w = (wK00 + wK10 + wN00 + wN10 + wK0m + wK1m + wN0m + wN1m) * 0.125;
            const double AA = (aK + A[bk + ojm] + apK + A[bk + oipjm]) * 0.25;
            const double BB = (bK + bjK + bN + bjN) * 0.25;
            const double Crra = qdiv(C[bk + oim] + C[bk + oip], dr2, rdr2);
            const double welr = qdiv((wK10 + wN10 + wK1m + wN1m) * 0.25 - (wK00 + wN00 + wK0m + wN0m) * 0.25, dr, rdr);
            double Czza, cnz;
            if (j == 1) {
                Czza = qdiv(3.2 * C[bk + ojm] + 2.0 * cjK - 0.2 * C[bk + ojp + li], dz2, rdz2);
                cnz = q.c25z;
            } else if (j == nz) {
                Czza = qdiv(3.2 * cjK + 2.0 * C[bk + ojm] - 0.2 * C[bk + ojm - li], dz2, rdz2);
                cnz = q.c25z;
            } else {
                Czza = qdiv(C[bk + ojm] + cjK, dz2, rdz2);
                cnz = -rdz2;
            }
            const double welz = qdiv((wK00 + wK10 + wN00 + wN10) * 0.25 - (wK0m + wK1m + wN0m + wN1m) * 0.25, dz, rdz);
            const double Chha = qdiv(cN + cM, dp2, rdp2);
            const double welh = qdiv((wN00 + wN10 + wN0m + wN1m) * 0.25 - (wK00 + wK10 + wK0m + wK1m) * 0.25, dp, rdp);
            const double cnr = -rdr2;
            const double cnh = -rdp2;
            const double Cx = F.dv[6][bk + o], Cy = F.dv[7][bk + o], Cz = F.dv[8][bk + o];
            const double RHS1 = Crra + Chha + Czza + qdiv(cK, pdt, rpdt);
            const double ULT = qdiv(u * Cx + v * Cy + w * Cz - welr * AA - welh * BB, Ptl, rptl);
            const double RHS2 = -2.0 * Re * welz;
            const double cf1 = rpdt - qdiv(welz, ptl2, rptl2) - cnr - cnz - cnh;
            const double cf2 = qdiv(welz, ptl2, rptl2) + cnr + cnz + cnh;
            F.Cn[bk + o] = (cf2 * F.Cp[bk + o] - ULT - RHS2 + RHS1) / cf1;
        }
        if (MASK & 2) {
            const double u = (uK00 + uK10) * 0.5;
            const double v = (vK00 + vM00) * 0.5;
            const double w = (wK00 + wK0m) * 0.5;
            const double thim = th[bk + oim], thip = th[bk + oip], thjm = th[bk + ojm];
            double thrra, cnr, thzza, cnz;
            if (i == 1) {
                thrra = qdiv(3.2 * thim + 2.0 * thip - 0.2 * th[bk + o + 2u], dr2, rdr2);
                cnr = q.c25r;
            } else if (i == nr) {
                thrra = qdiv(3.2 * thip + 2.0 * thim - 0.2 * th[bk + o - 2u], dr2, rdr2);
                cnr = q.c25r;
            } else {
                thrra = qdiv(thim + thip, dr2, rdr2);
                cnr = -rdr2;
            }
            if (j == 1) {
                thzza = qdiv(3.2 * thjm + 2.0 * tjK - 0.2 * th[bk + ojp + li], dz2, rdz2);
                cnz = q.c25z;
            } else if (j == nz) {
                thzza = qdiv(3.2 * tjK + 2.0 * thjm - 0.2 * th[bk + ojm - li], dz2, rdz2);
                cnz = q.c25z;
            } else {
                thzza = qdiv(thjm + tjK, dz2, rdz2);
                cnz = -rdz2;
            }
            const double thhha = qdiv(tN + tM, dp2, rdp2);
            const double cnp = -rdp2;
            const double Tx = F.dv[9][bk + o], Ty = F.dv[10][bk + o], Tz = F.dv[11][bk + o];
            const double RHS1 = thrra + thhha + thzza + qdiv(tK, dt, rdt);
            const double ULT = u * Tx + v * Ty + w * Tz;
            const double cf1 = rdt - cnr - cnz - cnp;
            const double cf2 = cnr + cnz + cnp;
            F.thn[bk + o] = (cf2 * F.thp[bk + o] - ULT + RHS1) / cf1;
        }
        // move the march on one plane
        aM = aK; aK = aN; apM = apK;
        bM = bK; bK = bN; bmK = bmN; bjK = bjN;
        cM = cK; cK = cN; cjM = cjK;
        tM = tK; tK = tN; tjK = tjN;
        uK00 = uN00; uK01 = uN01; uK10 = uN10; uK11 = uN11;
        vM00 = vK00; vM01 = vK01; vM10 = vK10; vM11 = vK11;
        wK00 = wN00; wK10 = wN10; wK0m = wN0m; wK1m = wN1m;
    }
}

// heat(trans1,trans2), velvor2dhorconvtank.f:1589-1620.  Its k-sum adds, for
// each k, the row's term at k and at k+1, each 0.5 dp (Fornberg row) dzi;
// k_t_heat_terms evaluates every (i, k) term once, in parallel (the term at
// k+1 of step k is the term at k of step k+1: the same operations), and
// k_t_heat_cols adds them per x column in the routine's order.
__global__ void __launch_bounds__(WG) k_t_heat_terms(TGeo g, const double *th, double *E)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const int i = 1 + blockIdx.x * WG + threadIdx.x, kk = blockIdx.y, k = np / 2 + kk;
    if (i > nr) return;
    const double dp = g.dp, dzi = 1.0 / g.dz;
    const size_t n = (size_t)(np - np / 2 + 1) * nr;   // rows k = np/2 .. np
    E[(size_t)kk * nr + (i - 1)] = 0.5 * dp *
                                    (-3.35238095 * X(th, i, 0, k) + 4.375 * X(th, i, 1, k) -
                                     1.45833333 * X(th, i, 2, k) + 0.525 * X(th, i, 3, k) -
                                     0.08928571 * X(th, i, 4, k)) *
                                    dzi;
    E[n + (size_t)kk * nr + (i - 1)] = 0.5 * dp *
                                        (-3.35238095 * X(th, i, nz + 1, k) + 4.375 * X(th, i, nz, k) -
                                         1.45833333 * X(th, i, nz - 1, k) + 0.525 * X(th, i, nz - 2, k) -
                                         0.08928571 * X(th, i, nz - 3, k)) *
                                        dzi;
}

__global__ void __launch_bounds__(WG) k_t_heat_cols(TGeo g, const double *E, double *cols)
{
    const int nr = g.nr, np = g.np;
    const int i = blockIdx.x * WG + threadIdx.x;
    if (i >= nr) return;
    const size_t n = (size_t)(np - np / 2 + 1) * nr;
    double trans1 = 0.0, trans2 = 0.0;
    for (int kk = 0; kk < np - np / 2; ++kk) {
        const size_t o = (size_t)kk * nr + i;
        trans1 = trans1 + E[o] + E[o + nr];
        trans2 = trans2 + E[n + o] + E[n + o + nr];
    }
    cols[2 * i] = trans1;
    cols[2 * i + 1] = trans2;
}

__global__ void k_t_heat_sum(int nr, double dr, const double *cols, double *out)
{
    if (threadIdx.x != 0) return;
    double s1 = 0.0, s2 = 0.0;
    for (int i = 0; i < nr; ++i) {
        s1 = s1 + cols[2 * i] * dr;
        s2 = s2 + cols[2 * i + 1] * dr;
    }
    out[0] = s1;
    out[1] = s2;
}

// rhsw / rhsv / rhsu (rot3dmg.f:1026-1134, Cartesian)
struct RF {
    const double *An, *Bn, *Cn;
    double *wrh, *vrh, *urh;
};
__global__ void __launch_bounds__(WG) k_t_rhs(TGeo g, RF F, int which)
{
    const int nr = g.nr, nz = g.nz;
    const double dr = g.dr, dz = g.dz, dp = g.dp;
    const int ni = (which == 3) ? nr - 1 : nr;
    const int nj = (which == 1) ? nz - 1 : nz;
    const int t = blockIdx.x * WG + threadIdx.x;
    if (t >= ni * nj) return;
    const int jj = t / ni, i = 1 + (t - jj * ni), j = 1 + jj, k = 1 + blockIdx.y;
    const double *An = F.An, *Bn = F.Bn, *Cn = F.Cn;
    if (which == 1) {
        const double Ah = (X(An, i, j, k) - X(An, i, j, k - 1)) / dp;
        const double Br = (X(Bn, i, j, k) - X(Bn, i - 1, j, k)) / dr;
        double v;
        if (i == 1) v = Ah - X(Bn, 1, j, k) / dr;
        else if (i == nr) v = Ah + X(Bn, nr - 1, j, k) / dr;
        else v = Ah - Br;
        X(F.wrh, i, j, k) = v;
    } else if (which == 2) {
        const double Az = (X(An, i, j, k) - X(An, i, j - 1, k)) / dz;
        const double Cr = (X(Cn, i, j, k) - X(Cn, i - 1, j, k)) / dr;
        double v;
        if (i == 1) v = X(Cn, 1, j, k) / dr;
        else if (i == nr) v = -X(Cn, nr - 1, j, k) / dr;
        else v = Cr;
        if (j == 1) v = v - X(An, i, 1, k) / dz;
        else if (j == nz) v = v + X(An, i, nz - 1, k) / dz;
        else v = v - Az;
        X(F.vrh, i, j, k) = v;
    } else {
        const double Bz = (X(Bn, i, j, k) - X(Bn, i, j - 1, k)) / dz;
        const double Ch = (X(Cn, i, j, k) - X(Cn, i, j, k - 1)) / dp;
        double v;
        if (j == 1) v = X(Bn, i, 1, k) / dz - Ch;
        else if (j == nz) v = -X(Bn, i, nz - 1, k) / dz - Ch;
        else v = Bz - Ch;
        X(F.urh, i, j, k) = v;
    }
}

// ------------------------------------------------------------ update
struct UF {
    const double *uel, *vel, *wel;
    double *An, *Bn, *Cn, *thn;
};

// wall vorticity at the z walls (rot3dmg.f:980-987) and x walls (989-996),
// one thread per (wall point, k); blockIdx.y = k - 1 (k = 1..np)
__global__ void __launch_bounds__(WG) k_t_upd_walls(TGeo g, UF F)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const double dr = g.dr, dz = g.dz;
    const double *uel = F.uel, *vel = F.vel, *wel = F.wel;
    const int t = blockIdx.x * WG + threadIdx.x, k = 1 + blockIdx.y;
    // t < nr: z walls of A (i = t+1) and B (i = t+1 < nr)
    if (t < nr) {
        const int i = t + 1;
        if (k <= np - 1) {
            X(F.An, i, 0, k) = -(3.0 * X(vel, i, 1, k) - X(vel, i, 2, k) / 3.0) / dz;
            X(F.An, i, nz, k) = -(X(vel, i, nz - 1, k) / 3.0 - 3.0 * X(vel, i, nz, k)) / dz;
        }
        if (i <= nr - 1) {
            X(F.Bn, i, 0, k) = (3.0 * X(uel, i, 1, k) - X(uel, i, 2, k) / 3.0) / dz;
            X(F.Bn, i, nz, k) = (-3.0 * X(uel, i, nz, k) + X(uel, i, nz - 1, k) / 3.0) / dz;
        }
    } else if (t < nr + nz) {
        const int j = t - nr + 1;   // 1..nz
        if (j <= nz - 1) {
            X(F.Bn, 0, j, k) = -(3.0 * X(wel, 1, j, k) - X(wel, 2, j, k) / 3.0) / dr;
            X(F.Bn, nr, j, k) = -(X(wel, nr - 1, j, k) / 3.0 - 3.0 * X(wel, nr, j, k)) / dr;
        }
        if (k <= np - 1) {
            X(F.Cn, 0, j, k) = (3.0 * X(vel, 1, j, k) - X(vel, 2, j, k) / 3.0) / dr;
            X(F.Cn, nr, j, k) = (X(vel, nr - 1, j, k) / 3.0 - 3.0 * X(vel, nr, j, k)) / dr;
        }
    }
}

// side walls, stress free: A = C = 0 (velvor2dhorconvtank.f:1720-1727 for
// C), B mirrored into the ghost planes; one thread per (i, j)
__global__ void __launch_bounds__(WG) k_t_upd_side(TGeo g, UF F)
{
    const int np = g.np;
    const int t = blockIdx.x * WG + threadIdx.x;
    if (t >= g.ldi * g.ldj) return;
    const int j = t / g.ldi, i = t - j * g.ldi;
    X(F.An, i, j, 0) = 0.0;
    X(F.An, i, j, np) = 0.0;
    X(F.Cn, i, j, 0) = 0.0;
    X(F.Cn, i, j, np) = 0.0;
    X(F.Bn, i, j, 0) = X(F.Bn, i, j, 1);
    X(F.Bn, i, j, np + 1) = X(F.Bn, i, j, np);
}

// temperature boundaries, velvor2dhorconvtank.f:1728-1744, in three phases
// (each reads what the previous one wrote): 0: the insulated bottom and the
// linear surface (i = 1..nr, k = 1..np); 1: the insulated x walls
// (j = 0..nz+1, k = 1..np); 2: the insulated side walls (all i, j)
__global__ void __launch_bounds__(WG) k_t_upd_temp(TGeo g, double *thn, int phase)
{
    const int nr = g.nr, nz = g.nz, np = g.np;
    const int t = blockIdx.x * WG + threadIdx.x;
    if (phase == 0) {
        if (t >= nr * np) return;
        const int kk = t / nr, i = 1 + (t - kk * nr), k = 1 + kk;
        X(thn, i, 0, k) =
            -30.0 / 92.0 * (-3.75 * X(thn, i, 1, k) + 5.0 / 6.0 * X(thn, i, 2, k) - 3.0 / 20.0 * X(thn, i, 3, k));
        X(thn, i, nz + 1, k) = (double)k / (double)np * 1.0;
    } else if (phase == 1) {
        if (t >= (nz + 2) * np) return;
        const int kk = t / (nz + 2), j = t - kk * (nz + 2), k = 1 + kk;
        X(thn, 0, j, k) =
            -30.0 / 92.0 * (-3.75 * X(thn, 1, j, k) + 5.0 / 6.0 * X(thn, 2, j, k) - 3.0 / 20.0 * X(thn, 3, j, k));
        X(thn, nr + 1, j, k) = -30.0 / 92.0 *
                               (-3.75 * X(thn, nr, j, k) + 5.0 / 6.0 * X(thn, nr - 1, j, k) -
                                3.0 / 20.0 * X(thn, nr - 2, j, k));
    } else {
        if (t >= g.ldi * g.ldj) return;
        const int j = t / g.ldi, i = t - j * g.ldi;
        X(thn, i, j, 0) = 2.0 / 1.5 * X(thn, i, j, 1) - 1.0 / 3.0 * X(thn, i, j, 2);
        X(thn, i, j, np + 1) = 2.0 / 1.5 * X(thn, i, j, np) - 1.0 / 3.0 * X(thn, i, j, np - 1);
    }
}

// rot3dmg.f:1003-1015 as a rotation of three buffers per quantity: the old
// Ap buffer, which becomes the next An, receives a copy of An (so every
// level holds what the reference's copies leave)
struct CopyF {
    const double *src[4];
    double *dst[4];
};
__global__ void __launch_bounds__(WG) k_t_copy(CopyF F, size_t n)
{
    const size_t t = (size_t)blockIdx.x * WG + threadIdx.x;
    if (t >= n) return;
    const int q = blockIdx.y;
    F.dst[q][t] = F.src[q][t];
}

// ------------------------------------------------------------- Poisson
// y transform: S[q][u] = sum_k W[q][k] rhs(i, j, k+1), u = (j-1) i1 + (i-1)
struct YFwd {
    const double *W;
    int M, N, K;
    const double *rhs;
    double *S;
    int i1, ldi, ldj;
    __device__ double a(int, int m, int k) const { return W[(size_t)m * K + k]; }
    __device__ double b(int, int k, int n) const
    {
        const int j = n / i1, i = n - j * i1;
        return rhs[((size_t)(k + 1) * ldj + (size_t)(j + 1)) * ldi + (size_t)(i + 1)];
    }
    __device__ void c(int, int m, int n, double v) const { S[(size_t)m * N + n] = v; }
};
// z transform T[q][i][l] = sum_j S[q][j][i] Vinv[l][j]
struct ZFwd {
    const double *Vinv;
    int M, N, K;
    const double *S;
    double *T;
    int mt;
    __device__ double a(int q, int m, int k) const { return S[(size_t)q * mt + (size_t)k * M + m]; }
    __device__ double b(int, int k, int n) const { return Vinv[(size_t)n * K + k]; }
    __device__ void c(int q, int m, int n, double v) const { T[(size_t)q * mt + (size_t)m * N + n] = v; }
};
// S[q][j][i] = sum_l V[j][l] T[q][i][l]
struct ZInv {
    const double *V;
    int M, N, K;
    const double *T;
    double *S;
    int mt;
    __device__ double a(int, int m, int k) const { return V[(size_t)m * K + k]; }
    __device__ double b(int q, int k, int n) const { return T[(size_t)q * mt + (size_t)n * K + k]; }
    __device__ void c(int q, int m, int n, double v) const { S[(size_t)q * mt + (size_t)m * N + n] = v; }
};
// psi(i, j, k+1) = sum_q Wi[k][q] S[q][u]
struct YInv {
    const double *Wi;
    int M, N, K;
    const double *S;
    double *psi;
    int i1, ldi, ldj;
    __device__ double a(int, int m, int k) const { return Wi[(size_t)m * K + k]; }
    __device__ double b(int, int k, int n) const { return S[(size_t)k * N + n]; }
    __device__ void c(int, int m, int n, double v) const
    {
        const int j = n / i1, i = n - j * i1;
        psi[((size_t)(m + 1) * ldj + (size_t)(j + 1)) * ldi + (size_t)(i + 1)] = v;
    }
};

// Even np: the y modes split by parity.  A mode's values at the mirrored
// planes k and k' (k' = np+1-k for cells, np-k for faces) are equal or
// opposite, so each transform is two half-size products: spectral row
// q = 2p + b (b = 0, 1) of the forward takes the sums (b = 0) or the
// differences (b = 1) of the mirrored planes k = 1..np/2 (for faces the
// middle plane np/2 is its own mirror: its sum counts it twice, and its
// weight is halved); the inverse forms E (b = 0) and O (b = 1) on the
// planes k = 1..np/2, and psi(k) = E + O, psi(k') = E - O.
struct YFwdSplit {
    const double *W;        // [2][h][h]
    int M, N, K;            // h = np/2, mt, h
    const double *rhs;
    double *S;
    int i1, ldi, ldj, np, faces, k1;
    __device__ double a(int b, int m, int k) const { return W[((size_t)b * M + m) * K + k]; }
    __device__ double b(int b, int k, int n) const
    {
        const int j = n / i1, i = n - j * i1;
        const size_t off = (size_t)(j + 1) * ldi + (size_t)(i + 1), pl = (size_t)ldj * ldi;
        const int kk = k + 1, km = faces ? np - kk : np + 1 - kk;
        const double x = rhs[(size_t)kk * pl + off], y = rhs[(size_t)km * pl + off];
        return b ? x - y : x + y;
    }
    __device__ void c(int b, int m, int n, double v) const
    {
        const int q = 2 * m + b;
        if (q < k1) S[(size_t)q * N + n] = v;
    }
};
struct YInvSplit {
    const double *Wi;       // [2][h][h]
    int M, N, K;            // h, mt, h
    const double *S;
    double *EO;             // [2][h][mt]
    int k1;
    __device__ double a(int b, int m, int k) const { return Wi[((size_t)b * M + m) * K + k]; }
    __device__ double b(int b, int k, int n) const
    {
        const int q = 2 * k + b;
        return (q < k1) ? S[(size_t)q * N + n] : 0.0;
    }
    __device__ void c(int b, int m, int n, double v) const { EO[((size_t)b * M + m) * N + n] = v; }
};
__global__ void __launch_bounds__(WG) k_t_ycombine(TGeo g, const double *EO, double *psi, int i1, int mt, int faces)
{
    const int n = blockIdx.x * WG + threadIdx.x;
    if (n >= mt) return;
    const int h = g.np / 2, m = blockIdx.y, kk = m + 1;
    const int j = n / i1, i = n - j * i1;
    const unsigned off = (unsigned)((j + 1) * g.ldi + (i + 1)), pl = (unsigned)(g.ldi * g.ldj);
    const double E = EO[(size_t)m * mt + n], O = EO[((size_t)h + m) * mt + n];
    const int km = faces ? g.np - kk : g.np + 1 - kk;
    psi[(unsigned)kk * pl + off] = E + O;
    if (km != kk) psi[(unsigned)km * pl + off] = E - O;
}

// y modes by FFT (np = N a power of two): a mode transform of a line along y
// is the 2N-point DFT of its extension -- the cells' lines (u, w: modes
// cos(pi q (k - 1/2)/N), mirrored at the side walls) even about the walls,
// y_j = x_{j+1}, y_{2N-1-j} = x_{j+1}, whose DFT is 2 e^{i pi q/(2N)} S_q;
// the faces' lines (v: modes sin(pi q k/N), zero at the walls) odd,
// y_k = x_k, y_{2N-k} = -x_k, y_0 = y_N = 0, whose DFT is -2i S_q.  Two real
// lines n = 2p, 2p+1 go through one complex FFT (LP pairs per workgroup, in
// LDS, the stages of vv_fft.h); the cells' spectra separate by conjugate
// symmetry, the faces' fall apart into the real and imaginary parts.  The
// inverse: the cells' DCT-III by the inverse DFT (times 1/(2N)) of the
// conjugate-symmetric spectrum 2 e^{i pi q/(2N)} C_q, which is
// sum_q C_q cos(pi q (k - 1/2)/N) / norm_q (norm_0 = N, else N/2); the
// faces' DST-I being its own inverse (times 2/N).  tw: (cos, -sin)(2 pi t/(2N)), t < N; cq: (cos,
// sin)(pi q/(2N)), q < N.
using vv::bitrev;
using vv::fft_stages;

template <bool FACES>
__global__ void __launch_bounds__(512) k_t_yfft_fwd(TGeo g, const double *__restrict__ rhs, double *__restrict__ S,
                                                    const double2 *tw, const double2 *cq, int i1, int mt, int LP,
                                                    int bits)
{
    extern __shared__ double2 z[];
    const int N = g.np, N2 = 2 * N, nl = 2 * LP;
    const int nb = blockIdx.x * nl;
    const unsigned pl = (unsigned)(g.ldi * g.ldj);
    double *zd = reinterpret_cast<double *>(z);
    double2 *twl = z + LP * N2;   // the stages' twiddles, staged in LDS
    for (int t = threadIdx.x; t < N; t += blockDim.x) twl[t] = tw[t];
    if (FACES)
        for (int e = threadIdx.x; e < nl * 2; e += blockDim.x) {   // y_0 = y_N = 0
            const int l = e % nl, w = e / nl;
            zd[2 * ((l >> 1) * N2 + (int)bitrev((unsigned)(w * N), bits)) + (l & 1)] = 0.0;
        }
    const int K = FACES ? N - 1 : N;
    // element e = (k-1) nl + l: line l of plane k; the line's offset in the
    // plane is fixed per thread (blockDim is a multiple of nl), and 16
    // loads are in flight before their LDS stores
    const int l = (int)threadIdx.x % nl, n = nb + l;
    unsigned off = 0;
    if (n < mt) {
        const int j = n / i1, i = n - j * i1;
        off = (unsigned)((j + 1) * g.ldi + (i + 1));
    }
    double *zl = zd + 2 * (l >> 1) * N2 + (l & 1);
    constexpr int B = 16;
    const int kstep = (int)blockDim.x / nl;
    for (int k0 = 1 + (int)threadIdx.x / nl; k0 <= K; k0 += B * kstep) {
        double x[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int k = k0 + u * kstep;
            x[u] = (n < mt && k <= K) ? rhs[(unsigned)k * pl + off] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int k = k0 + u * kstep;
            if (k > K) break;
            if (FACES) {
                zl[2 * (int)bitrev((unsigned)k, bits)] = x[u];
                zl[2 * (int)bitrev((unsigned)(N2 - k), bits)] = -x[u];
            } else {
                zl[2 * (int)bitrev((unsigned)(k - 1), bits)] = x[u];
                zl[2 * (int)bitrev((unsigned)(N2 - k), bits)] = x[u];
            }
        }
    }
    __syncthreads();
    fft_stages<false>(z, twl, N2, bits, LP);
    if (n >= mt) return;
    const double2 *zp = z + (l >> 1) * N2;
    for (int qi = (int)threadIdx.x / nl; qi < K; qi += kstep) {
        double v;
        if (FACES) {
            const double2 Z = zp[qi + 1];
            v = (l & 1) ? 0.5 * Z.x : -0.5 * Z.y;
        } else {
            const int q = qi;
            const double2 Zm = zp[q], Zc = zp[q == 0 ? 0 : N2 - q];   // Z_q, Z_{2N-q}
            double re, im;   // A = (Z_q + conj Z_{2N-q})/2 or B = (Z_q - conj Z_{2N-q})/(2i)
            if ((l & 1) == 0) { re = 0.5 * (Zm.x + Zc.x); im = 0.5 * (Zm.y - Zc.y); }
            else { re = 0.5 * (Zm.y + Zc.y); im = -0.5 * (Zm.x - Zc.x); }
            const double2 c = cq[q];
            v = 0.5 * (c.x * re + c.y * im);
        }
        S[(size_t)qi * mt + n] = v;
    }
}

template <bool FACES>
__global__ void __launch_bounds__(512) k_t_yfft_inv(TGeo g, const double *__restrict__ S, double *__restrict__ psi,
                                                    const double2 *tw, const double2 *cq, int i1, int mt, int LP,
                                                    int bits)
{
    extern __shared__ double2 z[];
    const int N = g.np, N2 = 2 * N, nl = 2 * LP;
    const int nb = blockIdx.x * nl;
    const unsigned pl = (unsigned)(g.ldi * g.ldj);
    double *zd = reinterpret_cast<double *>(z);
    double2 *twl = z + LP * N2;   // the stages' twiddles, staged in LDS
    for (int t = threadIdx.x; t < N; t += blockDim.x) twl[t] = tw[t];
    if (FACES) {
        for (int e = threadIdx.x; e < nl * 2; e += blockDim.x) {
            const int l = e % nl, w = e / nl;
            zd[2 * ((l >> 1) * N2 + (int)bitrev((unsigned)(w * N), bits)) + (l & 1)] = 0.0;
        }
        const int l = (int)threadIdx.x % nl, n = nb + l;
        double *zl = zd + 2 * (l >> 1) * N2 + (l & 1);
        constexpr int B = 16;
        const int qstep = (int)blockDim.x / nl;
        for (int q0 = 1 + (int)threadIdx.x / nl; q0 <= N - 1; q0 += B * qstep) {
            double x[B];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const int q = q0 + u * qstep;
                x[u] = (n < mt && q <= N - 1) ? S[(size_t)(q - 1) * mt + n] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const int q = q0 + u * qstep;
                if (q > N - 1) break;
                zl[2 * (int)bitrev((unsigned)q, bits)] = x[u];
                zl[2 * (int)bitrev((unsigned)(N2 - q), bits)] = -x[u];
            }
        }
    } else {
        // Z_q = Y^a_q + i Y^b_q, Z_{2N-q} = conj Y^a_q + i conj Y^b_q,
        // Y_q = 2 e^{i pi q/(2N)} C_q / norm_q, Z_N = 0
        const int pr = (int)threadIdx.x % LP, na = nb + 2 * pr, nb1 = na + 1;
        constexpr int B = 8;
        const int qstep = (int)blockDim.x / LP;
        for (int qq0 = (int)threadIdx.x / LP; qq0 < N; qq0 += B * qstep) {
        double va[B], vb[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int q = qq0 + u * qstep;
            va[u] = (na < mt && q < N) ? S[(size_t)q * mt + na] : 0.0;
            vb[u] = (nb1 < mt && q < N) ? S[(size_t)q * mt + nb1] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int q = qq0 + u * qstep;
            if (q >= N) break;
            const double ca = va[u], cb = vb[u];
            const double2 c = cq[q];
            const double ar = 2.0 * c.x * ca, ai = 2.0 * c.y * ca, br = 2.0 * c.x * cb, bi = 2.0 * c.y * cb;
            double2 *zp = z + pr * N2;
            zp[bitrev((unsigned)q, bits)] = make_double2(ar - bi, ai + br);
            if (q > 0) zp[bitrev((unsigned)(N2 - q), bits)] = make_double2(ar + bi, br - ai);
            else zp[bitrev((unsigned)N, bits)] = make_double2(0.0, 0.0);
        }
        }
    }
    __syncthreads();
    if (FACES) fft_stages<false>(z, twl, N2, bits, LP);
    else fft_stages<true>(z, twl, N2, bits, LP);
    const int K = FACES ? N - 1 : N;
    const double sc = FACES ? 1.0 / N : 1.0 / N2;
    const int lo = (int)threadIdx.x % nl, no = nb + lo;
    if (no >= mt) return;
    const int jo = no / i1, io = no - jo * i1;
    const unsigned off = (unsigned)((jo + 1) * g.ldi + (io + 1));
    const double2 *zp = z + (lo >> 1) * N2;
    for (int k = 1 + (int)threadIdx.x / nl; k <= K; k += (int)blockDim.x / nl) {
        double v;
        if (FACES) {
            const double2 Z = zp[k];
            v = (lo & 1) ? Z.x * sc : -Z.y * sc;
        } else {
            const double2 Z = zp[k - 1];
            v = ((lo & 1) ? Z.y : Z.x) * sc;
        }
        psi[(unsigned)k * pl + off] = v;
    }
}

// psip <- psi everywhere; psi <- 0 outside the unknowns
__global__ void __launch_bounds__(WG) k_t_psi_pre(TGeo g, int i1, int j1, int k1, double *psi, double *psip)
{
    const int tp = blockIdx.x * WG + threadIdx.x;
    if (tp >= g.ldi * g.ldj) return;
    const int j = tp / g.ldi, i = tp - j * g.ldi, k = blockIdx.y;
    const unsigned o = (unsigned)(k * g.ldi * g.ldj + tp);
    psip[o] = psi[o];
    if (!(i >= 1 && i <= i1 && j >= 1 && j <= j1 && k >= 1 && k <= k1)) psi[o] = 0.0;
}

// the ghost planes of u and w mirror their first and last cells
__global__ void __launch_bounds__(WG) k_t_psi_post(TGeo g, double *psi)
{
    const int tp = blockIdx.x * WG + threadIdx.x;
    if (tp >= g.ldi * g.ldj) return;
    const unsigned pl = (unsigned)(g.ldi * g.ldj);
    psi[tp] = psi[pl + tp];
    psi[(unsigned)(g.np + 1) * pl + tp] = psi[(unsigned)g.np * pl + tp];
}

// per (y mode q, z eigenvalue l): (D_x + mu_q + lz_l) x = y along i, x
// overwriting y in T (layout [q][i][l]).  Row 1 has an extra coefficient e1
// two points inward, removed with row 2; row i1 has en, removed with the
// eliminated row i1 - 2.  Thomas elimination, one thread per line (threads
// on consecutive l: coalesced), without a factor array: the forward sweep
// keeps only the state (c, y) entering every XCK-th row (in ckp, 2/XCK of
// T), and the back substitution takes the segments from the last: it reads
// the segment's right-hand sides again, repeats their elimination in
// registers from the state before it -- the same operations, so the same
// bits -- and substitutes back.  T is read twice and written once (against
// reading it twice, writing it twice and the factors once each).
constexpr int XCK = 16;

__global__ void __launch_bounds__(64) k_t_xsolve(double *__restrict__ T, double *__restrict__ ckp,
                                                     const double *__restrict__ rt, const double *__restrict__ lz,
                                                     const double *__restrict__ ly, int i1, int j1, int k1, double e1,
                                                     double en)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nlines = k1 * j1;
    if (t >= nlines) return;
    const int q = t / j1, l = t - q * j1;
    const size_t mt = (size_t)i1 * j1;
    double *__restrict__ x = T + (size_t)q * mt + l;
    // (c, y) entering segment sg at ck[(2 sg + w) nlines]
    double *__restrict__ ck = ckp + t;
    const double *sub = rt, *dia = rt + i1, *sup = rt + 2 * i1;
    const double shift = ly[q] + lz[l];
    const int nseg = (i1 + XCK - 1) / XCK;
    // row 0, its third coefficient folded in with row 1
    double d0 = dia[0] + shift, s0 = sup[0];
    const double x0 = x[0], x1 = x[j1];
    double f0 = 0.0;
    if (e1 != 0.0) {
        f0 = e1 / sup[1];
        d0 -= f0 * sub[1];
        s0 -= f0 * (dia[1] + shift);
    }
    // one elimination row: a and r as given (the last row's fold applied by the caller)
    auto row = [&](int i, double a, double r, double &c, double &y) {
        const double den = dia[i] + shift - a * c;
        c = sup[i] / den;
        y = (r - a * y) / den;
    };
    auto row0 = [&](double &c, double &y) {
        y = x0;
        if (e1 != 0.0) y -= f0 * x1;
        c = s0 / d0;
        y /= d0;
    };
    // forward sweep: rows in groups of XCK with their loads issued ahead
    double c, y, c3 = 0.0, y3 = 0.0;   // (c, y) of row i1-3 for the last row's fold
    row0(c, y);
    if (i1 == 3) { c3 = c; y3 = y; }
    for (int sg = 0; sg < nseg; ++sg) {
        const int r0 = sg * XCK;
        if (sg > 0) {
            ck[(size_t)(2 * sg) * nlines] = c;
            ck[(size_t)(2 * sg + 1) * nlines] = y;
        }
        double ru[XCK];
#pragma unroll
        for (int u = 0; u < XCK; ++u) ru[u] = (r0 + u < i1) ? x[(size_t)(r0 + u) * j1] : 0.0;
#pragma unroll
        for (int u = 0; u < XCK; ++u) {
            const int i = r0 + u;
            if (i == 0 || i >= i1) continue;
            double a = sub[i], r = ru[u];
            if (i == i1 - 1 && en != 0.0) {
                a -= en * c3;
                r -= en * y3;
            }
            row(i, a, r, c, y);
            if (i == i1 - 3) { c3 = c; y3 = y; }
        }
    }
    // back substitution, segment by segment from the last
    double xn = 0.0;
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int r0 = sg * XCK;
        double ru[XCK], cv[XCK], yv[XCK];
#pragma unroll
        for (int u = 0; u < XCK; ++u) ru[u] = (r0 + u < i1) ? x[(size_t)(r0 + u) * j1] : 0.0;
        if (sg > 0) {
            c = ck[(size_t)(2 * sg) * nlines];
            y = ck[(size_t)(2 * sg + 1) * nlines];
        }
#pragma unroll
        for (int u = 0; u < XCK; ++u) {
            const int i = r0 + u;
            if (i < i1) {
                if (i == 0) {
                    row0(c, y);
                } else {
                    double a = sub[i], r = ru[u];
                    if (i == i1 - 1 && en != 0.0) {
                        a -= en * c3;
                        r -= en * y3;
                    }
                    row(i, a, r, c, y);
                }
            }
            cv[u] = c;
            yv[u] = y;
        }
#pragma unroll
        for (int u = XCK - 1; u >= 0; --u) {
            const int i = r0 + u;
            if (i >= i1) continue;
            xn = (i == i1 - 1) ? yv[u] : yv[u] - cv[u] * xn;
            x[(size_t)i * j1] = xn;
        }
    }
}

__global__ void __launch_bounds__(WG) k_t_energy(TGeo g, const double *uel, const double *vel, const double *wel,
                                                  double *part)
{
    __shared__ double sh[WG / 64];
    double s = 0.0;
    const int pl = g.ldi * g.ldj;
    const long long n = (long long)pl * (g.np + 2);
    for (long long o = (long long)blockIdx.x * WG + threadIdx.x; o < n; o += (long long)gridDim.x * WG) {
        const int k = (int)(o / pl), r = (int)(o - (long long)k * pl);
        const int j = r / g.ldi, i = r - j * g.ldi;
        const bool kc = k >= 1 && k <= g.np, kf = k >= 1 && k <= g.np - 1;
        if (kc && i >= 1 && i <= g.nr - 1 && j >= 1 && j <= g.nz) s += uel[o] * uel[o];
        if (kf && i >= 1 && i <= g.nr && j >= 1 && j <= g.nz) s += vel[o] * vel[o];
        if (kc && i >= 1 && i <= g.nr && j >= 1 && j <= g.nz - 1) s += wel[o] * wel[o];
    }
    s = vv::block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(WG) k_t_energy_sum(const double *part, int nb, double scale, double *out)
{
    __shared__ double sh[WG / 64];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += WG) s += part[b];
    s = vv::block_sum(s, sh);
    if (threadIdx.x == 0) *out = scale * s;
}

}  // namespace

// the mode transforms of vv_ytrans.h: the kernels above with the tank's
// launch shape (2 LP lines a workgroup of 512 threads, LP = min(4, 2048/np))
bool ytrans_fits(int np) { return np >= 8 && np <= 2048 && (np & (np - 1)) == 0; }

int ytrans_setup(YTrans &T, int np, int ldi, int ldj)
{
    if (!ytrans_fits(np)) {
        vv::set_error("mode transform by FFT: np must be a power of two, 8..2048");
        return VV_EINVAL;
    }
    const double pi = std::acos(-1.0);
    const int N2 = 2 * np;
    T.np = np;
    T.ldi = ldi;
    T.ldj = ldj;
    T.bits = __builtin_ctz((unsigned)N2);
    T.lp = std::max(1, std::min(4, 4096 / N2));
    std::vector<double2> t(2 * (size_t)np);
    for (int q = 0; q < np; ++q) {
        t[q] = make_double2(std::cos(2.0 * pi * q / N2), -std::sin(2.0 * pi * q / N2));
        t[np + q] = make_double2(std::cos(pi * q / N2), std::sin(pi * q / N2));
    }
    VV_HIP(hipMalloc(&T.tw, t.size() * sizeof(double2)));
    VV_HIP(hipMemcpy(T.tw, t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice));
    return VV_OK;
}

void ytrans_free(YTrans &T)
{
    if (T.tw) (void)hipFree(T.tw);
    T.tw = nullptr;
}

static TGeo ytrans_geo(const YTrans &T)
{
    TGeo g{};
    g.np = T.np;
    g.ldi = T.ldi;
    g.ldj = T.ldj;
    return g;
}

void ytrans_fwd(const YTrans &T, bool faces, const double *rhs, double *S, int i1, int mt, hipStream_t s)
{
    const int LP = T.lp, nbk = (mt + 2 * LP - 1) / (2 * LP);
    const size_t flds = ((size_t)LP * 2 * T.np + T.np) * sizeof(double2);
    const TGeo g = ytrans_geo(T);
    const double2 *tw = T.tw, *cq = T.tw + T.np;
    if (faces)
        hipLaunchKernelGGL(k_t_yfft_fwd<true>, dim3(nbk), dim3(512), flds, s, g, rhs, S, tw, cq, i1, mt, LP, T.bits);
    else
        hipLaunchKernelGGL(k_t_yfft_fwd<false>, dim3(nbk), dim3(512), flds, s, g, rhs, S, tw, cq, i1, mt, LP, T.bits);
}

void ytrans_inv(const YTrans &T, bool faces, const double *S, double *psi, int i1, int mt, hipStream_t s)
{
    const int LP = T.lp, nbk = (mt + 2 * LP - 1) / (2 * LP);
    const size_t flds = ((size_t)LP * 2 * T.np + T.np) * sizeof(double2);
    const TGeo g = ytrans_geo(T);
    const double2 *tw = T.tw, *cq = T.tw + T.np;
    if (faces)
        hipLaunchKernelGGL(k_t_yfft_inv<true>, dim3(nbk), dim3(512), flds, s, g, S, psi, tw, cq, i1, mt, LP, T.bits);
    else
        hipLaunchKernelGGL(k_t_yfft_inv<false>, dim3(nbk), dim3(512), flds, s, g, S, psi, tw, cq, i1, mt, LP, T.bits);
}

}  // namespace vvt

using namespace vvt;

// one Poisson operator of the tank
struct TankOp {
    int i1 = 0, j1 = 0, k1 = 0, m = 0;
    double e1 = 0.0, en = 0.0;
    double *W = nullptr, *Wi = nullptr;   // k1 x k1: y modes (forward, inverse); for even np the
                                          // parity-split halves [2][np/2][np/2] instead
    double *ly = nullptr;                 // k1: y eigenvalues / dp^2
    double *V = nullptr, *Vinv = nullptr; // j1 x j1: z eigenvectors (row-major) and inverse; with
                                          // zsplit the parity halves [2][he][he] instead (G, F)
    double *lz = nullptr;                 // j1: z eigenvalues / dz^2 (zsplit: even modes first)
    bool zsplit = false;
    double *rt = nullptr;                 // 3 x i1: x rows (sub, diagonal, super)
};

struct vvt_run {
    std::recursive_mutex mu;
    vvt_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    int nr = 0, nz = 0, np = 0, ldi = 0, ldj = 0;
    double dr = 0, dz = 0, dp = 0;
    size_t nf = 0;
    double *f[VV3_NFIELD] = {nullptr};
    // per field: zero outside its solve's unknowns, ghost planes aside (the
    // fields start at zero; a put clears the flag); see launch_poisson
    bool clean[VV3_NFIELD] = {};
    double *dv[12] = {nullptr};
    double *tabs = nullptr;       // compact elimination tables
    size_t tab_off[12] = {0};
    TankOp op[3];
    double *S = nullptr, *T = nullptr, *cp = nullptr;   // Poisson work (kmax x mmax each)
    // v's right-hand side and solve run beside w's and u's (one_step): a
    // second stream with work arrays of its own, and the fork / join events
    hipStream_t stream2 = nullptr;
    double *S2 = nullptr, *T2 = nullptr, *cp2 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // the x derivatives run on the second stream beside the y and z ones,
    // with checkpoints of their own (4 jobs)
    double *ckpx = nullptr;
    hipEvent_t forkd = nullptr, joind = nullptr;
    // the stream and work arrays launch_rhs / launch_poisson use
    hipStream_t cur = nullptr;
    double *cS = nullptr, *cT = nullptr, *ccp = nullptr;
    double *ckp = nullptr;        // checkpoints of the y/z derivative lines: 8 jobs x ckp_seg x ckp_lines
    int ckp_seg = 0, ckp_lines = 0;
    double2 *ytw = nullptr;       // y FFT: twiddles of 2np points (np), then (cos, sin)(pi q/(2np)) (np)
    bool yfft = false;            // y modes by FFT (np a power of two, 8..2048)
    int ylp = 1, ybits = 0;       // line pairs per workgroup, log2(2np)
    double *heat = nullptr;       // 2 nr column sums, trans1, trans2, then the flux terms
    double *kepart = nullptr;     // KE_MAXB partials, then the total
    // profiling
    bool prof = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    long long prof_n[VVT_K_COUNT] = {0};
    double prof_ms[VVT_K_COUNT] = {0}, prof_bytes[VVT_K_COUNT] = {0};
};

namespace {

TGeo geo(const vvt_run *r)
{
    TGeo g;
    g.nr = r->nr; g.nz = r->nz; g.np = r->np; g.ldi = r->ldi; g.ldj = r->ldj;
    g.dt = r->cfg.dt; g.dr = r->dr; g.dz = r->dz; g.dp = r->dp;
    g.Re = r->cfg.Re; g.Ptl = r->cfg.Ptl; g.Gr = r->cfg.Gr;
    return g;
}

template <class T>
hipError_t upload(T **dst, const std::vector<T> &src)
{
    hipError_t e = hipMalloc(dst, src.size() * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
}

// the tables of a compact line, as the oracle (and the reference's tdma)
// forms them: rows 1-4-1, type 1's end rows (1, 2) and (2, 1)
void compact_tables(int type, int m, std::vector<double> &out)
{
    std::vector<double> a(m + 1, 1.0), b(m + 1, 4.0), c(m + 1, 1.0), fact(m + 1, 0.0);
    if (type == 1) {
        b[1] = 1.0;
        c[1] = 2.0;
        a[m] = 2.0;
        b[m] = 1.0;
    }
    for (int j = 2; j <= m; ++j) {
        fact[j] = a[j] / b[j - 1];
        b[j] = b[j] - fact[j] * c[j - 1];
    }
    std::vector<double> rb(m + 1, 0.0);
    for (int j = 1; j <= m; ++j) rb[j] = 1.0 / b[j];
    out.insert(out.end(), fact.begin(), fact.end());
    out.insert(out.end(), b.begin(), b.end());
    out.insert(out.end(), c.begin(), c.end());
    out.insert(out.end(), rb.begin(), rb.end());
}

// (field, axis) -> (type, m, n1, n2): the staggering of A, B, C, th
struct DSpec {
    int type, m, n1, n2;
};
DSpec dspec(const vvt_run *r, int d)
{
    const int nr = r->nr, nz = r->nz, np = r->np;
    switch (d) {
    case 0: return {1, nr, nz - 1, np - 1};        // A x: cells, lines (j, k)
    case 1: return {0, np - 1, nr, nz - 1};        // A y: faces, lines (i, j)
    case 2: return {0, nz - 1, nr, np - 1};        // A z: faces, lines (i, k)
    case 3: return {0, nr - 1, nz - 1, np};        // B x: faces
    case 4: return {0, np, nr - 1, nz - 1};        // B y: cells with ghost planes
    case 5: return {0, nz - 1, nr - 1, np};        // B z: faces
    case 6: return {0, nr - 1, nz, np - 1};        // C x: faces
    case 7: return {0, np - 1, nr - 1, nz};        // C y: faces
    case 8: return {1, nz, nr - 1, np - 1};        // C z: cells, walls half a spacing away
    case 9: return {1, nr, nz, np};                // th x
    case 10: return {0, np, nr, nz};               // th y
    default: return {1, nz, nr, np};               // th z
    }
}

int setup_poisson(vvt_run *r)
{
    using vv::set_error;
    const double pi = std::acos(-1.0);
    const int np = r->np;
    const double hhr = r->dr * r->dr, hhz = r->dz * r->dz, hhp = r->dp * r->dp;
    size_t mmax = 0, kmax = 0;
    r->yfft = np >= 8 && np <= 2048 && (np & (np - 1)) == 0;
    if (r->yfft) {
        const int N2 = 2 * np;
        r->ybits = __builtin_ctz((unsigned)N2);
        // 8 lines a workgroup (64 B of each plane) in 64 KB of LDS, plus the
        // twiddles: two workgroups per CU (16 lines in 128 KB measured slower)
        r->ylp = std::max(1, std::min(4, 4096 / N2));
        std::vector<double2> t(2 * (size_t)np);
        for (int q = 0; q < np; ++q) {
            t[q] = make_double2(std::cos(2.0 * pi * q / N2), -std::sin(2.0 * pi * q / N2));
            t[np + q] = make_double2(std::cos(pi * q / N2), std::sin(pi * q / N2));
        }
        VV_HIP(upload(&r->ytw, t));
    }
    for (int w = 1; w <= 3; ++w) {
        TankOp &O = r->op[w - 1];
        O.i1 = (w == 3) ? r->nr - 1 : r->nr;
        O.j1 = (w == 1) ? r->nz - 1 : r->nz;
        O.k1 = (w == 2) ? np - 1 : np;
        O.m = O.i1 * O.j1;
        mmax = std::max(mmax, (size_t)O.m);
        kmax = std::max(kmax, (size_t)O.k1);
        // y modes: u, w mirrored at the side walls (cos(pi q (k - 1/2)/np),
        // q = 0..np-1), v zero there (sin(pi q k/np), q = 1..np-1); each mode
        // is an eigenvector of the y rows with eigenvalue 2 cos(pi q/np) - 2
        const int k1 = O.k1;
        std::vector<double> W, Wi, ly(k1);
        if (!r->yfft) {
            W.resize((size_t)k1 * k1);
            Wi.resize((size_t)k1 * k1);
        }
        auto mode = [&](int q, int k) {   // mode q (0-based spectral row) at plane k
            const int qq = (w == 2) ? q + 1 : q;
            return (w == 2) ? std::sin(pi * qq * k / np) : std::cos(pi * qq * (k - 0.5) / np);
        };
        auto norm = [&](int q) { return (w != 2 && q == 0) ? (double)np : 0.5 * np; };
        for (int qi = 0; qi < k1; ++qi) {
            const int q = (w == 2) ? qi + 1 : qi;
            ly[qi] = (2.0 * std::cos(pi * q / np) - 2.0) / hhp;
            if (r->yfft) continue;
            for (int k = 1; k <= k1; ++k) {
                W[(size_t)qi * k1 + (k - 1)] = mode(qi, k);
                Wi[(size_t)(k - 1) * k1 + qi] = mode(qi, k) / norm(qi);
            }
        }
        if (!r->yfft && np % 2 == 0) {
            // the parity-split halves: forward row p of half b is spectral row
            // q = 2p + b at the planes k = 1..np/2 (the faces' middle plane
            // np/2 at half weight); inverse row k of half b holds the inverse
            // weights of the rows q = 2p + b at plane k
            const int h = np / 2;
            std::vector<double> Ws((size_t)2 * h * h, 0.0), Wis((size_t)2 * h * h, 0.0);
            for (int b = 0; b < 2; ++b)
                for (int pp = 0; pp < h; ++pp) {
                    const int q = 2 * pp + b;
                    if (q >= k1) continue;
                    for (int k = 1; k <= h; ++k) {
                        const double e = mode(q, k);
                        Ws[((size_t)b * h + pp) * h + (k - 1)] = (w == 2 && k == h) ? 0.5 * e : e;
                        Wis[((size_t)b * h + (k - 1)) * h + pp] = e / norm(q);
                    }
                }
            W.swap(Ws);
            Wi.swap(Wis);
        }
        if (!r->yfft) {
            VV_HIP(upload(&O.W, W));
            VV_HIP(upload(&O.Wi, Wi));
        }
        VV_HIP(upload(&O.ly, ly));
        // z rows: rot3dmg.f's (Dirichlet for w, the wall closure for v, u)
        std::vector<double> lam, V, Vinv;
        if (!vv3::eigen_banded(vv3::z_matrix(w, O.j1), O.j1, lam, V, Vinv)) {
            set_error("vvt_create: z eigen-decomposition of a Poisson operator failed");
            return VV_EINVAL;
        }
        for (double &v : lam) v /= hhz;
        std::vector<double> Fz, Gz;
        O.zsplit = vv3::zsplit_wanted(O.j1) && 2 * O.k1 <= 65535 && vv3::z_split(O.j1, lam, V, Vinv, Fz, Gz);
        VV_HIP(upload(&O.V, O.zsplit ? Gz : V));
        VV_HIP(upload(&O.Vinv, O.zsplit ? Fz : Vinv));
        VV_HIP(upload(&O.lz, lam));
        // x rows: the wall closure (-2, 5/3, -0.2) for the cells of w and v,
        // Dirichlet for the faces of u
        const int i1 = O.i1;
        std::vector<double> rt((size_t)3 * i1, 0.0);
        for (int i = 0; i < i1; ++i) {
            rt[i1 + i] = -2.0 / hhr;
            if (i > 0) rt[i] = 1.0 / hhr;
            if (i < i1 - 1) rt[2 * i1 + i] = 1.0 / hhr;
        }
        if (w != 3) {
            rt[2 * i1 + 0] = 5.0 / 3.0 / hhr;
            rt[i1 - 1] = 5.0 / 3.0 / hhr;
            O.e1 = -0.2 / hhr;
            O.en = -0.2 / hhr;
        }
        VV_HIP(upload(&O.rt, rt));
    }
    VV_HIP(hipMalloc(&r->S, kmax * mmax * sizeof(double)));
    VV_HIP(hipMalloc(&r->T, kmax * mmax * sizeof(double)));
    VV_HIP(hipMalloc(&r->cp, kmax * mmax * sizeof(double)));
    r->cur = r->stream; r->cS = r->S; r->cT = r->T; r->ccp = r->cp;
    // the second solve stream: without its memory the solves run in turn
    const size_t w2 = kmax * mmax * sizeof(double);
    if (hipMalloc(&r->S2, w2) == hipSuccess && hipMalloc(&r->T2, w2) == hipSuccess &&
        hipMalloc(&r->cp2, w2) == hipSuccess &&
        hipMalloc(&r->ckpx, 4 * (size_t)r->ckp_seg * r->ckp_lines * sizeof(double)) == hipSuccess &&
        hipStreamCreateWithFlags(&r->stream2, hipStreamNonBlocking) == hipSuccess &&
        hipEventCreateWithFlags(&r->fork, hipEventDisableTiming) == hipSuccess &&
        hipEventCreateWithFlags(&r->join, hipEventDisableTiming) == hipSuccess &&
        hipEventCreateWithFlags(&r->forkd, hipEventDisableTiming) == hipSuccess &&
        hipEventCreateWithFlags(&r->joind, hipEventDisableTiming) == hipSuccess)
        return VV_OK;
    (void)hipGetLastError();
    for (double **p : {&r->S2, &r->T2, &r->cp2, &r->ckpx})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (r->stream2) { (void)hipStreamDestroy(r->stream2); r->stream2 = nullptr; }
    for (hipEvent_t *e : {&r->fork, &r->join, &r->forkd, &r->joind})
        if (*e) { (void)hipEventDestroy(*e); *e = nullptr; }
    return VV_OK;
}

// profiling brackets around one group of launches
struct Prof {
    vvt_run *r;
    int id;
    double bytes;
    Prof(vvt_run *rr, int k, double b) : r(rr), id(k), bytes(b)
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
        r->prof_bytes[id] += bytes;
    }
};

double npts(const vvt_run *r) { return (double)r->nr * r->nz * r->np; }

VF vfields(vvt_run *r)
{
    VF F;
    F.A = r->f[VV3_A]; F.B = r->f[VV3_B]; F.C = r->f[VV3_C];
    F.Ap = r->f[VV3_AP]; F.Bp = r->f[VV3_BP]; F.Cp = r->f[VV3_CP];
    F.th = r->f[VV3_TH]; F.thp = r->f[VV3_THP];
    F.uel = r->f[VV3_UEL]; F.vel = r->f[VV3_VEL]; F.wel = r->f[VV3_WEL];
    for (int d = 0; d < 12; ++d) F.dv[d] = r->dv[d];
    F.An = r->f[VV3_AN]; F.Bn = r->f[VV3_BN]; F.Cn = r->f[VV3_CN]; F.thn = r->f[VV3_THN];
    return F;
}

int launch_derivatives(vvt_run *r)
{
    DJobs J;
    int nyz = 0, nx = 0, maxm = 0, maxmx = 0, nlx = 0;
    double byz = 0.0, bx = 0.0;
    const double *src[4] = {r->f[VV3_A], r->f[VV3_B], r->f[VV3_C], r->f[VV3_TH]};
    const double dri[3] = {1.0 / r->dr, 1.0 / r->dp, 1.0 / r->dz};
    for (int d = 0; d < 12; ++d) {
        const DSpec s = dspec(r, d);
        const int nl = (d % 3 == 0) ? 64 : 1;
        J.j[d] = DJob{src[d / 3], r->dv[d], r->tabs + r->tab_off[d], d % 3, s.type, s.m, s.n1, s.n2, nl, dri[d % 3]};
        const double b = 8.0 * (2.0 * s.m + 2.0) * s.n1 * s.n2;
        if (d % 3 == 0) bx += b;
        else byz += b;
        if (d % 3 == 0) {
            nx = std::max(nx, ((s.n1 + nl - 1) / nl) * s.n2);
            maxmx = std::max(maxmx, s.m);
            nlx = std::max(nlx, s.n1 * s.n2);
        } else {
            nyz = std::max(nyz, s.n1 * s.n2);
            maxm = std::max(maxm, s.m);
        }
    }
    // the y and z jobs in one launch: jobs 1, 2, 4, 5, ... are reached as
    // j0 + blockIdx.y with blockIdx.y over a compacted list
    DJobs Jyz;
    int c = 0;
    for (int d = 0; d < 12; ++d)
        if (d % 3 != 0) Jyz.j[c++] = J.j[d];
    if (r->stream2 && !r->prof) {   // the x jobs start from the state the y and z jobs read
        VV_HIP(hipEventRecord(r->forkd, r->stream));
        VV_HIP(hipStreamWaitEvent(r->stream2, r->forkd, 0));
    }
    {
        Prof p(r, VVT_K_DERIV_YZ, byz);
        if (c > 8 || maxm > r->ckp_seg * CK || nyz > r->ckp_lines) {
            vv::set_error("vvt: derivative checkpoints exceed their buffer");
            return VV_EINVAL;
        }
        if (maxm <= DYZ_LDS_MAXM)
            hipLaunchKernelGGL(k_t_deriv_yz<true>, dim3(nblk(nyz, WG), c), dim3(WG),
                               4 * (size_t)(maxm + 1) * sizeof(double), r->stream, geo(r), Jyz, 0, r->ckp,
                               r->ckp_seg, r->ckp_lines);
        else
            hipLaunchKernelGGL(k_t_deriv_yz<false>, dim3(nblk(nyz, WG), c), dim3(WG), 0, r->stream, geo(r), Jyz, 0,
                               r->ckp, r->ckp_seg, r->ckp_lines);
    }
    {
        Prof p(r, VVT_K_DERIV_X, bx);
        if ((maxmx + CX - 1) / CX > r->ckp_seg || nlx > r->ckp_lines) {
            vv::set_error("vvt: derivative checkpoints exceed their buffer");
            return VV_EINVAL;
        }
        if (r->stream2 && !r->prof) {
            // on the second stream (from the state the y and z jobs also
            // read), joined before the explicit updates read the derivatives
            hipLaunchKernelGGL(k_t_deriv_x, dim3(nx, 4), dim3(64), 0, r->stream2, geo(r), J, r->ckpx, r->ckp_seg,
                               r->ckp_lines);
            VV_HIP(hipEventRecord(r->joind, r->stream2));
            VV_HIP(hipStreamWaitEvent(r->stream, r->joind, 0));
        } else {
            hipLaunchKernelGGL(k_t_deriv_x, dim3(nx, 4), dim3(64), 0, r->stream, geo(r), J, r->ckp, r->ckp_seg,
                               r->ckp_lines);
        }
    }
    VV_HIP(hipGetLastError());
    return VV_OK;
}

// mask 1: A, B, C; 2: temperature; 3: all four
int launch_explicit(vvt_run *r, int mask)
{
    // algorithmic words per point: each field read or written once -- A, B,
    // C, their previous levels, th, u, v, w and nine derivatives, three
    // outputs (mask 1); th, thp, u, v, w, three derivatives, thn (mask 2)
    const double words = mask == 3 ? 27.0 : (mask == 1 ? 22.0 : 9.0);
    Prof p(r, VVT_K_EXPLICIT, 8.0 * words * npts(r));
    const dim3 grid(nblk((long long)r->nr * r->nz, WG), (r->np + KC - 1) / KC);
    const TGeo g = geo(r);
    const VF F = vfields(r);
    const XDiv q = xdiv(g);
    if (mask == 1) hipLaunchKernelGGL(k_t_explicit<1>, grid, dim3(WG), 0, r->stream, g, F, q);
    else if (mask == 2) hipLaunchKernelGGL(k_t_explicit<2>, grid, dim3(WG), 0, r->stream, g, F, q);
    else hipLaunchKernelGGL(k_t_explicit<3>, grid, dim3(WG), 0, r->stream, g, F, q);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

double *heat_out(vvt_run *r) { return r->heat + 2 * r->nr; }

int launch_heat(vvt_run *r)
{
    Prof p(r, VVT_K_HEAT, 8.0 * 10.0 * r->nr * (r->np / 2 + 1));
    double *E = r->heat + 2 * (size_t)r->nr + 2;
    hipLaunchKernelGGL(k_t_heat_terms, dim3(nblk(r->nr, WG), r->np - r->np / 2 + 1), dim3(WG), 0, r->cur, geo(r),
                       (const double *)r->f[VV3_TH], E);
    hipLaunchKernelGGL(k_t_heat_cols, dim3(nblk(r->nr, WG)), dim3(WG), 0, r->cur, geo(r), (const double *)E,
                       r->heat);
    hipLaunchKernelGGL(k_t_heat_sum, dim3(1), dim3(64), 0, r->cur, r->nr, r->dr, (const double *)r->heat,
                       heat_out(r));
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_rhs(vvt_run *r, int which)
{
    Prof p(r, VVT_K_RHS, 8.0 * 3.0 * npts(r));
    RF F{r->f[VV3_AN], r->f[VV3_BN], r->f[VV3_CN], r->f[VV3_WRH], r->f[VV3_VRH], r->f[VV3_URH]};
    const int ni = (which == 3) ? r->nr - 1 : r->nr, nj = (which == 1) ? r->nz - 1 : r->nz;
    const int nk = (which == 2) ? r->np - 1 : r->np;
    hipLaunchKernelGGL(k_t_rhs, dim3(nblk((long long)ni * nj, WG), nk), dim3(WG), 0, r->cur, geo(r), F, which);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_poisson(vvt_run *r, int which)
{
    const TankOp &O = r->op[which - 1];
    const int psi = (which == 1) ? VV3_WEL : (which == 2) ? VV3_VEL : VV3_UEL;
    const int psip = (which == 1) ? VV3_WELP : (which == 2) ? VV3_VELP : VV3_UELP;
    const int rhs = (which == 1) ? VV3_WRH : (which == 2) ? VV3_VRH : VV3_URH;
    const TGeo g = geo(r);
    hipStream_t s = r->cur;
    double *P = r->f[psi];
    const int mt = O.m;
    const double ftr_y = 2.0 * O.k1 * O.k1 * (double)mt, ftr_z = 2.0 * O.j1 * O.j1 * (double)O.i1 * O.k1;
    // psip <- psi and psi <- 0 outside the unknowns: when psip's buffer is
    // already zero there (it held psi before the previous solve), the two
    // buffers swap -- psip becomes the old psi, psi a buffer the solve then
    // fills (unknowns, then the ghost planes) -- with no copy at all
    if (r->clean[psip]) {
        std::swap(r->f[psi], r->f[psip]);
        std::swap(r->clean[psi], r->clean[psip]);
    } else {
        Prof p(r, VVT_K_PSI_PRE, 8.0 * 3.0 * (double)r->nf);
        hipLaunchKernelGGL(k_t_psi_pre, dim3(nblk((long long)r->ldi * r->ldj, WG), r->np + 2), dim3(WG), 0, s, g,
                           O.i1, O.j1, O.k1, P, r->f[psip]);
        r->clean[psip] = r->clean[psi];
        r->clean[psi] = true;
    }
    P = r->f[psi];
    const bool split = (r->np % 2 == 0);
    const int h = r->np / 2;
    const int LP = r->ylp, nbk = (mt + 2 * LP - 1) / (2 * LP);
    const size_t flds = ((size_t)LP * 2 * r->np + r->np) * sizeof(double2);
    const int yt = 512;    // threads: a multiple of the 2 LP lines (measured: 256 and 1024 slower)
    const double2 *tw = r->ytw, *cq = r->ytw + r->np;
    if (r->yfft) {   // reads the right-hand side's unknowns, writes the spectra
        Prof p(r, VVT_K_YFFT_FWD, 8.0 * 2.0 * O.k1 * (double)mt);
        if (which == 2)
            hipLaunchKernelGGL(k_t_yfft_fwd<true>, dim3(nbk), dim3(yt), flds, s, g, (const double *)r->f[rhs], r->cS,
                               tw, cq, O.i1, mt, LP, r->ybits);
        else
            hipLaunchKernelGGL(k_t_yfft_fwd<false>, dim3(nbk), dim3(yt), flds, s, g, (const double *)r->f[rhs], r->cS,
                               tw, cq, O.i1, mt, LP, r->ybits);
    } else {
        Prof p(r, VVT_K_YFWD, split ? 2.0 * (2.0 * h * h * (double)mt) : ftr_y);
        if (split)
            vv::gemm_f64<true, false>(YFwdSplit{O.W, h, mt, h, r->f[rhs], r->cS, O.i1, r->ldi, r->ldj, r->np,
                                                which == 2, O.k1}, 2, s);
        else
            vv::gemm_f64<true, false>(YFwd{O.W, O.k1, mt, O.k1, r->f[rhs], r->cS, O.i1, r->ldi, r->ldj}, 1, s);
    }
    const int zh = O.j1 / 2, zhe = O.j1 - zh;
    const double ftr_zs = 2.0 * ((double)zhe * zhe + (double)zh * zh) * O.i1 * O.k1;
    if (O.zsplit) {
        Prof p(r, VVT_K_ZFWD, ftr_zs);
        vv::zsplit_fwd(O.Vinv, r->cS, r->cT, O.i1, O.j1, O.k1, s);
    } else {
        Prof p(r, VVT_K_ZFWD, ftr_z);
        vv::gemm_f64<false, true>(ZFwd{O.Vinv, O.i1, O.j1, O.j1, r->cS, r->cT, mt}, O.k1, s);
    }
    {   // reads T and writes it (algorithmic: once each)
        Prof p(r, VVT_K_XSOLVE, 8.0 * 2.0 * O.k1 * (double)mt);
        // checkpoints in cp: 2 ceil(i1/XCK) k1 j1 <= i1 k1 j1 words (i1 >= 5)
        hipLaunchKernelGGL(k_t_xsolve, dim3(nblk((long long)O.k1 * O.j1, 64)), dim3(64), 0, s, r->cT, r->ccp,
                           (const double *)O.rt, (const double *)O.lz, (const double *)O.ly, O.i1, O.j1, O.k1, O.e1,
                           O.en);
    }
    if (O.zsplit) {
        Prof p(r, VVT_K_ZINV, ftr_zs);
        vv::zsplit_inv(O.V, r->cT, r->cS, O.i1, O.j1, O.k1, s);
    } else {
        Prof p(r, VVT_K_ZINV, ftr_z);
        vv::gemm_f64<true, true>(ZInv{O.V, O.j1, O.i1, O.j1, r->cT, r->cS, mt}, O.k1, s);
    }
    if (r->yfft) {
        Prof p(r, VVT_K_YFFT_INV, 8.0 * 2.0 * O.k1 * (double)mt);
        if (which == 2)
            hipLaunchKernelGGL(k_t_yfft_inv<true>, dim3(nbk), dim3(yt), flds, s, g, (const double *)r->cS, P, tw, cq,
                               O.i1, mt, LP, r->ybits);
        else
            hipLaunchKernelGGL(k_t_yfft_inv<false>, dim3(nbk), dim3(yt), flds, s, g, (const double *)r->cS, P, tw, cq,
                               O.i1, mt, LP, r->ybits);
        if (which != 2)
            hipLaunchKernelGGL(k_t_psi_post, dim3(nblk((long long)r->ldi * r->ldj, WG)), dim3(WG), 0, s, g, P);
    } else {
        Prof p(r, VVT_K_YINV, split ? 2.0 * (2.0 * h * h * (double)mt) : ftr_y);
        if (split) {
            // E and O go to T (free after the z inverse), then the planes
            vv::gemm_f64<true, false>(YInvSplit{O.Wi, h, mt, h, r->cS, r->cT, O.k1}, 2, s);
            hipLaunchKernelGGL(k_t_ycombine, dim3(nblk(mt, WG), h), dim3(WG), 0, s, g, (const double *)r->cT, P, O.i1,
                               mt, which == 2);
        } else {
            vv::gemm_f64<true, false>(YInv{O.Wi, O.k1, mt, O.k1, r->cS, P, O.i1, r->ldi, r->ldj}, 1, s);
        }
        if (which != 2)
            hipLaunchKernelGGL(k_t_psi_post, dim3(nblk((long long)r->ldi * r->ldj, WG)), dim3(WG), 0, s, g, P);
    }
    VV_HIP(hipGetLastError());
    return VV_OK;
}

int launch_update(vvt_run *r)
{
    Prof p(r, VVT_K_UPDATE, 8.0 * 8.0 * (double)r->nf);
    const TGeo g = geo(r);
    UF F{r->f[VV3_UEL], r->f[VV3_VEL], r->f[VV3_WEL], r->f[VV3_AN], r->f[VV3_BN], r->f[VV3_CN], r->f[VV3_THN]};
    hipStream_t s = r->stream;
    hipLaunchKernelGGL(k_t_upd_walls, dim3(nblk(r->nr + r->nz, WG), r->np), dim3(WG), 0, s, g, F);
    const int plane = r->ldi * r->ldj;
    hipLaunchKernelGGL(k_t_upd_side, dim3(nblk(plane, WG)), dim3(WG), 0, s, g, F);
    hipLaunchKernelGGL(k_t_upd_temp, dim3(nblk((long long)r->nr * r->np, WG)), dim3(WG), 0, s, g, r->f[VV3_THN], 0);
    hipLaunchKernelGGL(k_t_upd_temp, dim3(nblk((long long)(r->nz + 2) * r->np, WG)), dim3(WG), 0, s, g,
                       r->f[VV3_THN], 1);
    hipLaunchKernelGGL(k_t_upd_temp, dim3(nblk(plane, WG)), dim3(WG), 0, s, g, r->f[VV3_THN], 2);
    // Ap <- A <- An (and B, C, th): the old Ap buffer receives An, then the
    // three buffers rotate
    const int trip[4][3] = {{VV3_AP, VV3_A, VV3_AN}, {VV3_BP, VV3_B, VV3_BN}, {VV3_CP, VV3_C, VV3_CN},
                            {VV3_THP, VV3_TH, VV3_THN}};
    CopyF C;
    for (int q = 0; q < 4; ++q) {
        C.src[q] = r->f[trip[q][2]];
        C.dst[q] = r->f[trip[q][0]];
    }
    hipLaunchKernelGGL(k_t_copy, dim3(nblk((long long)r->nf, WG), 4), dim3(WG), 0, s, C, r->nf);
    VV_HIP(hipGetLastError());
    for (auto &t : trip) {
        double *pp = r->f[t[0]], *c = r->f[t[1]], *nn = r->f[t[2]];
        r->f[t[0]] = c;
        r->f[t[1]] = nn;
        r->f[t[2]] = pp;
    }
    return VV_OK;
}

int copy_field(vvt_run *r, double *dev, double *host, int ldi, int ldj, bool to_dev)
{
    const int di = r->ldi, dj = r->ldj, nk = r->np + 2;
    if (ldi == 0) ldi = di;
    if (ldj == 0) ldj = dj;
    if (ldi < di || ldj < dj) {
        vv::set_error("vvt_put/vvt_get: ldi must be >= nr+2 and ldj >= nz+2");
        return VV_EINVAL;
    }
    VV_HIP(hipSetDevice(r->device));
    const hipMemcpyKind kind = to_dev ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    if (ldi == di && ldj == dj) {
        if (to_dev) VV_HIP(hipMemcpyAsync(dev, host, r->nf * sizeof(double), kind, r->stream));
        else VV_HIP(hipMemcpyAsync(host, dev, r->nf * sizeof(double), kind, r->stream));
    } else {
        for (int k = 0; k < nk; ++k) {
            double *h = host + (size_t)k * ldi * ldj;
            double *d = dev + (size_t)k * di * dj;
            if (to_dev)
                VV_HIP(hipMemcpy2DAsync(d, di * sizeof(double), h, ldi * sizeof(double), di * sizeof(double), dj,
                                        kind, r->stream));
            else
                VV_HIP(hipMemcpy2DAsync(h, ldi * sizeof(double), d, di * sizeof(double), di * sizeof(double), dj,
                                        kind, r->stream));
        }
    }
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int one_step(vvt_run *r)
{
    int st;
    if ((st = launch_derivatives(r))) return st;
    if ((st = launch_explicit(r, 3))) return st;
    if (r->stream2 && !r->prof) {
        // the heat fluxes (they read th) and v's right-hand side (it reads
        // An and Cn only, rot3dmg.f:1066-1093) and solve run on the second
        // stream, v's solve with its own work arrays, while w's and u's (u's
        // right-hand side reads the new w) run on the first; the update
        // waits for both
        VV_HIP(hipEventRecord(r->fork, r->stream));
        VV_HIP(hipStreamWaitEvent(r->stream2, r->fork, 0));
        r->cur = r->stream2; r->cS = r->S2; r->cT = r->T2; r->ccp = r->cp2;
        st = launch_heat(r);
        if (!st) st = launch_rhs(r, 2);
        if (!st) st = launch_poisson(r, 2);
        r->cur = r->stream; r->cS = r->S; r->cT = r->T; r->ccp = r->cp;
        if (st) return st;
        VV_HIP(hipEventRecord(r->join, r->stream2));
        for (int w = 1; w <= 3; w += 2) {
            if ((st = launch_rhs(r, w))) return st;
            if ((st = launch_poisson(r, w))) return st;
        }
        VV_HIP(hipStreamWaitEvent(r->stream, r->join, 0));
    } else {
        if ((st = launch_heat(r))) return st;
        for (int w = 1; w <= 3; ++w) {
            if ((st = launch_rhs(r, w))) return st;
            if ((st = launch_poisson(r, w))) return st;
        }
    }
    return launch_update(r);
}

}  // namespace

extern "C" {

int vvt_create(const vvt_config *cfg, int device, vvt_run **out)
{
    using vv::set_error;
    if (!cfg || !out) { set_error("vvt_create: null argument"); return VV_EINVAL; }
    *out = nullptr;
    if (cfg->nr < 6 || cfg->nz < 6 || cfg->np < 6 || cfg->nr > 4096 || cfg->nz > 4096 || cfg->np > 8192) {
        set_error("vvt_create: need 6 <= nr, nz <= 4096 and 6 <= np <= 8192");
        return VV_EINVAL;
    }
    const double nf = (double)(cfg->nr + 2) * (cfg->nz + 2) * (cfg->np + 2);
    if (nf >= 4.0e9) { set_error("vvt_create: fields beyond 2^32 points"); return VV_EINVAL; }
    if (!(cfg->dt > 0) || !(cfg->lx > 0) || !(cfg->lz > 0) || !(cfg->ly > 0) || !(cfg->Ptl > 0) ||
        !std::isfinite(cfg->Re) || !std::isfinite(cfg->Gr)) {
        set_error("vvt_create: bad physical parameters");
        return VV_EINVAL;
    }
    VV_HIP(hipSetDevice(device));
    vvt_run *r = new vvt_run();
    r->cfg = *cfg;
    r->device = device;
    r->nr = cfg->nr; r->nz = cfg->nz; r->np = cfg->np;
    r->ldi = cfg->nr + 2; r->ldj = cfg->nz + 2;
    r->dr = cfg->lx / cfg->nr;
    r->dz = cfg->lz / cfg->nz;
    r->dp = cfg->ly / cfg->np;
    r->nf = (size_t)r->ldi * r->ldj * (cfg->np + 2);
    auto fail = [&](int s) { vvt_destroy(r); return s; };
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) return fail(VV_EHIP);
    for (int q = 0; q < VV3_NFIELD; ++q) {
        if (hipMalloc(&r->f[q], r->nf * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
        if (hipMemsetAsync(r->f[q], 0, r->nf * sizeof(double), r->stream) != hipSuccess) return fail(VV_EHIP);
        r->clean[q] = true;
    }
    for (int d = 0; d < 12; ++d) {
        if (hipMalloc(&r->dv[d], r->nf * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
        if (hipMemsetAsync(r->dv[d], 0, r->nf * sizeof(double), r->stream) != hipSuccess) return fail(VV_EHIP);
    }
    {
        std::vector<double> tabs;
        for (int d = 0; d < 12; ++d) {
            const DSpec s = dspec(r, d);
            r->tab_off[d] = tabs.size();
            compact_tables(s.type, s.m, tabs);
        }
        if (upload(&r->tabs, tabs) != hipSuccess) return fail(VV_ENOMEM);
        // checkpoints of the derivative lines (y and z: 8 jobs, one per CK
        // rows; then x: 4 jobs, one per CX rows, in the same buffer)
        int segs = 0, nl = 0;
        for (int d = 0; d < 12; ++d) {
            const DSpec s = dspec(r, d);
            segs = std::max(segs, (s.m + ((d % 3 == 0) ? CX : CK) - 1) / ((d % 3 == 0) ? CX : CK));
            nl = std::max(nl, s.n1 * s.n2);
        }
        r->ckp_seg = segs;
        r->ckp_lines = nl;
        if (hipMalloc(&r->ckp, 8 * (size_t)r->ckp_seg * nl * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
    }
    {   // column sums (2 nr), trans1 and trans2, then the terms of rows k = np/2 .. np
        const size_t nh = 2 * (size_t)r->nr + 2 + 2 * (size_t)(cfg->np - cfg->np / 2 + 1) * r->nr;
        if (hipMalloc(&r->heat, nh * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
        if (hipMemsetAsync(r->heat, 0, nh * sizeof(double), r->stream) != hipSuccess) return fail(VV_EHIP);
    }
    if (hipMalloc(&r->kepart, (KE_MAXB + 1) * sizeof(double)) != hipSuccess) return fail(VV_ENOMEM);
    if (hipEventCreate(&r->ev[0]) != hipSuccess || hipEventCreate(&r->ev[1]) != hipSuccess) return fail(VV_EHIP);
    {
        const int st = setup_poisson(r);
        if (st) return fail(st);
    }
    if (hipStreamSynchronize(r->stream) != hipSuccess) return fail(VV_EHIP);
    *out = r;
    return VV_OK;
}

int vvt_destroy(vvt_run *r)
{
    if (!r) return VV_OK;
    {
        std::lock_guard<std::recursive_mutex> lk(r->mu);
        (void)hipSetDevice(r->device);
        if (r->stream) (void)hipStreamSynchronize(r->stream);
        for (double *p : r->f)
            if (p) (void)hipFree(p);
        for (double *p : r->dv)
            if (p) (void)hipFree(p);
        for (auto &O : r->op)
            for (double *p : {O.W, O.Wi, O.ly, O.V, O.Vinv, O.lz, O.rt})
                if (p) (void)hipFree(p);
        if (r->stream2) (void)hipStreamSynchronize(r->stream2);
        for (double *p : {r->tabs, r->S, r->T, r->cp, r->S2, r->T2, r->cp2, r->ckpx, r->heat, r->kepart})
            if (p) (void)hipFree(p);
        for (hipEvent_t e : {r->fork, r->join, r->forkd, r->joind})
            if (e) (void)hipEventDestroy(e);
        if (r->stream2) (void)hipStreamDestroy(r->stream2);
        if (r->ytw) (void)hipFree(r->ytw);
        if (r->ckp) (void)hipFree(r->ckp);
        for (hipEvent_t e : r->ev)
            if (e) (void)hipEventDestroy(e);
        if (r->stream) (void)hipStreamDestroy(r->stream);
    }
    delete r;
    return VV_OK;
}

int vvt_put(vvt_run *r, int field, const double *host, int ldi, int ldj)
{
    if (!r || !host || field < 0 || field >= VV3_NFIELD) { vv::set_error("vvt_put: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    r->clean[field] = false;
    return copy_field(r, r->f[field], const_cast<double *>(host), ldi, ldj, true);
}

int vvt_get(vvt_run *r, int field, double *host, int ldi, int ldj)
{
    if (!r || !host || field < 0 || field >= VV3_NFIELD) { vv::set_error("vvt_get: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    return copy_field(r, r->f[field], host, ldi, ldj, false);
}

int vvt_get_derivative(vvt_run *r, int d, double *host)
{
    if (!r || !host || d < 0 || d >= 12) { vv::set_error("vvt_get_derivative: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    return copy_field(r, r->dv[d], host, 0, 0, false);
}

#define VVT_CALL(expr)                                      \
    do {                                                    \
        if (!r) { vv::set_error("null run"); return VV_EINVAL; } \
        std::lock_guard<std::recursive_mutex> lk(r->mu);    \
        VV_HIP(hipSetDevice(r->device));                    \
        int st_ = (expr);                                   \
        if (st_) return st_;                                \
        VV_HIP(hipStreamSynchronize(r->stream));            \
        return VV_OK;                                       \
    } while (0)

int vvt_derivatives(vvt_run *r) { VVT_CALL(launch_derivatives(r)); }
int vvt_vorticity(vvt_run *r) { VVT_CALL(launch_explicit(r, 1)); }
int vvt_temperature(vvt_run *r) { VVT_CALL(launch_explicit(r, 2)); }
int vvt_update(vvt_run *r) { VVT_CALL(launch_update(r)); }

int vvt_rhs(vvt_run *r, int which)
{
    if (which < 1 || which > 3) { vv::set_error("vvt_rhs: which must be 1, 2 or 3"); return VV_EINVAL; }
    VVT_CALL(launch_rhs(r, which));
}

int vvt_poisson(vvt_run *r, int which)
{
    if (which < 1 || which > 3) { vv::set_error("vvt_poisson: which must be 1, 2 or 3"); return VV_EINVAL; }
    VVT_CALL(launch_poisson(r, which));
}

static int read_heat(vvt_run *r, double *trans1, double *trans2)
{
    double h[2];
    VV_HIP(hipMemcpyAsync(h, heat_out(r), 2 * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    if (trans1) *trans1 = h[0];
    if (trans2) *trans2 = h[1];
    return VV_OK;
}

int vvt_heat(vvt_run *r, double *trans1, double *trans2)
{
    if (!r) { vv::set_error("null run"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    const int st = launch_heat(r);
    if (st) return st;
    return read_heat(r, trans1, trans2);
}

int vvt_step(vvt_run *r, int n, double *trans1, double *trans2)
{
    if (!r || n < 0) { vv::set_error("vvt_step: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    for (int s = 0; s < n; ++s) {
        const int st = one_step(r);
        if (st) return st;
    }
    return read_heat(r, trans1, trans2);
}

int vvt_energy(vvt_run *r, double *ke)
{
    if (!r || !ke) { vv::set_error("vvt_energy: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    VV_HIP(hipSetDevice(r->device));
    const int nb = std::min(KE_MAXB, std::max(1, nblk((long long)r->nf, WG)));
    hipLaunchKernelGGL(k_t_energy, dim3(nb), dim3(WG), 0, r->stream, geo(r), (const double *)r->f[VV3_UEL],
                       (const double *)r->f[VV3_VEL], (const double *)r->f[VV3_WEL], r->kepart);
    hipLaunchKernelGGL(k_t_energy_sum, dim3(1), dim3(WG), 0, r->stream, (const double *)r->kepart, nb,
                       0.5 * r->dr * r->dz * r->dp, r->kepart + KE_MAXB);
    VV_HIP(hipGetLastError());
    VV_HIP(hipMemcpyAsync(ke, r->kepart + KE_MAXB, sizeof(double), hipMemcpyDeviceToHost, r->stream));
    VV_HIP(hipStreamSynchronize(r->stream));
    return VV_OK;
}

int vvt_profile(vvt_run *r, int enable)
{
    if (!r) { vv::set_error("null run"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    if (enable) {
        for (int k = 0; k < VVT_K_COUNT; ++k) {
            r->prof_n[k] = 0;
            r->prof_ms[k] = 0.0;
            r->prof_bytes[k] = 0.0;
        }
    }
    r->prof = enable != 0;
    return VV_OK;
}

int vvt_profile_read(vvt_run *r, int kernel, long long *launches, double *ms, double *bytes)
{
    if (!r || kernel < 0 || kernel >= VVT_K_COUNT) { vv::set_error("vvt_profile_read: bad argument"); return VV_EINVAL; }
    std::lock_guard<std::recursive_mutex> lk(r->mu);
    if (launches) *launches = r->prof_n[kernel];
    if (ms) *ms = r->prof_ms[kernel];
    if (bytes) *bytes = r->prof_bytes[kernel];
    return VV_OK;
}

}  // extern "C"
