// vv_band.hip -- batched tdma / ptdma / pdma (include/vvband.h).
//
// One thread solves one line with the reference routine's operations in its
// order (rb.f:3046-3164; the library is compiled with -ffp-contract=off, so
// each a - f*c rounds twice, as in the reference).  The rows the reference
// overwrites in its argument arrays are written to a scratch buffer laid out
// row-major over the batch, scr[(w*n + j)*nsys + s] for word w of row j of
// system s, so consecutive threads touch consecutive addresses whatever the
// caller's layout.

#include <hip/hip_runtime.h>

#include "vv_internal.h"
#include "../../include/vvband.h"

namespace {

constexpr int WG = 256;

struct BandArgs {
    const double *a, *b, *c, *d, *e, *r;
    double *x;
    double *scr;
    long long ss, sr;
    int n, nsys;
};

// tdma (rb.f:3046-3064): forward elimination into b', r', back substitution
__global__ void __launch_bounds__(WG) k_band_tdma(BandArgs p)
{
    const int s = blockIdx.x * WG + threadIdx.x;
    if (s >= p.nsys) return;
    const long long o = (long long)s * p.ss, sr = p.sr;
    const size_t N = (size_t)p.nsys;
    double *bq = p.scr + s, *rq = p.scr + (size_t)p.n * N + s;
    double bj = p.b[o], rj = p.r[o];
    bq[0] = bj;
    rq[0] = rj;
    for (int j = 1; j < p.n; ++j) {
        const long long q = o + j * sr;
        const double fact = p.a[q] / bj;
        bj = p.b[q] - fact * p.c[q - sr];
        rj = p.r[q] - fact * rj;
        bq[j * N] = bj;
        rq[j * N] = rj;
    }
    double y = rj / bj;
    p.x[o + (long long)(p.n - 1) * sr] = y;
    for (int j = p.n - 2; j >= 0; --j) {
        const long long q = o + j * sr;
        y = (rq[j * N] - p.c[q] * y) / bq[j * N];
        p.x[q] = y;
    }
}

// ptdma (rb.f:3066-3116), n >= 3.  Rows 1..n-2 (1-based) are eliminated in
// order while row n accumulates the fill-in of the cyclic corner (cn, bn, rn
// stay in registers); a', b', r' of rows 1..n-2 go to scratch.
__global__ void __launch_bounds__(WG) k_band_ptdma(BandArgs p)
{
    const int s = blockIdx.x * WG + threadIdx.x;
    if (s >= p.nsys) return;
    const int n = p.n;
    const long long o = (long long)s * p.ss, sr = p.sr;
    const size_t N = (size_t)p.nsys;
    double *aq = p.scr + s, *bq = aq + (size_t)n * N, *rq = bq + (size_t)n * N;
    auto at = [&](const double *v, int j1) { return v[o + (long long)(j1 - 1) * sr]; };  // 1-based row
    // row 1 as given
    double aj = at(p.a, 1), bj = at(p.b, 1), rj = at(p.r, 1);
    aq[0] = aj;
    bq[0] = bj;
    rq[0] = rj;
    double cn = at(p.c, n), bn = at(p.b, n), rn = at(p.r, n);
    double cprev = at(p.c, 1);
    for (int j = 2; j <= n - 2; ++j) {
        double fact = at(p.a, j) / bj;
        const double bnew = at(p.b, j) - fact * cprev;
        const double anew = -fact * aj;
        const double rnew = at(p.r, j) - fact * rj;
        fact = cn / bj;
        cn = -fact * cprev;
        bn = bn - fact * aj;
        rn = rn - fact * rj;
        aj = anew;
        bj = bnew;
        rj = rnew;
        cprev = at(p.c, j);
        aq[(j - 1) * N] = aj;
        bq[(j - 1) * N] = bj;
        rq[(j - 1) * N] = rj;
    }
    // rows n-1 and n (aj, bj, rj, cprev now belong to row n-2)
    double fact = at(p.a, n - 1) / bj;
    const double b1 = at(p.b, n - 1) - fact * cprev;
    const double c1 = at(p.c, n - 1) - fact * aj;
    const double r1 = at(p.r, n - 1) - fact * rj;
    fact = cn / bj;
    double an = at(p.a, n) - fact * cprev;
    bn = bn - fact * aj;
    rn = rn - fact * rj;
    fact = an / b1;
    bn = bn - fact * c1;
    rn = rn - fact * r1;
    const double yn = rn / bn;
    double y = (r1 - c1 * yn) / b1;
    p.x[o + (long long)(n - 1) * sr] = yn;
    p.x[o + (long long)(n - 2) * sr] = y;
    for (int j = n - 2; j >= 1; --j) {
        const size_t w = (size_t)(j - 1) * N;
        y = (rq[w] - at(p.c, j) * y - aq[w] * yn) / bq[w];
        p.x[o + (long long)(j - 1) * sr] = y;
    }
}

// pdma (rb.f:3118-3164), n >= 2.  Row k is normalised by its pivot, then
// eliminated from rows k+1 and k+2, which are carried in registers; d', e',
// r' of rows 1..n-1 go to scratch.
__global__ void __launch_bounds__(WG) k_band_pdma(BandArgs p)
{
    const int s = blockIdx.x * WG + threadIdx.x;
    if (s >= p.nsys) return;
    const int n = p.n;
    const long long o = (long long)s * p.ss, sr = p.sr;
    const size_t N = (size_t)p.nsys;
    double *dq = p.scr + s, *eq = dq + (size_t)n * N, *rq = eq + (size_t)n * N;
    auto at = [&](const double *v, int j1) { return v[o + (long long)(j1 - 1) * sr]; };
    // row k (c, d, r) and row k+1 (b, c, d, r) as modified so far
    double ck = at(p.c, 1), dk = at(p.d, 1), rk = at(p.r, 1);
    double b1 = at(p.b, 2), c1 = at(p.c, 2), d1 = at(p.d, 2), r1 = at(p.r, 2);
    for (int k = 1; k <= n - 1; ++k) {
        const double dn = dk / ck, en = at(p.e, k) / ck, rnn = rk / ck;
        dq[(k - 1) * N] = dn;
        eq[(k - 1) * N] = en;
        rq[(k - 1) * N] = rnn;
        c1 = c1 - b1 * dn;
        if (k != n - 1) d1 = d1 - b1 * en;
        r1 = r1 - b1 * rnn;
        double b2 = 0.0, c2 = 0.0, d2 = 0.0, r2 = 0.0;
        if (k != n - 1) {
            const double a2 = at(p.a, k + 2);
            b2 = at(p.b, k + 2) - a2 * dn;
            c2 = at(p.c, k + 2) - a2 * en;
            r2 = at(p.r, k + 2) - a2 * rnn;
            d2 = at(p.d, k + 2);
        }
        ck = c1;
        dk = d1;
        rk = r1;
        b1 = b2;
        c1 = c2;
        d1 = d2;
        r1 = r2;
    }
    // ck, rk: row n
    double x2 = rk / ck;
    p.x[o + (long long)(n - 1) * sr] = x2;
    const size_t w1 = (size_t)(n - 2) * N;
    double x1 = rq[w1] - dq[w1] * x2;
    p.x[o + (long long)(n - 2) * sr] = x1;
    for (int j = n - 2; j >= 1; --j) {
        const size_t w = (size_t)(j - 1) * N;
        const double xj = rq[w] - dq[w] * x1 - eq[w] * x2;
        p.x[o + (long long)(j - 1) * sr] = xj;
        x2 = x1;
        x1 = xj;
    }
}

}  // namespace

namespace vv {

// scratch words per row of each kind
int band_words(int kind) { return kind == VV_BAND_TDMA ? 2 : 3; }

// argument checks shared by the entry points; returns VV_OK or VV_EINVAL
int band_check(int kind, int n, int nsys, long long ss, long long sr, const double *a, const double *b,
               const double *c, const double *d, const double *e, const double *r, const double *x)
{
    if (kind < VV_BAND_TDMA || kind > VV_BAND_PDMA) { set_error("vv_band: unknown kind"); return VV_EINVAL; }
    const int nmin = kind == VV_BAND_TDMA ? 1 : kind == VV_BAND_PTDMA ? 3 : 2;
    if (n < nmin || nsys < 0 || ss < 1 || sr < 1) { set_error("vv_band: bad sizes or strides"); return VV_EINVAL; }
    if (!b || !c || !r || !x || (n > 1 && !a) || (kind == VV_BAND_PTDMA && !a) ||
        (kind == VV_BAND_PDMA && (!a || !d || !e))) {
        set_error("vv_band: null array");
        return VV_EINVAL;
    }
    return VV_OK;
}

int band_launch(int kind, int n, int nsys, long long ss, long long sr, const double *a, const double *b,
                const double *c, const double *d, const double *e, const double *r, double *x, double *scr,
                hipStream_t st)
{
    if (nsys == 0) return VV_OK;
    BandArgs p{a, b, c, d, e, r, x, scr, ss, sr, n, nsys};
    const dim3 grid((nsys + WG - 1) / WG);
    if (kind == VV_BAND_TDMA) hipLaunchKernelGGL(k_band_tdma, grid, dim3(WG), 0, st, p);
    else if (kind == VV_BAND_PTDMA) hipLaunchKernelGGL(k_band_ptdma, grid, dim3(WG), 0, st, p);
    else hipLaunchKernelGGL(k_band_pdma, grid, dim3(WG), 0, st, p);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

}  // namespace vv

extern "C" int vv_band_solve(int kind, int n, int nsys, long long ss, long long sr, const double *a,
                             const double *b, const double *c, const double *d, const double *e,
                             const double *r, double *x, void *stream)
{
    int st = vv::band_check(kind, n, nsys, ss, sr, a, b, c, d, e, r, x);
    if (st != VV_OK || nsys == 0) return st;
    hipStream_t hs = (hipStream_t)stream;
    const size_t bytes = (size_t)vv::band_words(kind) * (size_t)n * (size_t)nsys * sizeof(double);
    double *scr = nullptr;
    if (hipMallocAsync((void **)&scr, bytes, hs) != hipSuccess) {
        (void)hipGetLastError();
        vv::set_error("vv_band_solve: scratch allocation failed");
        return VV_ENOMEM;
    }
    st = vv::band_launch(kind, n, nsys, ss, sr, a, b, c, d, e, r, x, scr, hs);
    const hipError_t fe = hipFreeAsync(scr, hs);
    if (st != VV_OK) return st;
    VV_HIP(fe);
    return VV_OK;
}

// ----------------------------------------------------- allocation-free form
// The scratch rows of vv_band_solve_noalloc live in this array, one copy per
// GPU for the life of the process.  Host code never passes its name to a
// kernel: the device address of each GPU's copy is looked up once per device
// (hipGetSymbolAddress, with that device current) and kept in g_scr.
namespace {
static __device__ __align__(256) double scratch[VV_BAND_NOALLOC_WORDS];

constexpr int MAX_DEV = 64;
std::mutex g_scr_mu;
double *g_scr[MAX_DEV] = {};
}  // namespace

extern "C" int vv_band_solve_noalloc(int kind, int n, int nsys, long long ss, long long sr, const double *a,
                                     const double *b, const double *c, const double *d, const double *e,
                                     const double *r, double *x, void *stream)
{
    int st = vv::band_check(kind, n, nsys, ss, sr, a, b, c, d, e, r, x);
    if (st != VV_OK) return st;
    const double words = (double)vv::band_words(kind) * (double)n * (double)nsys;
    if (words > (double)VV_BAND_NOALLOC_WORDS) {
        vv::set_error("vv_band_solve_noalloc: batch needs " + std::to_string((long long)words) +
                      " scratch words, more than VV_BAND_NOALLOC_WORDS (" +
                      std::to_string((long long)VV_BAND_NOALLOC_WORDS) + ")");
        return VV_EINVAL;
    }
    if (nsys == 0) return VV_OK;
    int dev = 0;
    VV_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEV) { vv::set_error("vv_band_solve_noalloc: device index out of range"); return VV_EINVAL; }
    double *scr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scr_mu);
        if (!g_scr[dev]) {
            void *p = nullptr;
            VV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            g_scr[dev] = (double *)p;
        }
        scr = g_scr[dev];
    }
    return vv::band_launch(kind, n, nsys, ss, sr, a, b, c, d, e, r, x, scr, (hipStream_t)stream);
}
