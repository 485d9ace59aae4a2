// vv3_poisson.hip -- exact solves of rot3dmg.f's three Poisson operators
// (the residuals of relax1/2/3, rot3dmg.f:1695-1723, 1927-1965, 2130-2165)
// on gfx950.
//
// Each operator is L = D_phi + I_z (x) R + Z (x) I_r on the unknowns
// i = 1..i1, j = 1..j1, k = 1..np (k periodic): D_phi is the periodic second
// difference in phi divided by (dp r_i)^2, R the r rows, Z the z rows.  A
// solve is
//   1. a real DFT over k (a GEMM with a cos/sin table): one (r,z) problem
//      per Fourier mode and part,
//   2. the z eigenbasis: y = V^-1 x along j (a batched GEMM),
//   3. per (mode, z eigenvalue) one banded system in r, eliminated by a
//      thread (Thomas, with the two wall rows' extra coefficient folded in),
//   4. x = V y along j, the inverse DFT into psi.
// The result is the exact solution of the reference's discrete equations to
// rounding (the reference iterates multigrid V-cycles to an absolute
// residual of 1e-5, tolcon at rot3dmg.f:1562).

#include "vv3_internal.h"
#include "vv_fft.h"
#include "vv_gemm.h"
#include "vv_zsplit.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

#pragma clang fp contract(fast)

namespace vv3 {

namespace {

using vv::bitrev;
using vv::fft_combine;
using vv::fft_pairs;
using vv::fft_stages;
using vv::mr_pos;
using vv::gemm_f64;

// 1. spectra S[q][u] = sum_k W[q][k] rhs(i, j, k+1), u = (j-1) i1 + (i-1)
struct PhiFwd {
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

// 2. T[q][i][l] = sum_j S[q][j][i] Vinv[l][j]  (m = i, k = j, n = l)
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

// 4. S[q][j][i] = sum_l V[j][l] T[q][i][l]  (m = j, k = l, n = i)
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

// 5. psi(i, j, k+1) = sum_q Wi[k][q] S[q][u]; k = np also fills k = 0
struct PhiInv {
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
        const size_t off = (size_t)(j + 1) * ldi + (size_t)(i + 1);
        psi[(size_t)(m + 1) * ldj * ldi + off] = v;
        if (m + 1 == M) psi[off] = v;
    }
};

// ---------------------------------------------------------------- phi FFT
// For even np (8 <= np <= 4096) the phi transforms are FFTs in LDS: radix 2
// for a power of two (R = 0), otherwise np = b 2^bits with b odd, b radix-2
// sub-lines per line and the in-place combination of fft_combine (R outputs
// per thread).  Two real
// lines (consecutive i at one j) share one complex transform, z = a + i b;
// Z_m and Z_{np-m} separate them afterwards.  A workgroup holds LP such pairs
// (LP np complex values, at most 64 KB); k' = k mod np indexes the line, so
// k = np is position 0.  Stage s pairs positions p and p + h (h = 2^s) with
// the twiddle e^{-i 2 pi (p mod h)/(2h)} (conjugated for the inverse).

// forward: S[m][u] = sum_k rhs(i,j,k) cos(2 pi m k/np) (m = 0..np/2) and
// S[np/2+m][u] = sum_k rhs(i,j,k) sin(...) (0 < m < np/2), u = (j-1) i1 + (i-1)
template <int R>
__global__ void __launch_bounds__(256) k3_fft_fwd(const double *rhs, double *S, const double2 *tw, int np, int bits,
                                                   int b, int LP, int i1, int j1, int ldi, int ldj, int mt)
{
    extern __shared__ double2 z[];
    const int j = blockIdx.y, ib = blockIdx.x * 2 * LP, nl = 2 * LP;
    for (int e = threadIdx.x; e < nl * np; e += blockDim.x) {
        const int l = e % nl, kq = e / nl;
        const int i = ib + l, k = (kq == 0) ? np : kq;
        const double v = (i < i1) ? rhs[((size_t)k * ldj + (size_t)(j + 1)) * ldi + (size_t)(i + 1)] : 0.0;
        const int pos = R ? mr_pos(kq, b, np / b, bits) : (int)bitrev(kq, bits);
        double *zz = reinterpret_cast<double *>(&z[(l >> 1) * np + pos]);
        zz[l & 1] = v;
    }
    __syncthreads();
    fft_stages<false>(z, tw, np / b, bits, LP * b);
    if (R) fft_combine<false, R ? R : 1>(z, tw + (np / b) / 2, np, np / b, b, LP);
    const int half = np >> 1;
    for (int e = threadIdx.x; e < nl * (half + 1); e += blockDim.x) {
        const int l = e % nl, m = e / nl;
        const int i = ib + l;
        if (i >= i1) continue;
        const double2 zm = z[(l >> 1) * np + m], zn = z[(l >> 1) * np + (m == 0 ? 0 : np - m)];
        // A = (Z_m + conj Z_{np-m})/2, B = (Z_m - conj Z_{np-m})/(2i)
        double re, im;
        if ((l & 1) == 0) { re = 0.5 * (zm.x + zn.x); im = 0.5 * (zm.y - zn.y); }
        else { re = 0.5 * (zm.y + zn.y); im = -0.5 * (zm.x - zn.x); }
        const size_t u = (size_t)j * i1 + i;
        S[(size_t)m * mt + u] = re;                                  // sum x cos
        if (m > 0 && m < half) S[(size_t)(half + m) * mt + u] = -im;  // sum x sin
    }
}

// inverse: psi(i,j,k) = (1/np) sum_m w_m (C_m cos + S_m sin), k = 1..np, and
// k = 0 copied from k = np
template <int R>
__global__ void __launch_bounds__(256) k3_fft_inv(const double *S, double *psi, const double2 *tw, int np, int bits,
                                                   int b, int LP, int i1, int j1, int ldi, int ldj, int mt)
{
    extern __shared__ double2 z[];
    const int j = blockIdx.y, ib = blockIdx.x * 2 * LP, nl = 2 * LP, half = np >> 1;
    // Z_m = X_m + i Y_m with X_m = C_m - i S_m (m <= np/2), X_{np-m} = conj X_m
    for (int e = threadIdx.x; e < LP * np; e += blockDim.x) {
        const int pr = e % LP, m = e / LP;   // neighbouring threads: neighbouring lines, one spectral row
        const int mm = (m <= half) ? m : np - m;
        const double sg = (m <= half) ? -1.0 : 1.0;   // imaginary part of X_m is sg * S_mm
        double xr = 0, xi = 0, yr = 0, yi = 0;
        for (int h2 = 0; h2 < 2; ++h2) {
            const int i = ib + 2 * pr + h2;
            if (i >= i1) continue;
            const size_t u = (size_t)j * i1 + i;
            const double c = S[(size_t)mm * mt + u];
            const double sn = (mm > 0 && mm < half) ? S[(size_t)(half + mm) * mt + u] : 0.0;
            if (h2 == 0) { xr = c; xi = sg * sn; } else { yr = c; yi = sg * sn; }
        }
        const int pos = R ? mr_pos(m, b, np / b, bits) : (int)bitrev(m, bits);
        z[pr * np + pos] = make_double2(xr - yi, xi + yr);
    }
    __syncthreads();
    fft_stages<true>(z, tw, np / b, bits, LP * b);
    if (R) fft_combine<true, R ? R : 1>(z, tw + (np / b) / 2, np, np / b, b, LP);
    const double inv = 1.0 / np;
    for (int e = threadIdx.x; e < nl * np; e += blockDim.x) {
        const int l = e % nl, kq = e / nl;
        const int i = ib + l;
        if (i >= i1) continue;
        const double2 v = z[(l >> 1) * np + kq];
        const double x = ((l & 1) ? v.y : v.x) * inv;
        const size_t off = (size_t)(j + 1) * ldi + (size_t)(i + 1);
        if (kq == 0) {
            psi[(size_t)np * ldj * ldi + off] = x;
            psi[off] = x;
        } else {
            psi[(size_t)kq * ldj * ldi + off] = x;
        }
    }
}


// psip <- psi on the range the reference's routine writes, and psi <- 0 on
// that range's wall points (rot3dmg.f:1294-1318, 1400-1424, 1509-1533)
__global__ void __launch_bounds__(256) k3_psi_pre(Geo g, int which, int i1, int j1, double *psi, double *psip)
{
    const int nr = g.nr, nz = g.nz;
    const int tp = blockIdx.x * blockDim.x + threadIdx.x;
    if (tp >= g.ldi * g.ldj) return;
    const int j = tp / g.ldi, i = tp - j * g.ldi;
    const unsigned t = (unsigned)(blockIdx.y * g.ldi * g.ldj + tp);
    const int ri = (which == 3) ? nr : nr + 1;
    const int rj = (which == 1) ? nz : nz + 1;
    if (i > ri || j > rj) return;
    psip[t] = psi[t];
    if (!(i >= 1 && i <= i1 && j >= 1 && j <= j1)) psi[t] = 0.0;
}

// 3. per (mode, l): (R + lphi(mode) diag(1/(dp r)^2) + lz(l)) x = y for the
// cos part (spectral row mode) and, for 0 < mode < np/2, the sin part (row
// np/2 + mode), x overwriting y in T (layout [q][i][l]).  Row 1 has an extra
// coefficient e1 at i = 3, removed with row 2; row i1 has en at i1 - 2,
// removed with the eliminated row i1 - 2.  Thomas factors go to cp.
__global__ void __launch_bounds__(64) k3_rsolve(double *__restrict__ T, double *__restrict__ cp,
                                                 const double *__restrict__ rt, const double *__restrict__ lz,
                                                 const double *__restrict__ lphi, int i1, int j1, int np, double e1,
                                                 double en)
{
    const int nm = np / 2 + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nm * j1) return;
    const int mode = t / j1, l = t - mode * j1;
    const size_t mt = (size_t)i1 * j1;
    const bool two = (mode > 0 && 2 * mode < np);
    double *__restrict__ xc = T + (size_t)mode * mt + l;
    double *__restrict__ xs = T + (size_t)(two ? np / 2 + mode : mode) * mt + l;
    double *__restrict__ c = cp + (size_t)mode * mt + l;
    const double *sub = rt, *dia = rt + i1, *sup = rt + 2 * i1, *ir2 = rt + 3 * i1;
    const double mu = lphi[mode], lam = lz[l];
    auto diag = [&](int i) { return dia[i] + mu * ir2[i] + lam; };
    double d0 = diag(0), s0 = sup[0], yc = xc[0], ys = two ? xs[0] : 0.0;
    if (e1 != 0.0) {
        const double f = e1 / sup[1];
        d0 -= f * sub[1];
        s0 -= f * diag(1);
        yc -= f * xc[j1];
        if (two) ys -= f * xs[j1];
    }
    double cprev = s0 / d0;
    yc /= d0;
    ys /= d0;
    c[0] = cprev;
    xc[0] = yc;
    if (two) xs[0] = ys;
    // rows 1 .. i1-2 in groups of U with their loads issued ahead
    constexpr int U = 8;
    int i = 1;
    for (; i + U <= i1 - 1; i += U) {
        double rcu[U], rsu[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            rcu[u] = xc[(size_t)(i + u) * j1];
            rsu[u] = two ? xs[(size_t)(i + u) * j1] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int ii = i + u;
            const size_t o = (size_t)ii * j1;
            const double a = sub[ii];
            const double den = diag(ii) - a * cprev;
            cprev = sup[ii] / den;
            yc = (rcu[u] - a * yc) / den;
            ys = (rsu[u] - a * ys) / den;
            c[o] = cprev;
            xc[o] = yc;
            if (two) xs[o] = ys;
        }
    }
    for (; i < i1; ++i) {
        const size_t o = (size_t)i * j1;
        double a = sub[i], rc = xc[o], rs = two ? xs[o] : 0.0;
        if (i == i1 - 1 && en != 0.0) {
            const size_t o2 = (size_t)(i - 2) * j1;
            a -= en * c[o2];
            rc -= en * xc[o2];
            if (two) rs -= en * xs[o2];
        }
        const double den = diag(i) - a * cprev;
        cprev = sup[i] / den;
        yc = (rc - a * yc) / den;
        ys = (rs - a * ys) / den;
        c[o] = cprev;
        xc[o] = yc;
        if (two) xs[o] = ys;
    }
    i = i1 - 2;
    for (; i - U + 1 >= 0; i -= U) {
        double fu[U], xcu[U], xsu[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t o = (size_t)(i - u) * j1;
            fu[u] = c[o];
            xcu[u] = xc[o];
            xsu[u] = two ? xs[o] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t o = (size_t)(i - u) * j1;
            yc = xcu[u] - fu[u] * yc;
            xc[o] = yc;
            if (two) {
                ys = xsu[u] - fu[u] * ys;
                xs[o] = ys;
            }
        }
    }
    for (; i >= 0; --i) {
        const size_t o = (size_t)i * j1;
        const double f = c[o];
        yc = xc[o] - f * yc;
        xc[o] = yc;
        if (two) {
            ys = xs[o] - f * ys;
            xs[o] = ys;
        }
    }
}

// ---------------------------------------------------------------- host side

// LU with partial pivoting of (M - x I) for a pentadiagonal M, in band
// storage: row r keeps columns r-2 .. r+4 (fill from the row swaps).
struct BandLU {
    int n = 0;
    std::vector<double> b;
    std::vector<int> piv;
    double &at(int r, int c) { return b[(size_t)r * 7 + (c - r + 2)]; }
    // returns the sign of det(M - x I) (0 if a pivot vanishes)
    int factor(const std::vector<double> &M, int nn, double x)
    {
        n = nn;
        b.assign((size_t)n * 7, 0.0);
        piv.assign(n, 0);
        for (int r = 0; r < n; ++r)
            for (int c = std::max(0, r - 2); c <= std::min(n - 1, r + 2); ++c)
                at(r, c) = M[(size_t)r * n + c] - (r == c ? x : 0.0);
        int sign = 1;
        for (int c = 0; c < n; ++c) {
            int p = c;
            for (int r = c + 1; r <= std::min(n - 1, c + 2); ++r)
                if (std::fabs(at(r, c)) > std::fabs(at(p, c))) p = r;
            piv[c] = p;
            if (at(p, c) == 0.0) return 0;
            if (p != c) {
                for (int q = c; q <= std::min(n - 1, c + 4); ++q) std::swap(at(c, q), at(p, q));
                sign = -sign;
            }
            if (at(c, c) < 0.0) sign = -sign;
            for (int r = c + 1; r <= std::min(n - 1, c + 2); ++r) {
                const double f = at(r, c) / at(c, c);
                at(r, c) = f;
                for (int q = c + 1; q <= std::min(n - 1, c + 4); ++q) at(r, q) -= f * at(c, q);
            }
        }
        return sign;
    }
    void solve(std::vector<double> &y)
    {
        for (int c = 0; c < n; ++c) {
            if (piv[c] != c) std::swap(y[c], y[piv[c]]);
            for (int r = c + 1; r <= std::min(n - 1, c + 2); ++r) y[r] -= at(r, c) * y[c];
        }
        for (int c = n - 1; c >= 0; --c) {
            double s = y[c];
            for (int q = c + 1; q <= std::min(n - 1, c + 4); ++q) s -= at(c, q) * y[q];
            y[c] = s / at(c, c);
        }
    }
};

// Gauss-Jordan inverse with partial pivoting (row-major, n x n)
bool dense_inverse(std::vector<double> a, int n, std::vector<double> &inv)
{
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)p * n + c])) p = r;
        if (a[(size_t)p * n + c] == 0.0) return false;
        if (p != c)
            for (int q = 0; q < n; ++q) {
                std::swap(a[(size_t)c * n + q], a[(size_t)p * n + q]);
                std::swap(inv[(size_t)c * n + q], inv[(size_t)p * n + q]);
            }
        const double d = 1.0 / a[(size_t)c * n + c];
        for (int q = 0; q < n; ++q) {
            a[(size_t)c * n + q] *= d;
            inv[(size_t)c * n + q] *= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = a[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int q = 0; q < n; ++q) {
                a[(size_t)r * n + q] -= f * a[(size_t)c * n + q];
                inv[(size_t)r * n + q] -= f * inv[(size_t)c * n + q];
            }
        }
    }
    return true;
}

template <class T>
hipError_t upload(T **dst, const std::vector<T> &src)
{
    hipError_t e = hipMalloc(dst, src.size() * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
}

}  // namespace

void r_rows(int which, int nr, double r0, double dr, std::vector<double> &sub, std::vector<double> &diag,
            std::vector<double> &sup, double &e1, double &en)
{
    const int i1 = (which == 3) ? nr - 1 : nr;
    const double hhr = dr * dr;
    sub.assign(i1, 0.0);
    diag.assign(i1, 0.0);
    sup.assign(i1, 0.0);
    e1 = en = 0.0;
    for (int i = 1; i <= i1; ++i) {
        double a = 0, d = 0, c = 0;
        if (which == 1) {   // rot3dmg.f:1698-1712
            const double rv = r0 + ((double)i - 0.5) * dr;
            if (i == 1) { d = -2.0 / hhr - 0.5 / dr / rv; c = 5.0 / 3.0 / hhr + 0.5 / dr / rv; e1 = -0.2 / hhr; }
            else if (i == nr) { a = 5.0 / 3.0 / hhr - 0.5 / dr / rv; d = -2.0 / hhr + 0.5 / dr / rv; en = -0.2 / hhr; }
            else { a = 1.0 / hhr - 0.5 / dr / rv; d = -2.0 / hhr; c = 1.0 / hhr + 0.5 / dr / rv; }
        } else if (which == 2) {   // rot3dmg.f:1935-1950
            const double rv = r0 + ((double)i - 0.5) * dr;
            const double rr = 1.0 / (rv * rv);
            if (i == 1) { d = -2.0 / hhr + rr; c = 5.0 / 3.0 / hhr + 4.0 / 3.0 / dr / rv; e1 = -0.2 / hhr; }
            else if (i == nr) { a = 5.0 / 3.0 / hhr - 4.0 / 3.0 / dr / rv; d = -2.0 / hhr + rr; en = -0.2 / hhr; }
            else { a = 1.0 / hhr - 1.5 / dr / rv; d = -2.0 / hhr + rr; c = 1.0 / hhr + 1.5 / dr / rv; }
        } else {   // rot3dmg.f:2133-2136
            const double rv = (double)i * dr + r0;
            a = 1.0 / hhr - 1.5 / dr / rv;
            d = 1.0 / (rv * rv) - 2.0 / hhr;
            c = 1.0 / hhr + 1.5 / dr / rv;
        }
        sub[i - 1] = (i == 1) ? 0.0 : a;
        diag[i - 1] = d;
        sup[i - 1] = (i == i1) ? 0.0 : c;
    }
}

std::vector<double> z_matrix(int which, int j1)
{
    std::vector<double> M((size_t)j1 * j1, 0.0);
    for (int j = 0; j < j1; ++j) {
        M[(size_t)j * j1 + j] = -2.0;
        if (which != 1 && j == 0) {   // rot3dmg.f:1951-1952, 2143-2144
            M[1] = 5.0 / 3.0;
            M[2] = -0.2;
        } else if (which != 1 && j == j1 - 1) {
            M[(size_t)j * j1 + j - 1] = 5.0 / 3.0;
            M[(size_t)j * j1 + j - 2] = -0.2;
        } else {
            if (j > 0) M[(size_t)j * j1 + j - 1] = 1.0;
            if (j < j1 - 1) M[(size_t)j * j1 + j + 1] = 1.0;
        }
    }
    return M;
}

bool eigen_banded(const std::vector<double> &M, int n, std::vector<double> &lam, std::vector<double> &V,
                  std::vector<double> &Vinv)
{
    // Gershgorin interval [lo, hi]
    double lo = 0.0, hi = 0.0;
    for (int r = 0; r < n; ++r) {
        double rad = 0.0;
        for (int c = 0; c < n; ++c)
            if (c != r) rad += std::fabs(M[(size_t)r * n + c]);
        const double d = M[(size_t)r * n + r];
        lo = (r == 0) ? d - rad : std::min(lo, d - rad);
        hi = (r == 0) ? d + rad : std::max(hi, d + rad);
    }
    const double pad = 1e-3 * (hi - lo) + 1e-300;
    lo -= pad;
    hi += pad;
    const double R = std::max(std::fabs(lo), std::fabs(hi));
    const double cen = 0.5 * (lo + hi), half = 0.5 * (hi - lo);
    // the sign of det(M - x I) changes exactly at each (real, simple)
    // eigenvalue: scan a Chebyshev-spaced grid over the Gershgorin interval
    // (denser at its ends, where the eigenvalues of these second-difference
    // matrices crowd), refining until n changes are seen, then bisect each
    // bracket
    BandLU lu;
    const double pi = std::acos(-1.0);
    lam.clear();
    for (int G = 8 * n + 64; G <= (1 << 22); G *= 4) {
        lam.clear();
        std::vector<double> xs(G + 1);
        std::vector<int> sg(G + 1);
        for (int g = 0; g <= G; ++g) {
            xs[g] = cen - half * std::cos(pi * g / G);
            sg[g] = lu.factor(M, n, xs[g]);
        }
        int prev = sg[0], gprev = 0;
        for (int g = 1; g <= G; ++g) {
            if (sg[g] == 0) continue;
            if (prev != 0 && sg[g] != prev) {
                double a = xs[gprev], b = xs[g];
                const int sa = prev;
                for (int it = 0; it < 200 && b - a > 4e-16 * std::max(std::fabs(a), std::fabs(b)); ++it) {
                    const double mid = 0.5 * (a + b);
                    const int sm = lu.factor(M, n, mid);
                    if (sm == 0) { a = b = mid; break; }
                    if (sm == sa) a = mid; else b = mid;
                }
                lam.push_back(0.5 * (a + b));
            }
            prev = sg[g];
            gprev = g;
        }
        if ((int)lam.size() == n) break;
    }
    if ((int)lam.size() != n) return false;
    // eigenvectors by inverse iteration at a slightly shifted eigenvalue
    V.assign((size_t)n * n, 0.0);
    std::vector<double> y(n);
    for (int l = 0; l < n; ++l) {
        const double x = lam[l] + 1e-11 * R;
        if (lu.factor(M, n, x) == 0) return false;
        for (int i = 0; i < n; ++i) y[i] = 1.0 + 0.01 * std::sin(1.0 + i);
        for (int it = 0; it < 3; ++it) {
            lu.solve(y);
            double nrm = 0.0;
            for (double v : y) nrm += v * v;
            nrm = std::sqrt(nrm);
            if (!(nrm > 0.0)) return false;
            for (double &v : y) v /= nrm;
        }
        for (int i = 0; i < n; ++i) V[(size_t)i * n + l] = y[i];
    }
    if (!dense_inverse(V, n, Vinv)) return false;
    double e1 = 0.0, e2 = 0.0;
    for (int i = 0; i < n; ++i)
        for (int l = 0; l < n; ++l) {
            double mv = 0.0, iv = 0.0;
            for (int q = std::max(0, i - 2); q <= std::min(n - 1, i + 2); ++q) mv += M[(size_t)i * n + q] * V[(size_t)q * n + l];
            for (int q = 0; q < n; ++q) iv += Vinv[(size_t)i * n + q] * V[(size_t)q * n + l];
            e1 = std::max(e1, std::fabs(mv - lam[l] * V[(size_t)i * n + l]));
            e2 = std::max(e2, std::fabs(iv - (i == l ? 1.0 : 0.0)));
        }
    return e1 <= 1e-11 * R && e2 <= 1e-10;
}

bool z_split(int n, std::vector<double> &lam, const std::vector<double> &V, const std::vector<double> &Vinv,
             std::vector<double> &F, std::vector<double> &G)
{
    const int h = n / 2, he = n - h;
    std::vector<int> ev, od;
    for (int l = 0; l < n; ++l) {
        // parity of column l of V and of row l of V^-1 (the same, since
        // the reversal J commutes with the matrix: J V = V D, V^-1 J = D V^-1)
        double s = 0.0, d = 0.0, nv = 0.0, si = 0.0, di = 0.0, ni = 0.0;
        for (int j = 0; j < n; ++j) {
            const double a = V[(size_t)j * n + l], b = V[(size_t)(n - 1 - j) * n + l];
            const double c = Vinv[(size_t)l * n + j], e = Vinv[(size_t)l * n + (n - 1 - j)];
            s = std::max(s, std::fabs(a + b)); d = std::max(d, std::fabs(a - b)); nv = std::max(nv, std::fabs(a));
            si = std::max(si, std::fabs(c + e)); di = std::max(di, std::fabs(c - e)); ni = std::max(ni, std::fabs(c));
        }
        if (d <= 1e-9 * nv && di <= 1e-9 * ni) ev.push_back(l);
        else if (s <= 1e-9 * nv && si <= 1e-9 * ni) od.push_back(l);
        else return false;
    }
    if ((int)ev.size() != he || (int)od.size() != h) return false;
    std::vector<double> lp(n);
    F.assign((size_t)2 * he * he, 0.0);
    G.assign((size_t)2 * he * he, 0.0);
    for (int b = 0; b < 2; ++b) {
        const std::vector<int> &P = b ? od : ev;
        const int nj = b ? h : he;   // odd modes vanish on the middle row of odd n
        for (int l = 0; l < (int)P.size(); ++l) {
            lp[(size_t)b * he + l] = lam[P[l]];
            for (int j = 0; j < nj; ++j) {
                F[((size_t)b * he + l) * he + j] = Vinv[(size_t)P[l] * n + j];
                G[((size_t)b * he + j) * he + l] = V[(size_t)j * n + P[l]];
            }
        }
    }
    lam.swap(lp);
    return true;
}

bool zsplit_wanted(int n)
{
    const char *e = std::getenv("VV_ZSPLIT");
    if (e && e[0] == '0') return false;
    if (e && e[0] == '1') return true;
    return n >= 96;
}

int poisson3_setup(Poisson3 &P, const Geo &g)
{
    using vv::set_error;
    P.nr = g.nr; P.nz = g.nz; P.np = g.np; P.ldi = g.ldi; P.ldj = g.ldj;
    const int np = g.np, nm = np / 2 + 1;
    P.nm = nm;
    const double pi = std::acos(-1.0);
    // phi transforms by the LDS FFT (even np up to 4096), else dense GEMMs
    // with the tables W, Wi: the angle of mode m at k is 2 pi ((m k) mod
    // np) / np; spectral rows q = m (cos, m = 0..np/2), q = np/2 + m (sin,
    // m = 1..np/2-1)
    P.fft = (np >= 8 && np <= 4096 && np % 2 == 0);
    std::vector<double> lphi(nm);
    for (int m = 0; m < nm; ++m) lphi[m] = 2.0 * std::cos(2.0 * pi * m / np) - 2.0;
    if (!P.fft) {
        std::vector<double> W((size_t)np * np), Wi((size_t)np * np);
        for (int m = 0; m < nm; ++m) {
            const double wgt = (m == 0 || 2 * m == np) ? 1.0 : 2.0;
            const bool sin_row = (m > 0 && 2 * m < np);
            for (int k = 1; k <= np; ++k) {
                const double ang = 2.0 * pi * (double)(((long long)m * k) % np) / np;
                const double cs = std::cos(ang), sn = std::sin(ang);
                W[(size_t)m * np + (k - 1)] = cs;
                Wi[(size_t)(k - 1) * np + m] = wgt * cs / np;
                if (sin_row) {
                    W[(size_t)(np / 2 + m) * np + (k - 1)] = sn;
                    Wi[(size_t)(k - 1) * np + np / 2 + m] = wgt * sn / np;
                }
            }
        }
        VV_HIP(upload(&P.W, W));
        VV_HIP(upload(&P.Wi, Wi));
    }
    if (P.fft) {
        // (cos, -sin)(2 pi t/n1), t < n1/2, for the n1 = 2^bits point
        // radix-2 sub-lines, then (np not a power of two) (cos, -sin)(2 pi t/np),
        // t < np, for the mixed-radix combination
        const int n1 = np & -np;
        std::vector<double> tw;
        for (int t = 0; t < n1 / 2; ++t) {
            tw.push_back(std::cos(2.0 * pi * t / n1));
            tw.push_back(-std::sin(2.0 * pi * t / n1));
        }
        if (n1 != np)
            for (int t = 0; t < np; ++t) {
                tw.push_back(std::cos(2.0 * pi * t / np));
                tw.push_back(-std::sin(2.0 * pi * t / np));
            }
        VV_HIP(upload(&P.tw, tw));
    }
    VV_HIP(upload(&P.lphi, lphi));
    const double hhz = g.dz * g.dz, hhp = g.dp * g.dp;
    size_t mmax = 0;
    for (int w = 1; w <= 3; ++w) {
        PoissonOp &O = P.op[w - 1];
        O.which = w;
        O.i1 = (w == 3) ? g.nr - 1 : g.nr;
        O.j1 = (w == 1) ? g.nz - 1 : g.nz;
        O.m = O.i1 * O.j1;
        mmax = std::max(mmax, (size_t)O.m);
        std::vector<double> sub, dia, sup;
        r_rows(w, g.nr, g.r0, g.dr, sub, dia, sup, O.e1, O.en);
        std::vector<double> rt((size_t)4 * O.i1);
        for (int i = 0; i < O.i1; ++i) {
            const double rv = (w == 3) ? (double)(i + 1) * g.dr + g.r0 : g.r0 + ((double)(i + 1) - 0.5) * g.dr;
            rt[i] = sub[i];
            rt[O.i1 + i] = dia[i];
            rt[2 * O.i1 + i] = sup[i];
            rt[3 * O.i1 + i] = 1.0 / hhp / (rv * rv);
        }
        std::vector<double> lam, V, Vinv;
        if (!eigen_banded(z_matrix(w, O.j1), O.j1, lam, V, Vinv)) {
            set_error("vv3_create: z eigen-decomposition of a Poisson operator failed");
            return VV_EINVAL;
        }
        for (double &v : lam) v /= hhz;
        std::vector<double> Fz, Gz;
        O.zsplit = zsplit_wanted(O.j1) && 2 * np <= 65535 && z_split(O.j1, lam, V, Vinv, Fz, Gz);
        VV_HIP(upload(&O.V, O.zsplit ? Gz : V));
        VV_HIP(upload(&O.Vinv, O.zsplit ? Fz : Vinv));
        VV_HIP(upload(&O.rt, rt));
        VV_HIP(upload(&O.lz, lam));
    }
    VV_HIP(hipMalloc(&P.S, (size_t)np * mmax * sizeof(double)));
    VV_HIP(hipMalloc(&P.T, (size_t)np * mmax * sizeof(double)));
    VV_HIP(hipMalloc(&P.cp, (size_t)nm * mmax * sizeof(double)));
    if (hipMalloc(&P.S2, (size_t)np * mmax * sizeof(double)) != hipSuccess ||
        hipMalloc(&P.T2, (size_t)np * mmax * sizeof(double)) != hipSuccess ||
        hipMalloc(&P.cp2, (size_t)nm * mmax * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        for (double **q : {&P.S2, &P.T2, &P.cp2})
            if (*q) { (void)hipFree(*q); *q = nullptr; }
    }
    return VV_OK;
}

void poisson3_free(Poisson3 &P)
{
    for (double *p : {P.W, P.Wi, P.lphi, P.S, P.T, P.cp, P.S2, P.T2, P.cp2, P.tw})
        if (p) (void)hipFree(p);
    P.W = P.Wi = P.lphi = P.S = P.T = P.cp = P.S2 = P.T2 = P.cp2 = P.tw = nullptr;
    for (auto &O : P.op) {
        for (double *p : {O.V, O.Vinv, O.rt, O.lz})
            if (p) (void)hipFree(p);
        O.V = O.Vinv = O.rt = O.lz = nullptr;
    }
}

int poisson3_launch(const Poisson3 &P, const Geo &g, int which, double *psi, double *psip, const double *rhs,
                    hipStream_t s, bool second)
{
    const PoissonOp &O = P.op[which - 1];
    double *const S = second ? P.S2 : P.S, *const T = second ? P.T2 : P.T, *const cpw = second ? P.cp2 : P.cp;
    const int nm = P.nm, np = P.np, i1 = O.i1, j1 = O.j1, mt = O.m;
    hipLaunchKernelGGL(k3_psi_pre, dim3((unsigned)((g.ldi * g.ldj + 255) / 256), np + 1), dim3(256), 0, s, g, which,
                       i1, j1, psi, psip);
    // LP line pairs per workgroup, few enough for about 256 workgroups
    const int bits = P.fft ? __builtin_ctz((unsigned)np) : 0;
    const int LP = std::max(1, std::min(fft_pairs(np), (i1 * j1) / 512));
    const int fodd = np >> bits;                      // odd factor of np (1: radix 2 only)
    // registers of the mixed-radix combination: LP np <= R 256
    const int fR = (fodd == 1) ? 0 : (LP * np <= 8 * 256) ? 8 : 32;
    if (P.fft && LP * np > 32 * 256) { vv::set_error("phi FFT line longer than 8192 points"); return VV_EINVAL; }
    const dim3 fgrid((unsigned)((i1 + 2 * LP - 1) / (2 * LP)), (unsigned)j1);
    const size_t flds = (size_t)LP * np * sizeof(double2);
    const double2 *tw = (const double2 *)P.tw;
    if (P.fft && fR == 0)
        hipLaunchKernelGGL(k3_fft_fwd<0>, fgrid, dim3(256), flds, s, rhs, S, tw, np, bits, 1, LP, i1, j1, g.ldi,
                           g.ldj, mt);
    else if (P.fft && fR == 8)
        hipLaunchKernelGGL(k3_fft_fwd<8>, fgrid, dim3(256), flds, s, rhs, S, tw, np, bits, fodd, LP, i1, j1, g.ldi,
                           g.ldj, mt);
    else if (P.fft)
        hipLaunchKernelGGL(k3_fft_fwd<32>, fgrid, dim3(256), flds, s, rhs, S, tw, np, bits, fodd, LP, i1, j1,
                           g.ldi, g.ldj, mt);
    else
        gemm_f64<true, false>(PhiFwd{P.W, np, mt, np, rhs, S, i1, g.ldi, g.ldj}, 1, s);
    if (O.zsplit) vv::zsplit_fwd(O.Vinv, S, T, i1, j1, np, s);
    else gemm_f64<false, true>(ZFwd{O.Vinv, i1, j1, j1, S, T, mt}, np, s);
    hipLaunchKernelGGL(k3_rsolve, dim3((unsigned)((nm * j1 + 63) / 64)), dim3(64), 0, s, T, cpw,
                       (const double *)O.rt, (const double *)O.lz, (const double *)P.lphi, i1, j1, np, O.e1, O.en);
    if (O.zsplit) vv::zsplit_inv(O.V, T, S, i1, j1, np, s);
    else gemm_f64<true, true>(ZInv{O.V, j1, i1, j1, T, S, mt}, np, s);
    if (P.fft && fR == 0)
        hipLaunchKernelGGL(k3_fft_inv<0>, fgrid, dim3(256), flds, s, (const double *)S, psi, tw, np, bits, 1, LP, i1,
                           j1, g.ldi, g.ldj, mt);
    else if (P.fft && fR == 8)
        hipLaunchKernelGGL(k3_fft_inv<8>, fgrid, dim3(256), flds, s, (const double *)S, psi, tw, np, bits, fodd, LP,
                           i1, j1, g.ldi, g.ldj, mt);
    else if (P.fft)
        hipLaunchKernelGGL(k3_fft_inv<32>, fgrid, dim3(256), flds, s, (const double *)S, psi, tw, np, bits, fodd,
                           LP, i1, j1, g.ldi, g.ldj, mt);
    else
        gemm_f64<true, false>(PhiInv{P.Wi, np, mt, np, S, psi, i1, g.ldi, g.ldj}, 1, s);
    VV_HIP(hipGetLastError());
    return VV_OK;
}

}  // namespace vv3
