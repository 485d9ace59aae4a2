// vv_fft.h -- radix-2 FFT stages in LDS over periodic lines, shared by the
// 2D and 3D solvers (device code; included by the .hip files only).
//
// A workgroup holds LP pairs of real lines; each pair is one complex line of
// n points (n a power of two, loaded in bit-reversed order; other even n
// through the mixed-radix step below).  Stage s (radix 2) pairs
// positions p and p + h (h = 2^s) with the twiddle e^{-i 2 pi (p mod h)/(2h)},
// taken from tw[t] = (cos, -sin)(2 pi t/n), t < n/2 (conjugated for the
// inverse).  Z_m and Z_{n-m} separate the pair's two real spectra.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

namespace vv {

__device__ __forceinline__ unsigned bitrev(unsigned x, int bits) { return __builtin_bitreverse32(x) >> (32 - bits); }

template <bool INV>
__device__ __forceinline__ void fft_bfly(double2 &u, double2 &v, double2 w)
{
    if (INV) w.y = -w.y;
    const double2 vw = make_double2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
    const double2 a = u;
    u = make_double2(a.x + vw.x, a.y + vw.y);
    v = make_double2(a.x - vw.x, a.y - vw.y);
}

// The radix-2 stages, three per pass through LDS: the 8 points a + m h
// (m < 8, h = 2^s, a = (group) 8h + pos, pos < h) pair only among themselves
// in stages s, s+1 and s+2, so one thread loads them, applies the three
// stages' butterflies in registers -- each the radix-2 butterfly of that
// stage with its twiddle, so the same operations and the same bits -- and
// stores them; one or two stages left over take a pass of their own.
template <bool INV>
__device__ void fft_stages(double2 *z, const double2 *tw, int np, int bits, int LP)
{
    int s = 0;
    for (; s + 3 <= bits; s += 3) {
        const int h = 1 << s;
        const int nb = LP * (np >> 3);
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const int pr = b >> (bits - 3), t = b & ((np >> 3) - 1);
            const int pos = t & (h - 1), grp = t >> s;
            double2 *za = z + pr * np + grp * 8 * h + pos;
            double2 x[8];
#pragma unroll
            for (int m = 0; m < 8; ++m) x[m] = za[m * h];
            const double2 w0 = tw[pos << (bits - 1 - s)];   // stage s: pairs (m, m+1), m even
            fft_bfly<INV>(x[0], x[1], w0);
            fft_bfly<INV>(x[2], x[3], w0);
            fft_bfly<INV>(x[4], x[5], w0);
            fft_bfly<INV>(x[6], x[7], w0);
            // stage s+1: pairs (m, m+2), m = 0, 1, 4, 5, at positions pos + (m mod 4) h
            const double2 w1a = tw[pos << (bits - 2 - s)], w1b = tw[(pos + h) << (bits - 2 - s)];
            fft_bfly<INV>(x[0], x[2], w1a);
            fft_bfly<INV>(x[1], x[3], w1b);
            fft_bfly<INV>(x[4], x[6], w1a);
            fft_bfly<INV>(x[5], x[7], w1b);
            // stage s+2: pairs (m, m+4), m < 4, at positions pos + m h
#pragma unroll
            for (int m = 0; m < 4; ++m) fft_bfly<INV>(x[m], x[m + 4], tw[(pos + m * h) << (bits - 3 - s)]);
#pragma unroll
            for (int m = 0; m < 8; ++m) za[m * h] = x[m];
        }
        __syncthreads();
    }
    if (bits - s == 2) {   // stages s, s+1 (the last two)
        const int h = 1 << s;
        const int nb = LP * (np >> 2);
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const int pr = b >> (bits - 2), t = b & ((np >> 2) - 1);
            const int pos = t & (h - 1), grp = t >> s;
            double2 *za = z + pr * np + grp * 4 * h + pos;
            double2 x0 = za[0], x1 = za[h], x2 = za[2 * h], x3 = za[3 * h];
            const double2 w0 = tw[pos << (bits - 1 - s)];
            fft_bfly<INV>(x0, x1, w0);
            fft_bfly<INV>(x2, x3, w0);
            fft_bfly<INV>(x0, x2, tw[pos << (bits - 2 - s)]);
            fft_bfly<INV>(x1, x3, tw[(pos + h) << (bits - 2 - s)]);
            za[0] = x0; za[h] = x1; za[2 * h] = x2; za[3 * h] = x3;
        }
        __syncthreads();
    } else if (bits - s == 1) {   // the last stage
        const int h = 1 << s;
        const int nb = LP * (np >> 1);
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const int pr = b >> (bits - 1), t = b & ((np >> 1) - 1);
            const int pos = t & (h - 1), grp = t >> s;
            double2 *za = z + pr * np + grp * 2 * h + pos;
            double2 x0 = za[0], x1 = za[h];
            fft_bfly<INV>(x0, x1, tw[pos << (bits - 1 - s)]);
            za[0] = x0; za[h] = x1;
        }
        __syncthreads();
    }
}

// Mixed radix, n = b * n1 points (n1 = 2^bits, b odd > 1): point k of a line
// is loaded into sub-line k mod b at position bitrev(k / b), fft_stages
// transforms the LP*b sub-lines of n1 points (twiddles of n1), and
// fft_combine forms, in place and in natural order,
//   X_k = sum_{j<b} W_n^{jk} A_j[k mod n1],  W_n^t = twc[t] = (cos, -sin)(2 pi t/n),
// which is the DFT of the line because W_n^{(b q + j) k} = W_{n1}^{q k} W_n^{j k}
// (conjugated twiddles give the inverse).  b = 1 is the plain radix-2 line.
__device__ __forceinline__ int mr_pos(int k, int b, int n1, int bits)
{
    if (b == 1) return (int)bitrev((unsigned)k, bits);
    const int q = k / b;
    return (k - q * b) * n1 + (int)bitrev((unsigned)q, bits);
}

// R outputs per thread: LP*n <= R*blockDim.x (checked by the host)
template <bool INV, int R>
__device__ void fft_combine(double2 *z, const double2 *twc, int n, int n1, int b, int LP)
{
    double2 acc[R];
    const int total = LP * n;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = (int)threadIdx.x + r * (int)blockDim.x;
        double re = 0.0, im = 0.0;
        if (e < total) {
            const int pr = e / n, k = e - pr * n;
            const double2 *zl = z + pr * n + (k & (n1 - 1));
            int t = 0;
            for (int j = 0; j < b; ++j) {
                double2 w = twc[t];
                if (INV) w.y = -w.y;
                const double2 v = zl[j * n1];
                re += v.x * w.x - v.y * w.y;
                im += v.x * w.y + v.y * w.x;
                t += k;
                if (t >= n) t -= n;
            }
        }
        acc[r] = make_double2(re, im);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = (int)threadIdx.x + r * (int)blockDim.x;
        if (e < total) z[e] = acc[r];
    }
    __syncthreads();
}

// pairs of lines per workgroup: LP np complex values within 32 KB (64 KB for
// the longest lines)
inline int fft_pairs(int np) { return np >= 4096 ? 1 : std::max(1, 2048 / np); }

}  // namespace vv
