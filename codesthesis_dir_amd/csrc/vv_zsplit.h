// vv_zsplit.h -- the z transforms of the 3D solvers, split by parity
// (device code and launchers; included by the .hip files only).
//
// The z rows of every Poisson operator read the same from either wall, so
// each eigenvector is even or odd under j -> n-1-j (vv3::z_split orders the
// modes even first and forms the half tables F, G: he = n - n/2 square,
// zero-padded).  Each transform is then two half-size products, batch
// 2q + b (b = 0 the even modes, 1 the odd ones) over the nq planes q of
// mt = i1 n values:
//   forward  T[q][i][b he + l] = sum_j F[b][l][j] (S[q][j][i] +- S[q][n-1-j][i])
//            over j < n/2, the middle row of odd n joining the even half alone;
//   inverse  E (b = 0) on the rows j < he and O (b = 1) on the rows n-1-j,
//            j < n/2, of S[q]; k_zcombine then forms S = E + O at j and
//            E - O at n-1-j.
#pragma once

#include "vv_gemm.h"

namespace vv {

struct ZFwdSplit {
    const double *F;        // [2][he][he]
    int M, N, K;            // i1, he, he
    const double *S;
    double *T;
    int mt, n, h;
    __device__ double a(int bq, int m, int k) const
    {
        // branch-free: the middle row of odd n (k = h) reads itself twice
        const int b = bq & 1;
        const bool pair = k < h;
        const double *X = S + (size_t)(bq >> 1) * mt + m;
        const double x = X[(size_t)k * M], y = X[(size_t)(pair ? n - 1 - k : k) * M];
        return b ? (pair ? x - y : 0.0) : (pair ? x + y : x);
    }
    __device__ double b(int bq, int k, int n_) const { return F[((size_t)(bq & 1) * K + n_) * K + k]; }
    __device__ void c(int bq, int m, int n_, double v) const
    {
        const int b = bq & 1;
        if (b == 0 || n_ < h) T[(size_t)(bq >> 1) * mt + (size_t)m * n + (size_t)b * K + n_] = v;
    }
};

struct ZInvSplit {
    const double *G;        // [2][he][he]
    int M, N, K;            // he, i1, he
    const double *T;
    double *S;
    int mt, n, h;
    __device__ double a(int bq, int m, int k) const { return G[((size_t)(bq & 1) * M + m) * K + k]; }
    __device__ double b(int bq, int k, int n_) const
    {
        const int b = bq & 1;
        return (b == 0 || k < h) ? T[(size_t)(bq >> 1) * mt + (size_t)n_ * n + (size_t)b * K + k] : 0.0;
    }
    __device__ void c(int bq, int m, int n_, double v) const
    {
        const int b = bq & 1;
        if (b == 0) S[(size_t)(bq >> 1) * mt + (size_t)m * N + n_] = v;
        else if (m < h) S[(size_t)(bq >> 1) * mt + (size_t)(n - 1 - m) * N + n_] = v;
    }
};

template <int NT>
__global__ void __launch_bounds__(NT) k_zcombine(double *S, int i1, int n, int mt)
{
    const int t = blockIdx.x * NT + threadIdx.x, h = n / 2;
    if (t >= h * i1) return;
    const int j = t / i1, i = t - j * i1;
    double *X = S + (size_t)blockIdx.y * mt + i;
    const double e = X[(size_t)j * i1], o = X[(size_t)(n - 1 - j) * i1];
    X[(size_t)j * i1] = e + o;
    X[(size_t)(n - 1 - j) * i1] = e - o;
}

// he > 64: 64 x 128 (forward) and 128 x 64 (inverse) tiles, so that for
// he <= 128 one tile holds all of a half's modes and each workgroup loads
// its rows of S, or of T, once; smaller he: the usual tiles.  2 nq <= 65535.
inline void zsplit_fwd(const double *F, const double *S, double *T, int i1, int n, int nq, hipStream_t s)
{
    const int h = n / 2, he = n - h;
    const ZFwdSplit op{F, i1, he, he, S, T, i1 * n, n, h};
    if (he > 64) gemm_f64_tile<false, true, 2, 4>(op, 2 * nq, s);
    else gemm_f64<false, true>(op, 2 * nq, s);
}

inline void zsplit_inv(const double *G, const double *T, double *S, int i1, int n, int nq, hipStream_t s)
{
    const int h = n / 2, he = n - h;
    const ZInvSplit op{G, he, i1, he, T, S, i1 * n, n, h};
    if (he > 64) gemm_f64_tile<true, true, 4, 2>(op, 2 * nq, s);
    else gemm_f64<true, true>(op, 2 * nq, s);
    if (h > 0)
        hipLaunchKernelGGL(k_zcombine<256>, dim3((unsigned)((h * i1 + 255) / 256), (unsigned)nq), dim3(256), 0, s, S,
                           i1, n, i1 * n);
}

}  // namespace vv
