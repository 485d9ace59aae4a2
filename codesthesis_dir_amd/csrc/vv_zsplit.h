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
//   inverse  S[q][j][i] = E + O and S[q][n-1-j][i] = E - O, j < n/2, with
//            E = sum_l G[0][j][l] T[q][i][l] (even modes) and
//            O = sum_l G[1][j][l] T[q][i][he + l] (odd modes); the middle
//            row of odd n is E alone.
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

// The inverse in one pass: a workgroup computes E and O of one 64 x 64
// block (rows j < he of the half tables, columns i) as two products, over
// the even modes and then over the odd ones, with two accumulators, and
// writes E + O at row j and E - O at row n-1-j (E alone on the middle row of
// odd n).  MFMA lane layout and operand staging as in k_gemm_f64.
template <int GK>
__global__ void __launch_bounds__(256) k_zinv_pair(const double *G, const double *T, double *S, int i1, int n)
{
    constexpr int TM = 64, TN = 64, NT = 256;
    constexpr int LA = GK * TM / NT, LB = GK * TN / NT;
    __shared__ double As[GK][TM + 1];
    __shared__ double Bs[GK][TN + 1];
    const int h = n / 2, he = n - h;
    const unsigned nmt = gridDim.y, nnt = gridDim.x;
    const unsigned lb = xcd_block((blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x,
                                  gridDim.x * gridDim.y * gridDim.z);
    const int q = (int)(lb / (nmt * nnt));
    const int m0 = (int)(lb % nmt) * TM, n0 = (int)((lb / nmt) % nnt) * TN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    const size_t mt = (size_t)i1 * n;
    const double *Tq = T + (size_t)q * mt;
    d4 acc[2][2][2];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) acc[b][r][c] = d4{0.0, 0.0, 0.0, 0.0};
    double ra[LA], rb[LB];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const int K = b ? h : he;
        const double *Gb = G + (size_t)b * he * he;
        const double *Tb = Tq + (size_t)b * he;
        auto fetch = [&](int k0) {
#pragma unroll
            for (int t = 0; t < LA; ++t) {
                const int e = tid + NT * t, m = m0 + e / GK, k = k0 + e % GK;
                ra[t] = (m < he && k < K) ? Gb[(size_t)m * he + k] : 0.0;
            }
#pragma unroll
            for (int t = 0; t < LB; ++t) {
                const int e = tid + NT * t, i = n0 + e / GK, k = k0 + e % GK;
                rb[t] = (i < i1 && k < K) ? Tb[(size_t)i * n + k] : 0.0;
            }
        };
        fetch(0);
        for (int k0 = 0; k0 < K; k0 += GK) {
#pragma unroll
            for (int t = 0; t < LA; ++t) {
                const int e = tid + NT * t;
                As[e % GK][e / GK] = ra[t];
            }
#pragma unroll
            for (int t = 0; t < LB; ++t) {
                const int e = tid + NT * t;
                Bs[e % GK][e / GK] = rb[t];
            }
            __syncthreads();
            if (k0 + GK < K) fetch(k0 + GK);
#pragma unroll
            for (int ks = 0; ks < GK; ks += 4) {
                double av[2], bv[2];
#pragma unroll
                for (int r = 0; r < 2; ++r) av[r] = As[ks + lk][wr + 16 * r + li];
#pragma unroll
                for (int c = 0; c < 2; ++c) bv[c] = Bs[ks + lk][wc + 16 * c + li];
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int c = 0; c < 2; ++c)
                        acc[b][r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r], bv[c], acc[b][r][c], 0, 0, 0);
            }
            __syncthreads();
        }
    }
    double *Sq = S + (size_t)q * mt;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + wr + 16 * r + lk + 4 * e;
                const int i = n0 + wc + 16 * c + li;
                if (i >= i1) continue;
                const double E = acc[0][r][c][e], O = acc[1][r][c][e];
                if (m < h) {
                    Sq[(size_t)m * i1 + i] = E + O;
                    Sq[(size_t)(n - 1 - m) * i1 + i] = E - O;
                } else if (m < he) {
                    Sq[(size_t)m * i1 + i] = E;
                }
            }
}

// The forward: for he > 64 on 128 x 128 tiles (for he <= 128 one tile holds
// all of a half's modes; measured faster than 64 x 128, 64 x 64 and
// 128 x 64 tiles and than one workgroup per block of both halves); smaller
// he: the usual tiles.  2 nq <= 65535.  The inverse: k_zinv_pair,
// nq <= 65535.
inline void zsplit_fwd(const double *F, const double *S, double *T, int i1, int n, int nq, hipStream_t s)
{
    const int h = n / 2, he = n - h;
    const ZFwdSplit op{F, i1, he, he, S, T, i1 * n, n, h};
    if (he > 64) gemm_f64_tile<false, true, 4, 4>(op, 2 * nq, s);
    else gemm_f64<false, true>(op, 2 * nq, s);
}

inline void zsplit_inv(const double *G, const double *T, double *S, int i1, int n, int nq, hipStream_t s)
{
    const int he = n - n / 2;
    hipLaunchKernelGGL(k_zinv_pair<16>, dim3((unsigned)((i1 + 63) / 64), (unsigned)((he + 63) / 64), (unsigned)nq),
                       dim3(256), 0, s, G, T, S, i1, n);
}

}  // namespace vv
