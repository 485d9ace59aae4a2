// vv_gemm.h -- the FP64 matrix-core GEMM shared by the 2D and 3D solvers
// (device code and its launcher; included by the .hip files only).
#pragma once

#include "vv_internal.h"

namespace vv {

// Workgroups are dispatched to the 8 XCDs (each with its own L2) in turn:
// xcd_block maps the dispatch index b of nb workgroups to a logical index so
// that each XCD works through one contiguous range, and neighbouring rows
// and planes of a field are read through the same L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb)
{
    const unsigned x = b & 7u, q = b >> 3, per = nb >> 3, rem = nb & 7u;
    return x * per + (x < rem ? x : rem) + q;
}


typedef double d4 __attribute__((ext_vector_type(4)));

// C[q] = A[q] * B[q] in FP64 on the matrix cores, operands reached through
// the functor: op.a(q, m, k), op.b(q, k, n) load, op.c(q, m, n, v) stores; C
// is written with consecutive n in consecutive lanes.  A workgroup of WAVES
// waves computes a (32 WR) x (32 WC) tile (WR x WC = WAVES), each wave a
// 32 x 32 block as 2 x 2 v_mfma_f64_16x16x4f64 products.  MFMA lane layout:
// A[i][k] from lane i + 16k, B[k][j] from lane j + 16k, element e of D in
// lane l is D[l/16 + 4e][l%16].  The next k step's operands are loaded into
// registers while the current one is multiplied.  AKFAST / BKFAST: the tile
// loads run along k (operand contiguous in k) rather than along m / n.
// RB x CB: the 16 x 16 blocks of a wave (2 x 2 by default; wider waves
// load each operand element for more products).
template <class Op, bool AKFAST, bool BKFAST, int WAVES, int GK, int RB = 2, int CB = 2>
__global__ void __launch_bounds__(64 * WAVES) k_gemm_f64(Op op)
{
    constexpr int WR = (WAVES == 4) ? 2 : 1, WC = WAVES / WR;
    constexpr int TM = 16 * RB * WR, TN = 16 * CB * WC, NT = 64 * WAVES;
    constexpr int LA = GK * TM / NT, LB = GK * TN / NT;   // tile elements per thread
    // rows padded by one double: k-fast tile stores (consecutive threads one
    // row apart) then fall on different banks
    __shared__ double As[GK][TM + 1];
    __shared__ double Bs[GK][TN + 1];
    // logical tile order: m tiles fastest, then n tiles, then the batch, so
    // the m tiles that share a B tile run together on one XCD
    const unsigned nmt = gridDim.y, nnt = gridDim.x;
    const unsigned lb = xcd_block((blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x,
                                  gridDim.x * gridDim.y * gridDim.z);
    const int q = (int)(lb / (nmt * nnt));
    const int m0 = (int)(lb % nmt) * TM, n0 = (int)((lb / nmt) % nnt) * TN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave / WC) * 16 * RB, wc = (wave % WC) * 16 * CB;
    const int li = lane & 15, lk = lane >> 4;
    d4 acc[RB][CB];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < CB; ++c) acc[r][c] = d4{0.0, 0.0, 0.0, 0.0};
    double ra[LA], rb[LB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int t = 0; t < LA; ++t) {
            const int e = tid + NT * t;
            const int ka = AKFAST ? (e % GK) : (e / TM);
            const int ma = AKFAST ? (e / GK) : (e % TM);
            const int m = m0 + ma, k = k0 + ka;
            ra[t] = (m < op.M && k < op.K) ? op.a(q, m, k) : 0.0;
        }
#pragma unroll
        for (int t = 0; t < LB; ++t) {
            const int e = tid + NT * t;
            const int kb = BKFAST ? (e % GK) : (e / TN);
            const int nb = BKFAST ? (e / GK) : (e % TN);
            const int n = n0 + nb, k = k0 + kb;
            rb[t] = (n < op.N && k < op.K) ? op.b(q, k, n) : 0.0;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int t = 0; t < LA; ++t) {
            const int e = tid + NT * t;
            As[AKFAST ? (e % GK) : (e / TM)][AKFAST ? (e / GK) : (e % TM)] = ra[t];
        }
#pragma unroll
        for (int t = 0; t < LB; ++t) {
            const int e = tid + NT * t;
            Bs[BKFAST ? (e % GK) : (e / TN)][BKFAST ? (e / GK) : (e % TN)] = rb[t];
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < op.K; k0 += GK) {
        stash();
        __syncthreads();
        if (k0 + GK < op.K) fetch(k0 + GK);
#pragma unroll
        for (int ks = 0; ks < GK; ks += 4) {
            double a[RB], b[CB];
#pragma unroll
            for (int r = 0; r < RB; ++r) a[r] = As[ks + lk][wr + 16 * r + li];
#pragma unroll
            for (int c = 0; c < CB; ++c) b[c] = Bs[ks + lk][wc + 16 * c + li];
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int c = 0; c < CB; ++c) acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < CB; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + wr + 16 * r + lk + 4 * e;
                const int n = n0 + wc + 16 * c + li;
                if (m < op.M && n < op.N) op.c(q, m, n, acc[r][c][e]);
            }
}

// 64 x 64 tiles of four waves when they fill the chip, else 32 x 32 tiles of
// one wave (four times the workgroups for small planes).  GK is the k step:
// 16 by default; a deeper step (fewer, larger load rounds) suits short,
// latency-bound k loops.
template <bool AKFAST, bool BKFAST, class Op, int GK = 16>
void gemm_f64(const Op &op, int batch, hipStream_t s)
{
    const long long big = (long long)((op.N + 63) / 64) * ((op.M + 63) / 64) * batch;
    if (big >= 256) {
        dim3 grid((op.N + 63) / 64, (op.M + 63) / 64, batch);
        hipLaunchKernelGGL((k_gemm_f64<Op, AKFAST, BKFAST, 4, GK>), grid, dim3(256), 0, s, op);
    } else {
        dim3 grid((op.N + 31) / 32, (op.M + 31) / 32, batch);
        hipLaunchKernelGGL((k_gemm_f64<Op, AKFAST, BKFAST, 1, GK>), grid, dim3(64), 0, s, op);
    }
}

// Tiles of (32 RB) x (32 CB) for four waves, for products whose operand
// loads bound them (a short k loop over a large operand).
template <bool AKFAST, bool BKFAST, int RB, int CB, class Op, int GK = 16>
void gemm_f64_tile(const Op &op, int batch, hipStream_t s)
{
    constexpr int TM = 32 * RB, TN = 32 * CB;
    dim3 grid((op.N + TN - 1) / TN, (op.M + TM - 1) / TM, batch);
    hipLaunchKernelGGL((k_gemm_f64<Op, AKFAST, BKFAST, 4, GK, RB, CB>), grid, dim3(256), 0, s, op);
}

}  // namespace vv

namespace vv {

// Small products with a long k (few output tiles, K in the hundreds): each
// 32 x 32 output tile gets a workgroup of four waves, wave w accumulating
// the k range [w kc, (w+1) kc) with its own LDS tiles; the four partial
// tiles are summed through LDS at the end.  Four times the waves of the
// one-wave tiles on the same output, each with a quarter of the k loop.
template <class Op, bool AKFAST, bool BKFAST, int GK>
__global__ void __launch_bounds__(256) k_gemm_f64_ks(Op op)
{
    constexpr int TM = 32, TN = 32, NT = 64;
    constexpr int LA = GK * TM / NT, LB = GK * TN / NT;
    __shared__ double As[4][GK][TM + 1];
    __shared__ double Bs[4][GK][TN + 1];
    __shared__ double Red[3][TM * TN];
    const unsigned nmt = gridDim.y, nnt = gridDim.x;
    const unsigned lb = xcd_block((blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x,
                                  gridDim.x * gridDim.y * gridDim.z);
    const int q = (int)(lb / (nmt * nnt));
    const int m0 = (int)(lb % nmt) * TM, n0 = (int)((lb / nmt) % nnt) * TN;
    const int wave = threadIdx.x >> 6, tid = threadIdx.x & 63, lane = tid;
    const int li = lane & 15, lk = lane >> 4;
    const int kc = ((op.K + 4 * GK - 1) / (4 * GK)) * GK;
    const int kb = wave * kc, ke = min(op.K, kb + kc);
    d4 acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[r][c] = d4{0.0, 0.0, 0.0, 0.0};
    double ra[LA], rb[LB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int t = 0; t < LA; ++t) {
            const int e = tid + NT * t;
            const int ka = AKFAST ? (e % GK) : (e / TM);
            const int ma = AKFAST ? (e / GK) : (e % TM);
            const int m = m0 + ma, k = k0 + ka;
            ra[t] = (m < op.M && k < ke) ? op.a(q, m, k) : 0.0;
        }
#pragma unroll
        for (int t = 0; t < LB; ++t) {
            const int e = tid + NT * t;
            const int kk = BKFAST ? (e % GK) : (e / TN);
            const int nb = BKFAST ? (e / GK) : (e % TN);
            const int n = n0 + nb, k = k0 + kk;
            rb[t] = (n < op.N && k < ke) ? op.b(q, k, n) : 0.0;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int t = 0; t < LA; ++t) {
            const int e = tid + NT * t;
            As[wave][AKFAST ? (e % GK) : (e / TM)][AKFAST ? (e / GK) : (e % TM)] = ra[t];
        }
#pragma unroll
        for (int t = 0; t < LB; ++t) {
            const int e = tid + NT * t;
            Bs[wave][BKFAST ? (e % GK) : (e / TN)][BKFAST ? (e / GK) : (e % TN)] = rb[t];
        }
    };
    // every wave runs kc / GK rounds (the same count: the barriers match)
    fetch(kb);
    for (int k0 = kb; k0 < kb + kc; k0 += GK) {
        stash();
        __syncthreads();
        if (k0 + GK < kb + kc) fetch(k0 + GK);
#pragma unroll
        for (int ks = 0; ks < GK; ks += 4) {
            double a[2], b[2];
#pragma unroll
            for (int r = 0; r < 2; ++r) a[r] = As[wave][ks + lk][16 * r + li];
#pragma unroll
            for (int c = 0; c < 2; ++c) b[c] = Bs[wave][ks + lk][16 * c + li];
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int c = 0; c < 2; ++c) acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
        }
        __syncthreads();
    }
    if (wave > 0) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    Red[wave - 1][(16 * r + lk + 4 * e) * TN + 16 * c + li] = acc[r][c][e];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int mm = 16 * r + lk + 4 * e, nn = 16 * c + li;
                    const double v = acc[r][c][e] + Red[0][mm * TN + nn] + Red[1][mm * TN + nn] + Red[2][mm * TN + nn];
                    const int m = m0 + mm, n = n0 + nn;
                    if (m < op.M && n < op.N) op.c(q, m, n, v);
                }
    }
}

template <bool AKFAST, bool BKFAST, class Op, int GK = 16>
void gemm_f64_ksplit(const Op &op, int batch, hipStream_t s)
{
    dim3 grid((op.N + 31) / 32, (op.M + 31) / 32, batch);
    hipLaunchKernelGGL((k_gemm_f64_ks<Op, AKFAST, BKFAST, GK>), grid, dim3(256), 0, s, op);
}

}  // namespace vv
