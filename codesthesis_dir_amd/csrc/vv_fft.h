// vv_fft.h -- radix-2 FFT stages in LDS over periodic lines, shared by the
// 2D and 3D solvers (device code; included by the .hip files only).
//
// A workgroup holds LP pairs of real lines; each pair is one complex line of
// n points (n a power of two, loaded in bit-reversed order).  Stage s pairs
// positions p and p + h (h = 2^s) with the twiddle e^{-i 2 pi (p mod h)/(2h)},
// taken from tw[t] = (cos, -sin)(2 pi t/n), t < n/2 (conjugated for the
// inverse).  Z_m and Z_{n-m} separate the pair's two real spectra.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

namespace vv {

__device__ __forceinline__ unsigned bitrev(unsigned x, int bits) { return __builtin_bitreverse32(x) >> (32 - bits); }

template <bool INV>
__device__ void fft_stages(double2 *z, const double2 *tw, int np, int bits, int LP)
{
    for (int s = 0; s < bits; ++s) {
        const int h = 1 << s;
        const int nb = LP * (np >> 1);
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const int pr = b / (np >> 1), t = b - pr * (np >> 1);
            const int pos = t & (h - 1), grp = t >> s;
            const int i0 = pr * np + grp * 2 * h + pos, i1 = i0 + h;
            double2 w = tw[pos << (bits - 1 - s)];
            if (INV) w.y = -w.y;
            const double2 u = z[i0], v = z[i1];
            const double2 vw = make_double2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
            z[i0] = make_double2(u.x + vw.x, u.y + vw.y);
            z[i1] = make_double2(u.x - vw.x, u.y - vw.y);
        }
        __syncthreads();
    }
}

// pairs of lines per workgroup: LP np complex values within 32 KB (64 KB for
// the longest lines)
inline int fft_pairs(int np) { return np >= 4096 ? 1 : std::max(1, 2048 / np); }

}  // namespace vv
