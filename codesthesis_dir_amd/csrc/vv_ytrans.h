// vv_ytrans.h -- the mode transforms along k of fields with np cells between
// two walls, by FFT in LDS (kernels in vvt.hip): the tank's y axis and the
// barrier sector's phi.  Cells (u, w: zero flux through the walls) take the
// modes cos(pi q (k - 1/2)/np), q = 0..np-1, k = 1..np; faces (v: zero on
// the walls) sin(pi q k/np), q = 1..np-1, k = 1..np-1.
//   forward: S[q'][u] = sum_k e_q(k) rhs(i, j, k), q' = q (cells) or q - 1
//            (faces), u = (j-1) i1 + (i-1), i = 1..i1, j = 1..mt/i1;
//   inverse: psi(i, j, k) = sum_q w_q e_q(k) S[q'][u] on the same points,
//            w_0 = 1/np, w_q = 2/np (cells), w_q = 2/np (faces).
// np must be a power of two, 8..2048 (ytrans_fits).
#pragma once

#include <hip/hip_runtime.h>

namespace vvt {

struct YTrans {
    int np = 0, ldi = 0, ldj = 0, bits = 0, lp = 0;
    double2 *tw = nullptr;   // (cos, -sin)(2 pi t/(2 np)), t < np, then (cos, sin)(pi q/(2 np)), q < np
};

bool ytrans_fits(int np);
// tables on the current device; fields of ldi x ldj x (np+2) doubles
int ytrans_setup(YTrans &T, int np, int ldi, int ldj);
void ytrans_free(YTrans &T);
void ytrans_fwd(const YTrans &T, bool faces, const double *rhs, double *S, int i1, int mt, hipStream_t s);
void ytrans_inv(const YTrans &T, bool faces, const double *S, double *psi, int i1, int mt, hipStream_t s);

}  // namespace vvt
