// vv3_internal.h -- shared declarations of the 3D step (not part of the ABI).
#pragma once

#include "vv_internal.h"

#include "../../include/vvmi3.h"

#include <vector>

namespace vv3 {


// grid and parameters handed to every kernel by value
struct Geo {
    int nr, nz, np, ldi, ldj;
    double dt, dr, dz, dp, Ra, Re, Ptl, Gr, r0;
    int model;   // enum vv3_model
};

// One of the three Poisson operators (1: w, 2: v, 3: u).  Unknowns are
// i = 1..i1, j = 1..j1, k = 1..np; an (r,z) plane holds m = i1*j1 of them.
// The operator is L = D_phi + I_z (x) R + Z (x) I_r: a real DFT in phi turns
// D_phi into a shift per Fourier mode, the z matrix Z = V diag(lz) V^-1 is
// diagonalised once, and what remains per (mode, z eigenvalue) is a banded
// system in r.
struct PoissonOp {
    int which = 0, i1 = 0, j1 = 0, m = 0;
    double e1 = 0.0, en = 0.0;   // r-rows 1 and i1: coefficient two points inward
    double *V = nullptr;         // j1 x j1, row-major: eigenvectors of Z
    double *Vinv = nullptr;      // j1 x j1, row-major
    double *rt = nullptr;        // 4 x i1: sub, diagonal, super, 1/(dp^2 r^2)
    double *lz = nullptr;        // j1 eigenvalues of Z
    bool zsplit = false;         // V, Vinv hold the parity halves G, F of z_split; lz
                                 // the even modes first
};

struct Poisson3 {
    int nr = 0, nz = 0, np = 0, nm = 0, ldi = 0, ldj = 0;
    PoissonOp op[3];
    double *W = nullptr;     // np x np: forward real DFT over k = 1..np; rows m = 0..np/2
                             // are cos(m phi), rows np/2 + m (0 < m < np/2) sin(m phi)
    double *Wi = nullptr;    // np x np: the inverse (weights and 1/np folded in)
    double *lphi = nullptr;  // nm: 2 cos(2 pi mode/np) - 2
    double *S = nullptr;     // np x mmax: phi spectra, then the z-inverse result
    double *T = nullptr;     // np x mmax: z spectra, (i, l) with l fastest
    double *cp = nullptr;    // nm x mmax: elimination factors of the r solves
    double *S2 = nullptr, *T2 = nullptr, *cp2 = nullptr;   // a second set (a solve beside another), or none
    double *tw = nullptr;    // np/2 complex (cos, -sin)(2 pi t/np): FFT twiddles (np a power of two)
    bool fft = false;        // phi transforms by FFT (np a power of two, 8..4096) instead of GEMM
};

// Host-side pieces, exposed so that they can be checked without a GPU.
// Rows of the r operator of `which` (rot3dmg.f:1698-1719, 1935-1961,
// 2133-2154 without their z and phi terms): sub, diagonal and super
// coefficients of row i (1..i1), and the coefficient two points inward on
// rows 1 and i1.
void r_rows(int which, int nr, double r0, double dr, std::vector<double> &sub, std::vector<double> &diag,
            std::vector<double> &sup, double &e1, double &en);
// The z matrix of `which` (scaled by dz^2): Dirichlet (1,-2,1) for w, with
// the (-2, 5/3, -0.2) wall closures for v and u; j1 x j1 row-major.
std::vector<double> z_matrix(int which, int j1);
// Real eigen-decomposition of a pentadiagonal n x n matrix with real,
// distinct eigenvalues: eigenvalues ascending, V (columns = eigenvectors,
// row-major), V^-1.  Returns false if the eigenvalues cannot be separated or
// the decomposition does not reproduce the matrix.
bool eigen_banded(const std::vector<double> &M, int n, std::vector<double> &lam, std::vector<double> &V,
                  std::vector<double> &Vinv);
// The z matrices are persymmetric (the same read from the other wall), so
// each eigenvector is even or odd under j -> n-1-j, and so is the matching
// row of V^-1.  The transforms then split into two half-size products: the
// even modes take the sums x_j + x_{n-1-j} (and the middle row of odd n),
// the odd modes the differences; back, E (even modes) and O (odd modes) give
// x_j = E_j + O_j and x_{n-1-j} = E_j - O_j.  z_split orders the eigenpairs
// even first (he = n - n/2 of them), then odd (n/2), permuting lam, and
// forms the half tables, he = n - n/2 square each, zero-padded:
// F[b][l][j] = V^-1[mode l of half b][j], G[b][j][l] = V[j][mode l of half b]
// (b = 0 even, 1 odd; j < he, resp. n/2).  False (lam unchanged) if a mode
// is not of one parity to 1e-9 of its largest entry.
bool z_split(int n, std::vector<double> &lam, const std::vector<double> &V, const std::vector<double> &Vinv,
             std::vector<double> &F, std::vector<double> &G);
// Whether a solver with n z rows takes the split transforms: from 96 rows
// (measured faster than the whole products at 127, 128, 255 and 256 rows,
// about 1 % slower at 31 and 32); VV_ZSPLIT=1 in the environment at create
// splits at every size the rows allow, VV_ZSPLIT=0 never.
bool zsplit_wanted(int n);

int poisson3_setup(Poisson3 &P, const Geo &g);
void poisson3_free(Poisson3 &P);
// solve L psi = rhs for operator `which`; psip <- old psi and walls 0 on the
// range the reference writes (rot3dmg.f:1294-1318, 1400-1424, 1509-1533)
// second: with the second set of work arrays (P.S2 != nullptr)
int poisson3_launch(const Poisson3 &P, const Geo &g, int which, double *psi, double *psip, const double *rhs,
                    hipStream_t s, bool second = false);

}  // namespace vv3
