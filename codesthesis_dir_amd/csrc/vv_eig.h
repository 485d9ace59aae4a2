// vv_eig.h -- eigen-decomposition of a real tridiagonal matrix whose
// off-diagonal products are positive (host code, header only).
//
// Such a matrix T (sub a[i] = T(i,i-1), diagonal b[i], super c[i] =
// T(i,i+1), a[i+1] c[i] > 0) is similar to the symmetric tridiagonal
// S = D T D^-1, D = diag(d), d[0] = 1, d[i+1] = d[i] sqrt(c[i] / a[i+1]),
// whose off-diagonal is sqrt(a[i+1] c[i]).  S = Q diag(lam) Q^T is found by
// the implicit QL iteration with Wilkinson shifts (EISPACK tql2; Bowdler,
// Martin, Reinsch and Wilkinson, Numer. Math. 11 (1968) 293-306), and
// T = V diag(lam) V^-1 with V = D^-1 Q, V^-1 = Q^T D: no matrix inverse is
// formed.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <vector>

namespace vv {

// Eigenvalues ascending; V (columns = eigenvectors) and Vinv row-major n x n.
// Returns false when the iteration does not converge or a product a[i+1]
// c[i] is not positive.
inline bool tridiag_eigen(const std::vector<double> &a, const std::vector<double> &b, const std::vector<double> &c,
                          std::vector<double> &lam, std::vector<double> &V, std::vector<double> &Vinv)
{
    const int n = (int)b.size();
    std::vector<double> dsc(n, 1.0), d(b), e(n, 0.0);
    for (int i = 0; i + 1 < n; ++i) {
        const double p = a[i + 1] * c[i];
        if (!(p > 0.0) || !std::isfinite(p)) return false;
        dsc[i + 1] = dsc[i] * std::sqrt(c[i] / a[i + 1]);
        e[i] = std::sqrt(p);
    }
    std::vector<double> Z((size_t)n * n, 0.0);   // Z[k*n + i]: row k of Q
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
    const double eps = std::numeric_limits<double>::epsilon();
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= eps * dd) break;
            }
            if (m != l) {
                if (++iter > 100) return false;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
                double s = 1.0, cc = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i];
                    const double bb = cc * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    cc = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * cc * bb;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = cc * r - bb;
                    for (int k = 0; k < n; ++k) {
                        double *zk = &Z[(size_t)k * n];
                        f = zk[i + 1];
                        zk[i + 1] = s * zk[i] + cc * f;
                        zk[i] = cc * zk[i] - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    std::vector<int> o(n);
    std::iota(o.begin(), o.end(), 0);
    std::sort(o.begin(), o.end(), [&](int x, int y) { return d[x] < d[y]; });
    lam.resize(n);
    V.assign((size_t)n * n, 0.0);
    Vinv.assign((size_t)n * n, 0.0);
    for (int l = 0; l < n; ++l) {
        lam[l] = d[o[l]];
        for (int k = 0; k < n; ++k) {
            const double q = Z[(size_t)k * n + o[l]];
            V[(size_t)k * n + l] = q / dsc[k];
            Vinv[(size_t)l * n + k] = q * dsc[k];
        }
    }
    for (double x : lam)
        if (!std::isfinite(x)) return false;
    return true;
}

}  // namespace vv
