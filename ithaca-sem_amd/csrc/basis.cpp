// Host-side 1D basis tables: Gauss-Lobatto-Legendre and Gauss-Radau-Jacobi points and
// weights, collocation derivative matrices, and the Modified_A / Modified_B (C0
// hierarchical, collapsed triangle direction) bases.
//
// Restates LibUtilities/Polylib/Polylib.cpp (jacobfd, Jacobz, zwglj :195-232, zwgrjm
// :119-144, Dglj :690-730) and LibUtilities/Foundations/Basis.cpp:392-509 (eModified_A,
// eModified_B), which is what the operators need at set-up time.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/ithaca_mf.h"

namespace {

void jacobfd(int np, const double *z, double *poly, double *polyd, int n, double a,
             double b)
{
    if (!np) return;
    if (n == 0) {
        for (int i = 0; i < np; ++i) {
            if (poly) poly[i] = 1.0;
            if (polyd) polyd[i] = 0.0;
        }
        return;
    }
    if (n == 1) {
        for (int i = 0; i < np; ++i) {
            if (poly) poly[i] = 0.5 * (a - b + (a + b + 2.0) * z[i]);
            if (polyd) polyd[i] = 0.5 * (a + b + 2.0);
        }
        return;
    }
    const double apb = a + b;
    std::vector<double> p1(np), p2(np), pv(np);
    double *P = poly ? poly : pv.data();
    for (int i = 0; i < np; ++i) {
        p2[i] = 1.0;
        p1[i] = 0.5 * (a - b + (a + b + 2.0) * z[i]);
    }
    for (int k = 2; k <= n; ++k) {
        double a1 = 2.0 * k * (k + apb) * (2.0 * k + apb - 2.0);
        double a2 = (2.0 * k + apb - 1.0) * (a * a - b * b) / a1;
        double a3 = (2.0 * k + apb - 2.0) * (2.0 * k + apb - 1.0) * (2.0 * k + apb) / a1;
        double a4 = 2.0 * (k + a - 1.0) * (k + b - 1.0) * (2.0 * k + apb) / a1;
        for (int i = 0; i < np; ++i) {
            P[i] = (a2 + a3 * z[i]) * p1[i] - a4 * p2[i];
            p2[i] = p1[i];
            p1[i] = P[i];
        }
    }
    if (polyd) {
        const double c4 = 2.0 * n + a + b;
        const double c1 = n * (a - b) / c4, c2 = n * (2.0 * n + a + b) / c4,
                     c3 = 2.0 * (n + a) * (n + b) / c4;
        for (int i = 0; i < np; ++i)
            polyd[i] = ((c1 - c2 * z[i]) * P[i] + c3 * p2[i]) / (1.0 - z[i] * z[i]);
    }
}

// Newton iteration with deflation from Chebyshev guesses (Polylib Jacobz).
void jacobz(int n, double *z, double a, double b)
{
    const double eps = 100 * 2.220446049250313e-16;
    const double dth = M_PI / (2.0 * n);
    double rlast = 0.0;
    for (int k = 0; k < n; ++k) {
        double r = -std::cos((2.0 * k + 1.0) * dth);
        if (k) r = 0.5 * (r + rlast);
        for (int it = 1; it < 30; ++it) {
            double p, dp, sum = 0.0;
            jacobfd(1, &r, &p, &dp, n, a, b);
            for (int i = 0; i < k; ++i) sum += 1.0 / (r - z[i]);
            const double delr = -p / (dp - sum * p);
            r += delr;
            if (std::fabs(delr) < eps) break;
        }
        z[k] = r;
        rlast = r;
    }
}

}  // namespace

extern "C" int imf_basis_modified_a(int nm, int nq, double *z, double *w, double *D,
                                    double *bdata, double *dbdata)
{
    if (nm < 2 || nq < 2 || nq > 64 || !z || !w || !D || !bdata || !dbdata) return IMF_ERR_ARG;
    // GLL zeros / weights
    z[0] = -1.0;
    z[nq - 1] = 1.0;
    if (nq > 2) jacobz(nq - 2, z + 1, 1.0, 1.0);
    jacobfd(nq, z, w, nullptr, nq - 1, 0.0, 0.0);
    const double fac = 2.0 / (double(nq) * double(nq - 1));
    for (int i = 0; i < nq; ++i) w[i] = fac / (w[i] * w[i]);
    // derivative matrix, flat D[a*nq + b] = l_a'(z_b)
    std::vector<double> pd(nq);
    pd[0] = 2.0 * std::pow(-1.0, nq) * double(nq - 1);
    if (nq > 2) {
        jacobfd(nq - 2, z + 1, pd.data() + 1, nullptr, nq - 3, 2.0, 2.0);
        for (int i = 1; i < nq - 1; ++i) pd[i] *= 0.5 * double(nq + 1) * (1.0 - z[i] * z[i]);
    }
    pd[nq - 1] = -2.0 * double(nq - 1);
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < nq; ++j) {
            if (i != j)
                D[i * nq + j] = pd[j] / (pd[i] * (z[j] - z[i]));
            else if (j == 0)
                D[i * nq + j] = -double(nq - 1) * double(nq) / 4.0;
            else if (j == nq - 1)
                D[i * nq + j] = double(nq - 1) * double(nq) / 4.0;
            else
                D[i * nq + j] = 0.0;
        }
    // Modified_A modes: vertices (1 -+ z)/2, interior (1-z)(1+z)/4 P_{p-2}^{(1,1)}(z)
    for (int i = 0; i < nq; ++i) {
        bdata[i] = 0.5 * (1 - z[i]);
        bdata[nq + i] = 0.5 * (1 + z[i]);
    }
    for (int p = 2; p < nm; ++p) {
        double *mode = bdata + p * nq;
        jacobfd(nq, z, mode, nullptr, p - 2, 1.0, 1.0);
        for (int i = 0; i < nq; ++i) mode[i] *= bdata[i] * bdata[nq + i];
    }
    // dbdata = D * bdata with D(i,k) = flat[k*nq + i]
    for (int p = 0; p < nm; ++p)
        for (int i = 0; i < nq; ++i) {
            double s = 0.0;
            for (int k = 0; k < nq; ++k) s += D[k * nq + i] * bdata[p * nq + k];
            dbdata[p * nq + i] = s;
        }
    return IMF_OK;
}

namespace {

// Gauss-Radau-Jacobi with z = -1 and weight (1 - z)^alpha (beta = 0): -1 and the zeros of
// P_{nq-1}^{(alpha,1)}; w_i = 2^alpha / (nq (nq + alpha)) (1 - z_i) / P_{nq-1}^{(alpha,0)}(z_i)^2
// (beta = 0: no end factor)
void gauss_radau_m(int nq, double alpha, double *z, double *w)
{
    if (nq == 1) {
        z[0] = 0.0;
        w[0] = 2.0;
        return;
    }
    z[0] = -1.0;
    jacobz(nq - 1, z + 1, alpha, 1.0);
    jacobfd(nq, z, w, nullptr, nq - 1, alpha, 0.0);
    const double fac = std::pow(2.0, alpha) / (double(nq) * (double(nq) + alpha));
    for (int i = 0; i < nq; ++i) w[i] = fac * (1.0 - z[i]) / (w[i] * w[i]);
}

// derivative matrix of the Lagrange basis through z, D[a*nq + b] = l_a'(z_b), from the
// barycentric weights; each l' row sums to zero
void lagrange_deriv(int nq, const double *z, double *D)
{
    std::vector<double> bw(nq);
    for (int a = 0; a < nq; ++a) {
        double p = 1.0;
        for (int c = 0; c < nq; ++c)
            if (c != a) p *= z[a] - z[c];
        bw[a] = 1.0 / p;
    }
    for (int b = 0; b < nq; ++b) {
        double diag = 0.0;
        for (int a = 0; a < nq; ++a) {
            if (a == b) continue;
            const double v = (bw[a] / bw[b]) / (z[b] - z[a]);
            D[a * nq + b] = v;
            diag -= v;
        }
        D[b * nq + b] = diag;
    }
}

// Modified_B on the points z, modes packed p-major (q fastest):
//   p = 0: (1-z)/2, (1+z)/2, (1-z)/2 (1+z)/2 P_{q-2}^{(1,1)}
//   p = 1: (1-z)/2, (1-z)/2 (1+z)/2 P_{q-1}^{(1,1)}
//   p >= 2: ((1-z)/2)^p, ((1-z)/2)^p (1+z)/2 P_{q-1}^{(2p-1,1)}
void modified_b_values(int nm, int nq, const double *z, double *bdata)
{
    std::vector<double> a(nq), bp(nq), pw(nq), P(nq);
    for (int i = 0; i < nq; ++i) {
        a[i] = 0.5 * (1.0 - z[i]);
        bp[i] = 0.5 * (1.0 + z[i]);
    }
    int mode = 0;
    for (int p = 0; p < nm; ++p) {
        if (p >= 2)
            for (int i = 0; i < nq; ++i) pw[i] = p == 2 ? a[i] * a[i] : pw[i] * a[i];
        for (int q = 0; q < nm - p; ++q, ++mode) {
            double *out = bdata + (size_t)mode * nq;
            if (p == 0 && q < 2) {
                for (int i = 0; i < nq; ++i) out[i] = q == 0 ? a[i] : bp[i];
            } else if (p == 1 && q == 0) {
                for (int i = 0; i < nq; ++i) out[i] = a[i];
            } else if (p >= 2 && q == 0) {
                for (int i = 0; i < nq; ++i) out[i] = pw[i];
            } else {
                const int n = p == 0 ? q - 2 : q - 1;
                const double alpha = p <= 1 ? 1.0 : 2.0 * p - 1.0;
                jacobfd(nq, z, P.data(), nullptr, n, alpha, 1.0);
                for (int i = 0; i < nq; ++i)
                    out[i] = P[i] * (p <= 1 ? a[i] : pw[i]) * bp[i];
            }
        }
    }
}

void apply_deriv(int nrows, int nq, const double *D, const double *bdata, double *dbdata)
{
    for (int m = 0; m < nrows; ++m)
        for (int i = 0; i < nq; ++i) {
            double sum = 0.0;
            for (int k = 0; k < nq; ++k) sum += D[k * nq + i] * bdata[m * nq + k];
            dbdata[m * nq + i] = sum;
        }
}

}  // namespace

extern "C" int imf_basis_modified_b(int nm, int nq, double *z, double *w, double *D,
                                    double *bdata, double *dbdata)
{
    if (nm < 2 || nq < 1 || nq > 64 || !z || !w || !D || !bdata || !dbdata) return IMF_ERR_ARG;
    gauss_radau_m(nq, 1.0, z, w);
    lagrange_deriv(nq, z, D);
    modified_b_values(nm, nq, z, bdata);
    apply_deriv(nm * (nm + 1) / 2, nq, D, bdata, dbdata);
    return IMF_OK;
}

extern "C" int imf_basis_modified_pyr_c(int nm, int nq, double *z, double *w, double *D,
                                        double *bdata, double *dbdata)
{
    if (nm < 2 || nq < 1 || nq > 64 || !z || !w || !D || !bdata || !dbdata) return IMF_ERR_ARG;
    gauss_radau_m(nq, 2.0, z, w);
    lagrange_deriv(nq, z, D);
    // mode (p, q, r), r < nm - max(p, q): phi^b_{max(p,q), r} when min(p, q) <= 1, else
    // ((1-z)/2)^{p+q-2}, then times (1+z)/2 P_{r-1}^{(2p+2q-3,1)}
    std::vector<double> bb((size_t)nm * (nm + 1) / 2 * nq), P(nq);
    modified_b_values(nm, nq, z, bb.data());
    double *mode = bdata;
    for (int p = 0; p < nm; ++p)
        for (int q = 0; q < nm; ++q) {
            const int nr = nm - std::max(p, q);
            if (p <= 1 || q <= 1) {
                const int s = std::max(p, q), start = s * nm - s * (s - 1) / 2;
                std::copy(bb.begin() + (size_t)start * nq, bb.begin() + (size_t)(start + nr) * nq,
                          mode);
                mode += (size_t)nr * nq;
            } else {
                const double *pw = mode;
                for (int i = 0; i < nq; ++i) mode[i] = std::pow(0.5 * (1.0 - z[i]), p + q - 2);
                mode += nq;
                for (int r = 1; r < nr; ++r, mode += nq) {
                    jacobfd(nq, z, P.data(), nullptr, r - 1, 2.0 * p + 2.0 * q - 3.0, 1.0);
                    for (int i = 0; i < nq; ++i) mode[i] = P[i] * pw[i] * 0.5 * (1.0 + z[i]);
                }
            }
        }
    apply_deriv(nm * (nm + 1) * (2 * nm + 1) / 6, nq, D, bdata, dbdata);
    return IMF_OK;
}

extern "C" int imf_basis_modified_c(int nm, int nq, double *z, double *w, double *D,
                                    double *bdata, double *dbdata)
{
    if (nm < 2 || nq < 1 || nq > 64 || !z || !w || !D || !bdata || !dbdata) return IMF_ERR_ARG;
    gauss_radau_m(nq, 2.0, z, w);
    lagrange_deriv(nq, z, D);
    // phi^c_{pqr} = phi^b_{p+q,r}: block p is the Modified_B groups p .. nm-1 in order
    std::vector<double> bb((size_t)nm * (nm + 1) / 2 * nq);
    modified_b_values(nm, nq, z, bb.data());
    size_t off = 0, boff = 0;
    for (int p = 0; p < nm; ++p) {
        const size_t n = (size_t)nq * (nm - p) * (nm - p + 1) / 2;
        std::copy(bb.begin() + boff, bb.begin() + boff + n, bdata + off);
        boff += (size_t)nq * (nm - p);
        off += n;
    }
    apply_deriv(nm * (nm + 1) * (nm + 2) / 6, nq, D, bdata, dbdata);
    return IMF_OK;
}
