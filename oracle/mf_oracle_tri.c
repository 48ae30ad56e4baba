/*
 * CPU oracle, triangles (TEST INFRASTRUCTURE; see mf_oracle.h).  Plain-C restatement of
 * the reference's collapsed-coordinate triangle path:
 *   Gauss-Radau-Jacobi points  LibUtilities/Polylib/Polylib.cpp:119-144 (zwgrjm)
 *   Modified_B basis           LibUtilities/Foundations/Basis.cpp:423-509
 *   straight-sided TriGeom     SpatialDomains/TriGeom.cpp (linear map of the vertices),
 *                              GeomFactors.cpp:420-474 (derivative factors)
 *   kernels                    MatrixFreeOps/BwdTransKernels.hpp:77-126,
 *                              IProductKernels.hpp:134-236, PhysDerivKernels.hpp:40-90,
 *                              :150-216, Helmholtz.h:504-620 (HelmholtzTriImpl)
 * with the loop and summation order of those kernels (element interleave width 1).
 *
 * Layouts: coefficients u[e*M + mode], mode = sum_{p'<p} (nm - p') + q (q fastest),
 * M = nm (nm + 1) / 2; points f[e*Q + j*nq0 + i], i along eta0 (nq0 GLL points), j along
 * eta1 (nq1 Gauss-Radau points), Q = nq0 nq1.  w1 is the Helper's scaled weight (0.5 times
 * the Gauss-Radau alpha = 1 weight, MatrixFreeOps/Operator.hpp:244-258).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mf_oracle.h"

void or_zwgrjm(double *z, double *w, int np, double alpha, double beta)
{
    if (np == 1) {
        z[0] = 0.0;
        w[0] = 2.0;
        return;
    }
    z[0] = -1.0;
    or_jacobz(np - 1, z + 1, alpha, beta + 1.0);
    or_jacobfd(np, z, w, NULL, np - 1, alpha, beta);
    double fac = pow(2.0, alpha + beta) * tgamma(alpha + np) * tgamma(beta + np) /
                 (tgamma((double)np) * tgamma(np + alpha + beta + 1.0));
    fac /= beta + np;
    for (int i = 0; i < np; ++i) w[i] = fac * (1.0 - z[i]) / (w[i] * w[i]);
    w[0] *= beta + 1.0;
}

void or_Dnodes(double *D, const double *z, int np)
{
    /* barycentric weights b_a = 1 / prod_{c != a} (z_a - z_c); l_a'(z_b) = (b_a / b_b) /
     * (z_b - z_a) for a != b, and the diagonal makes every row of l' sum to zero */
    double *bw = (double *)malloc(sizeof(double) * np);
    for (int a = 0; a < np; ++a) {
        double p = 1.0;
        for (int c = 0; c < np; ++c)
            if (c != a) p *= z[a] - z[c];
        bw[a] = 1.0 / p;
    }
    for (int b = 0; b < np; ++b) {
        double diag = 0.0;
        for (int a = 0; a < np; ++a) {
            if (a == b) continue;
            const double v = (bw[a] / bw[b]) / (z[b] - z[a]);
            D[a * np + b] = v;
            diag -= v;
        }
        D[b * np + b] = diag;
    }
    free(bw);
}

void or_modified_b_tables(int nm, int nq, const double *z, double *bdata)
{
    /* p = 0: (1-z)/2, (1+z)/2, then P_{q-2}^{(1,1)} (1-z)/2 (1+z)/2 */
    for (int i = 0; i < nq; ++i) {
        bdata[i] = 0.5 * (1.0 - z[i]);
        bdata[nq + i] = 0.5 * (1.0 + z[i]);
    }
    double *mode = bdata + 2 * nq;
    for (int q = 2; q < nm; ++q, mode += nq) {
        or_jacobfd(nq, z, mode, NULL, q - 2, 1.0, 1.0);
        for (int i = 0; i < nq; ++i) mode[i] *= bdata[i] * bdata[nq + i];
    }
    /* p = 1: (1-z)/2, then P_{q-1}^{(1,1)} (1-z)/2 (1+z)/2 */
    for (int i = 0; i < nq; ++i) mode[i] = 0.5 * (1.0 - z[i]);
    mode += nq;
    for (int q = 2; q < nm; ++q, mode += nq) {
        or_jacobfd(nq, z, mode, NULL, q - 2, 1.0, 1.0);
        for (int i = 0; i < nq; ++i) mode[i] *= bdata[i] * bdata[nq + i];
    }
    /* p >= 2: ((1-z)/2)^p, then P_{q-1}^{(2p-1,1)} ((1-z)/2)^p (1+z)/2 */
    const double *one_m_z_pow = bdata, *one_p_z = bdata + nq;
    for (int p = 2; p < nm; ++p) {
        for (int i = 0; i < nq; ++i) mode[i] = bdata[i] * one_m_z_pow[i];
        one_m_z_pow = mode;
        mode += nq;
        for (int q = 1; q < nm - p; ++q, mode += nq) {
            or_jacobfd(nq, z, mode, NULL, q - 1, 2.0 * p - 1.0, 1.0);
            for (int i = 0; i < nq; ++i) mode[i] *= one_m_z_pow[i] * one_p_z[i];
        }
    }
}

void or_basis_modified_b(int nm, int nq, double *z, double *w, double *D, double *bdata,
                         double *dbdata)
{
    or_zwgrjm(z, w, nq, 1.0, 0.0);
    or_Dnodes(D, z, nq);
    or_modified_b_tables(nm, nq, z, bdata);
    const int M = nm * (nm + 1) / 2;
    for (int m = 0; m < M; ++m)
        for (int i = 0; i < nq; ++i) {
            double s = 0.0;
            for (int k = 0; k < nq; ++k) s += D[k * nq + i] * bdata[m * nq + k];
            dbdata[m * nq + i] = s;
        }
}

/* x(xi) = -(xi0 + xi1)/2 v0 + (1 + xi0)/2 v1 + (1 + xi1)/2 v2: a constant map derivative
 * deriv[a][c] = dx_c / dxi_a */
static void tri_map_deriv(const double *v, double *deriv)
{
    for (int c = 0; c < 2; ++c) {
        deriv[0 * 2 + c] = 0.5 * (v[2 + c] - v[c]);
        deriv[1 * 2 + c] = 0.5 * (v[4 + c] - v[c]);
    }
}

void or_geom_tri(const double *verts, int E, int nq0, int nq1, int deformed, double *jac,
                 double *df)
{
    const int Q = nq0 * nq1;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
    for (int e = 0; e < E; ++e) {
        double deriv[4], J, fac[4];
        tri_map_deriv(verts + (size_t)e * 6, deriv);
        or_deriv_factors(2, deriv, &J, fac);
        const int nqq = deformed ? Q : 1;
        for (int q = 0; q < nqq; ++q) {
            const size_t idx = deformed ? (size_t)e * Q + q : (size_t)e;
            jac[idx] = J;
            for (int d = 0; d < 4; ++d) df[d * np + idx] = fac[d];
        }
    }
}

void or_coords_tri(const double *verts, int E, int nq0, int nq1, const double *z0,
                   const double *z1, double *xy)
{
    const int Q = nq0 * nq1;
    for (int e = 0; e < E; ++e) {
        const double *v = verts + (size_t)e * 6;
        for (int j = 0; j < nq1; ++j)
            for (int i = 0; i < nq0; ++i) {
                const double xi0 = 0.5 * (1.0 + z0[i]) * (1.0 - z1[j]) - 1.0, xi1 = z1[j];
                for (int c = 0; c < 2; ++c)
                    xy[c * (size_t)E * Q + (size_t)e * Q + j * nq0 + i] =
                        -0.5 * (xi0 + xi1) * v[c] + 0.5 * (1.0 + xi0) * v[2 + c] +
                        0.5 * (1.0 + xi1) * v[4 + c];
            }
    }
}

/* BwdTransTriKernel (CORRECT: the collapsed vertex mode 1 also carries phi_1(eta0)) */
static void bwd_tri_elmt(int nm, int nq0, int nq1, const double *b0, const double *b1,
                         const double *in, double *psum, double *out)
{
    for (int j = 0, idx = 0; j < nq1; ++j) {
        for (int p = 0, mode = 0; p < nm; ++p) {
            double s = 0.0;
            for (int q = 0; q < nm - p; ++q, ++mode) s = fma(b1[mode * nq1 + j], in[mode], s);
            psum[p] = s;
        }
        for (int i = 0; i < nq0; ++i, ++idx) {
            double s = 0.0;
            for (int p = 0; p < nm; ++p) s = fma(psum[p], b0[p * nq0 + i], s);
            s = fma(in[1] * b0[nq0 + i], b1[nq1 + j], s);
            out[idx] = s;
        }
    }
}

static void store(double *o, double v, int scale, int append, double sc)
{
    if (scale && append) *o = fma(v, sc, *o);
    else if (append) *o = *o + v;
    else if (scale) *o = v * sc;
    else *o = v;
}

/* IProductTriKernel with the CORRECT term on mode 1 (always appended) */
static void iprod_tri_elmt(int nm, int nq0, int nq1, const double *b0, const double *b1,
                           const double *w0, const double *w1, int deformed,
                           const double *jac, const double *in, double *e0sum, double *out,
                           int scale, int append, double sc)
{
    int mode = 0;
    for (int p = 0; p < nm; ++p) {
        for (int j = 0, idx = 0; j < nq1; ++j) {
            double s = 0.0;
            for (int i = 0; i < nq0; ++i, ++idx) {
                const double jv = deformed ? jac[j * nq0 + i] : jac[0];
                const double prod = in[idx] * b0[p * nq0 + i] * jv;
                s = fma(prod, w0[i], s);
            }
            e0sum[j] = s;
        }
        for (int q = 0; q < nm - p; ++q, ++mode) {
            double s = 0.0;
            for (int j = 0; j < nq1; ++j) s = fma(e0sum[j] * b1[mode * nq1 + j], w1[j], s);
            store(&out[mode], s, scale, append, sc);
        }
    }
    double c01 = 0.0;
    for (int j = 0, idx = 0; j < nq1; ++j) {
        const double pre = deformed ? w1[j] * b1[nq1 + j] : w1[j] * jac[0] * b1[nq1 + j];
        for (int i = 0; i < nq0; ++i, ++idx) {
            double prod = in[idx] * pre * w0[i];
            if (deformed) prod = prod * jac[j * nq0 + i];
            c01 = fma(prod, b0[nq0 + i], c01);
        }
    }
    store(&out[1], c01, scale, 1, sc);
}

/* PhysDerivTensor2DKernel: d0 = D0 in along eta0, d1 = in D1^T along eta1 */
static void deriv_tri_elmt(int nq0, int nq1, const double *D0, const double *D1,
                           const double *in, double *d0, double *d1)
{
    for (int i = 0; i < nq0; ++i)
        for (int j = 0; j < nq1; ++j) {
            double s = 0.0;
            for (int k = 0; k < nq0; ++k) s = fma(D0[k * nq0 + i], in[j * nq0 + k], s);
            d0[j * nq0 + i] = s;
        }
    for (int i = 0; i < nq0; ++i)
        for (int j = 0; j < nq1; ++j) {
            double s = 0.0;
            for (int k = 0; k < nq1; ++k) s = fma(in[k * nq0 + i], D1[k * nq1 + j], s);
            d1[j * nq0 + i] = s;
        }
}

void or_bwdtrans_tri(int nm, int nq0, int nq1, const double *b0, const double *b1, int E,
                     const double *in, double *out)
{
    const int M = nm * (nm + 1) / 2, Q = nq0 * nq1;
#pragma omp parallel
    {
        double *psum = (double *)malloc(sizeof(double) * nm);
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e)
            bwd_tri_elmt(nm, nq0, nq1, b0, b1, in + (size_t)e * M, psum, out + (size_t)e * Q);
        free(psum);
    }
}

void or_iproduct_tri(int nm, int nq0, int nq1, const double *b0, const double *b1,
                     const double *w0, const double *w1, int E, int deformed,
                     const double *jac, const double *in, double *out)
{
    const int M = nm * (nm + 1) / 2, Q = nq0 * nq1;
#pragma omp parallel
    {
        double *e0 = (double *)malloc(sizeof(double) * nq1);
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e)
            iprod_tri_elmt(nm, nq0, nq1, b0, b1, w0, w1, deformed,
                           jac + (deformed ? (size_t)e * Q : (size_t)e), in + (size_t)e * Q,
                           e0, out + (size_t)e * M, 0, 0, 1.0);
        free(e0);
    }
}

/* PhysDerivTriKernel: collapsed to standard coordinates, then the derivative factors */
void or_physderiv_tri(int nq0, int nq1, const double *D0, const double *D1,
                      const double *z0, const double *z1, int E, int deformed,
                      const double *df, const double *in, double *out0, double *out1)
{
    const int Q = nq0 * nq1;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        double *d0 = out0 + (size_t)e * Q, *d1 = out1 + (size_t)e * Q;
        deriv_tri_elmt(nq0, nq1, D0, D1, in + (size_t)e * Q, d0, d1);
        double f0 = 0, f1 = 0, f2 = 0, f3 = 0;
        if (!deformed) {
            f0 = df[0 * np + e]; f1 = df[1 * np + e]; f2 = df[2 * np + e]; f3 = df[3 * np + e];
        }
        for (int j = 0, c = 0; j < nq1; ++j) {
            const double x0 = 2.0 / (1.0 - z1[j]);
            for (int i = 0; i < nq0; ++i, ++c) {
                if (deformed) {
                    const size_t k = (size_t)e * Q + c;
                    f0 = df[0 * np + k]; f1 = df[1 * np + k]; f2 = df[2 * np + k]; f3 = df[3 * np + k];
                }
                const double a0 = x0 * d0[c];
                const double a1 = fma(a0, 0.5 * (1.0 + z0[i]), d1[c]);
                d0[c] = fma(a1, f1, a0 * f0);
                d1[c] = fma(a1, f3, a0 * f2);
            }
        }
    }
}

void or_helmholtz_tri(int nm, int nq0, int nq1, const double *b0, const double *db0,
                      const double *b1, const double *db1, const double *D0,
                      const double *D1, const double *z0, const double *z1,
                      const double *w0, const double *w1, int E, int deformed,
                      const double *jac, const double *df, double lambda,
                      const double *in, double *out)
{
    const int M = nm * (nm + 1) / 2, Q = nq0 * nq1;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
#pragma omp parallel
    {
        double *bwd = (double *)malloc(sizeof(double) * 3 * Q);
        double *dv0 = bwd + Q, *dv1 = dv0 + Q;
        double *wsp = (double *)malloc(sizeof(double) * (nm > nq1 ? nm : nq1));
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e) {
            const double *jp = jac + (deformed ? (size_t)e * Q : (size_t)e);
            double *o = out + (size_t)e * M;
            double f0 = 0, f1 = 0, f2 = 0, f3 = 0;
            if (!deformed) {
                f0 = df[0 * np + e]; f1 = df[1 * np + e]; f2 = df[2 * np + e]; f3 = df[3 * np + e];
            }
            bwd_tri_elmt(nm, nq0, nq1, b0, b1, in + (size_t)e * M, wsp, bwd);
            iprod_tri_elmt(nm, nq0, nq1, b0, b1, w0, w1, deformed, jp, bwd, wsp, o, 1, 0, lambda);
            deriv_tri_elmt(nq0, nq1, D0, D1, bwd, dv0, dv1);
            for (int j = 0, c = 0; j < nq1; ++j) {
                const double h1 = 2.0 / (1.0 - z1[j]);
                for (int i = 0; i < nq0; ++i, ++c) {
                    if (deformed) {
                        const size_t k = (size_t)e * Q + c;
                        f0 = df[0 * np + k]; f1 = df[1 * np + k]; f2 = df[2 * np + k]; f3 = df[3 * np + k];
                    }
                    const double h0 = 0.5 * (1.0 + z0[i]);
                    double m00 = h1 * (f0 + h0 * f1);
                    double m01 = m00 * f1;
                    m00 = m00 * m00;
                    const double t = h1 * (f2 + h0 * f3);
                    m01 = fma(t, f3, m01);
                    m00 = fma(t, t, m00);
                    const double m11 = fma(f3, f3, f1 * f1);
                    const double a0 = dv0[c], a1 = dv1[c];
                    bwd[c] = fma(m01, a1, m00 * a0);
                    dv0[c] = fma(m11, a1, m01 * a0);
                }
            }
            iprod_tri_elmt(nm, nq0, nq1, db0, b1, w0, w1, deformed, jp, bwd, wsp, o, 0, 1, 1.0);
            iprod_tri_elmt(nm, nq0, nq1, b0, db1, w0, w1, deformed, jp, dv0, wsp, o, 0, 1, 1.0);
        }
        free(wsp);
        free(bwd);
    }
}

/* IProductWRTDerivBaseTriImpl (MatrixFreeOps/IProductWRTDerivBase.h:671-, 2D coordinates):
 * t_a = sum_c df[2c+a] in_c per point, then the collapsed chain rule
 * tmp0 = 2/(1-eta1) (t0 + (1+eta0)/2 t1), tmp1 = t1, and the inner products with
 * (dB0, B1) and (B0, dB1) */
void or_iprodderiv_tri(int nm, int nq0, int nq1, const double *b0, const double *db0,
                       const double *b1, const double *db1, const double *z0, const double *z1,
                       const double *w0, const double *w1, int E, int deformed,
                       const double *jac, const double *df, const double *in0,
                       const double *in1, double *out)
{
    const int M = nm * (nm + 1) / 2, Q = nq0 * nq1;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
    double *tmp0 = (double *)malloc(sizeof(double) * 2 * Q), *tmp1 = tmp0 + Q;
    double *wsp = (double *)malloc(sizeof(double) * (nm > nq1 ? nm : nq1));
    for (int e = 0; e < E; ++e) {
        const double *jp = jac + (deformed ? (size_t)e * Q : (size_t)e);
        double f[4] = {0, 0, 0, 0};
        if (!deformed)
            for (int a = 0; a < 4; ++a) f[a] = df[a * np + e];
        for (int j = 0, c = 0; j < nq1; ++j) {
            const double f0 = 2.0 / (1.0 - z1[j]);
            for (int i = 0; i < nq0; ++i, ++c) {
                if (deformed)
                    for (int a = 0; a < 4; ++a) f[a] = df[a * np + (size_t)e * Q + c];
                const double x0 = in0[(size_t)e * Q + c], x1 = in1[(size_t)e * Q + c];
                double t0 = f[0] * x0 + f[2] * x1;
                const double t1 = f[1] * x0 + f[3] * x1;
                const double hf1 = 0.5 * (1.0 + z0[i]);
                t0 = t0 * f0;
                t0 = fma(hf1 * t1, f0, t0);
                tmp0[c] = t0;
                tmp1[c] = t1;
            }
        }
        double *o = out + (size_t)e * M;
        iprod_tri_elmt(nm, nq0, nq1, db0, b1, w0, w1, deformed, jp, tmp0, wsp, o, 0, 0, 1.0);
        iprod_tri_elmt(nm, nq0, nq1, b0, db1, w0, w1, deformed, jp, tmp1, wsp, o, 0, 1, 1.0);
    }
    free(tmp0);
    free(wsp);
}
