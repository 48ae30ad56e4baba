/*
 * CPU oracle, pyramids (TEST INFRASTRUCTURE; see mf_oracle.h).  Plain-C restatement of the
 * reference's collapsed-coordinate pyramid path:
 *   ModifiedPyr_C basis        LibUtilities/Foundations/Basis.cpp:569-668
 *   Gauss-Radau alpha = 2      Polylib.cpp:119-144 (zwgrjm), GaussPoints.cpp:123-125
 *   straight-sided PyrGeom     the vertex modes of the pyramid expansion, GeomFactors.cpp
 *                              (derivatives of the map through the collapsed chain rule)
 *   kernels                    MatrixFreeOps/BwdTransKernels.hpp:131-222 (BwdTransPyr),
 *                              IProductKernels.hpp:455-598 (IProductPyr),
 *                              PhysDerivKernels.hpp:464-550 (PhysDerivPyr),
 *                              Helmholtz.h:1478-1510, :1771-1935 (HelmholtzPyrImpl)
 * with the loop and summation order of those kernels (element interleave width 1).
 *
 * Expansion: Modified_A along eta0 and eta1 (nq0, nq1 GLL), ModifiedPyr_C along eta2 (nq2
 * Gauss-Radau alpha = 2), nm modes per direction.  Coefficients u[e*M + mode], mode
 * (p, q, r) with r < nm - max(p, q), r fastest, then q, then p;
 * M = nm (nm + 1) (2 nm + 1) / 6; b2 has one row per mode.  Points
 * f[e*Q + (k*nq1 + j)*nq0 + i].  w2 is the Helper's 0.25-scaled weight.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mf_oracle.h"

static int imax(int a, int b) { return a > b ? a : b; }

void or_basis_modified_pyr_c(int nm, int nq, double *z, double *w, double *D, double *bdata,
                             double *dbdata)
{
    or_zwgrjm(z, w, nq, 2.0, 0.0);
    or_Dnodes(D, z, nq);
    const int T = nm * (nm + 1) / 2;
    double *bb = (double *)malloc(sizeof(double) * (size_t)T * nq);
    or_modified_b_tables(nm, nq, z, bb);
    /* start of Modified_B group s: sum_{s'<s} (nm - s') rows */
    double *mode = bdata;
    for (int p = 0; p < nm; ++p)
        for (int q = 0; q < nm; ++q) {
            const int nr = nm - imax(p, q);
            if (p <= 1 || q <= 1) {  /* phi^b_{max(p,q), r} */
                const int s = imax(p, q);
                const int start = s * nm - s * (s - 1) / 2;
                memcpy(mode, bb + (size_t)start * nq, sizeof(double) * nr * nq);
                mode += (size_t)nr * nq;
            } else {  /* ((1-z)/2)^{p+q-2}, then times (1+z)/2 P_{r-1}^{(2p+2q-3,1)} */
                const double *pw = mode;
                for (int i = 0; i < nq; ++i) mode[i] = pow(0.5 * (1.0 - z[i]), p + q - 2);
                mode += nq;
                for (int r = 1; r < nr; ++r, mode += nq) {
                    or_jacobfd(nq, z, mode, NULL, r - 1, 2.0 * p + 2.0 * q - 3.0, 1.0);
                    for (int i = 0; i < nq; ++i) mode[i] *= pw[i] * 0.5 * (1.0 + z[i]);
                }
            }
        }
    free(bb);
    const int M = nm * (nm + 1) * (2 * nm + 1) / 6;
    for (int m = 0; m < M; ++m)
        for (int i = 0; i < nq; ++i) {
            double s = 0.0;
            for (int k = 0; k < nq; ++k) s += D[k * nq + i] * bdata[m * nq + k];
            dbdata[m * nq + i] = s;
        }
}

/* the map at collapsed point eta: x = sum_v phi_v(eta) x_v with the vertex modes
 * a0 b0 c0, a1 b0 c0, a1 b1 c0, a0 b1 c0, c1 (a = (1 -/+ eta0)/2, b the same in eta1,
 * c in eta2); deriv[a*3 + c] = dx_c / dxi_a through the collapsed chain rule */
static void pyr_map(const double *v, double e0, double e1, double e2, double *deriv, double *x)
{
    const double a0 = 0.5 * (1 - e0), a1 = 0.5 * (1 + e0), b0 = 0.5 * (1 - e1),
                 b1 = 0.5 * (1 + e1), c0 = 0.5 * (1 - e2), c1 = 0.5 * (1 + e2);
    const double N[5] = {a0 * b0 * c0, a1 * b0 * c0, a1 * b1 * c0, a0 * b1 * c0, c1};
    const double dN[5][3] = {{-0.5 * b0 * c0, -0.5 * a0 * c0, -0.5 * a0 * b0},
                             {0.5 * b0 * c0, -0.5 * a1 * c0, -0.5 * a1 * b0},
                             {0.5 * b1 * c0, 0.5 * a1 * c0, -0.5 * a1 * b1},
                             {-0.5 * b1 * c0, 0.5 * a0 * c0, -0.5 * a0 * b1},
                             {0.0, 0.0, 0.5}};
    const double h2 = 2.0 / (1.0 - e2), h0 = 0.5 * (1 + e0), h1 = 0.5 * (1 + e1);
    for (int c = 0; c < 3; ++c) {
        double de[3] = {0, 0, 0}, s = 0;
        for (int n = 0; n < 5; ++n) {
            for (int a = 0; a < 3; ++a) de[a] += dN[n][a] * v[n * 3 + c];
            s += N[n] * v[n * 3 + c];
        }
        if (x) x[c] = s;
        if (deriv) {
            deriv[0 * 3 + c] = h2 * de[0];
            deriv[1 * 3 + c] = h2 * de[1];
            deriv[2 * 3 + c] = h0 * h2 * de[0] + h1 * h2 * de[1] + de[2];
        }
    }
}

void or_geom_pyr(const double *verts, int E, int nq0, int nq1, int nq2, const double *z0,
                 const double *z1, const double *z2, int deformed, double *jac, double *df)
{
    const int Q = nq0 * nq1 * nq2;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
    for (int e = 0; e < E; ++e) {
        const int nqq = deformed ? Q : 1;
        for (int q = 0; q < nqq; ++q) {
            const int i = q % nq0, j = (q / nq0) % nq1, k = q / (nq0 * nq1);
            double deriv[9], J, fac[9];
            pyr_map(verts + (size_t)e * 15, z0[i], z1[j], z2[k], deriv, NULL);
            or_deriv_factors(3, deriv, &J, fac);
            const size_t idx = deformed ? (size_t)e * Q + q : (size_t)e;
            jac[idx] = J;
            for (int d = 0; d < 9; ++d) df[d * np + idx] = fac[d];
        }
    }
}

void or_coords_pyr(const double *verts, int E, int nq0, int nq1, int nq2, const double *z0,
                   const double *z1, const double *z2, double *xyz)
{
    const int Q = nq0 * nq1 * nq2;
    for (int e = 0; e < E; ++e)
        for (int q = 0; q < Q; ++q) {
            const int i = q % nq0, j = (q / nq0) % nq1, k = q / (nq0 * nq1);
            double x[3];
            pyr_map(verts + (size_t)e * 15, z0[i], z1[j], z2[k], NULL, x);
            for (int c = 0; c < 3; ++c) xyz[(size_t)c * E * Q + (size_t)e * Q + q] = x[c];
        }
}

/* BwdTransPyrKernel with the top-vertex CORRECT term */
static void bwd_pyr_elmt(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                         const double *b2, const double *in, double *fpq, double *fp,
                         double *out)
{
    for (int k = 0, cnt = 0; k < nq2; ++k) {
        int mode_pqr = 0, mode_pq = 0;
        for (int p = 0; p < nm; ++p) {
            for (int q = 0; q < p; ++q, ++mode_pq) {
                double prod = 0.0;
                for (int r = 0; r < nm - p; ++r, ++mode_pqr)
                    prod = fma(in[mode_pqr], b2[mode_pqr * nq2 + k], prod);
                fpq[mode_pq] = prod;
            }
            for (int q = p; q < nm; ++q, ++mode_pq) {
                double prod = 0.0;
                for (int r = 0; r < nm - q; ++r, ++mode_pqr)
                    prod = fma(in[mode_pqr], b2[mode_pqr * nq2 + k], prod);
                fpq[mode_pq] = prod;
            }
        }
        for (int j = 0; j < nq1; ++j) {
            mode_pq = 0;
            for (int p = 0; p < nm; ++p) {
                double prod = 0.0;
                for (int q = 0; q < nm; ++q, ++mode_pq) prod = fma(fpq[mode_pq], b1[q * nq1 + j], prod);
                fp[p] = prod;
            }
            for (int i = 0; i < nq0; ++i, ++cnt) {
                double v = 0.0;
                for (int p = 0; p < nm; ++p) v = fma(fp[p], b0[p * nq0 + i], v);
                double t1 = b0[i] * b1[nq1 + j];
                t1 = fma(b0[nq0 + i], b1[j], t1);
                t1 = fma(b0[nq0 + i], b1[nq1 + j], t1);
                t1 = t1 * b2[nq2 + k];
                v = fma(t1, in[1], v);
                out[cnt] = v;
            }
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

/* IProductPyrKernel with the top-vertex CORRECT term */
static void iprod_pyr_elmt(int nm, int nq0, int nq1, int nq2, const double *b0,
                           const double *b1, const double *b2, const double *w0,
                           const double *w1, const double *w2, int deformed, const double *jac,
                           const double *in, double *skj, double *sk, double *out, int scale,
                           int append, double sc)
{
    int mode_pqr = 0;
    for (int p = 0; p < nm; ++p) {
        int cnt_kji = 0, cnt_kj = 0;
        for (int k = 0; k < nq2; ++k)
            for (int j = 0; j < nq1; ++j, ++cnt_kj) {
                double s = 0.0;
                for (int i = 0; i < nq0; ++i, ++cnt_kji) {
                    const double jv = deformed ? jac[nq0 * nq1 * k + nq0 * j + i] : jac[0];
                    const double prod = b0[nq0 * p + i] * jv * w0[i];
                    s = fma(prod, in[cnt_kji], s);
                }
                skj[cnt_kj] = s;
            }
        for (int q = 0; q < nm; ++q) {
            cnt_kj = 0;
            for (int k = 0; k < nq2; ++k) {
                double s = 0.0;
                for (int j = 0; j < nq1; ++j, ++cnt_kj) s = fma(b1[q * nq1 + j] * w1[j], skj[cnt_kj], s);
                sk[k] = s;
            }
            for (int r = 0; r < nm - imax(p, q); ++r, ++mode_pqr) {
                double s = 0.0;
                for (int k = 0; k < nq2; ++k) s = fma(b2[mode_pqr * nq2 + k] * w2[k], sk[k], s);
                store(&out[mode_pqr], s, scale, append, sc);
            }
        }
    }
    for (int k = 0, cnt = 0; k < nq2; ++k) {
        double q2 = w2[k];
        if (!deformed) q2 = q2 * jac[0];
        for (int j = 0; j < nq1; ++j) {
            const double q1 = q2 * w1[j];
            for (int i = 0; i < nq0; ++i, ++cnt) {
                double tq = q1 * w0[i];
                const double tin = in[cnt];
                if (deformed) tq = tq * jac[cnt];
                double t = b0[i] * b1[nq1 + j];
                t = fma(b0[nq0 + i], b1[j], t);
                t = fma(b0[nq0 + i], b1[nq1 + j], t);
                t = t * b2[nq2 + k];
                t = t * tin;
                store(&out[1], t * tq, scale, 1, sc);
            }
        }
    }
}

void or_bwdtrans_pyr(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                     const double *b2, int E, const double *in, double *out)
{
    const int M = nm * (nm + 1) * (2 * nm + 1) / 6, Q = nq0 * nq1 * nq2;
    double *fpq = (double *)malloc(sizeof(double) * nm * nm);
    double *fp = (double *)malloc(sizeof(double) * nm);
    for (int e = 0; e < E; ++e)
        bwd_pyr_elmt(nm, nq0, nq1, nq2, b0, b1, b2, in + (size_t)e * M, fpq, fp,
                     out + (size_t)e * Q);
    free(fpq);
    free(fp);
}

void or_iproduct_pyr(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                     const double *b2, const double *w0, const double *w1, const double *w2,
                     int E, int deformed, const double *jac, const double *in, double *out)
{
    const int M = nm * (nm + 1) * (2 * nm + 1) / 6, Q = nq0 * nq1 * nq2;
    double *ws = (double *)malloc(sizeof(double) * (nq1 * nq2 + nq2));
    for (int e = 0; e < E; ++e) {
        const double *jp = deformed ? jac + (size_t)e * Q : jac + e;
        iprod_pyr_elmt(nm, nq0, nq1, nq2, b0, b1, b2, w0, w1, w2, deformed, jp,
                       in + (size_t)e * Q, ws, ws + nq1 * nq2, out + (size_t)e * M, 0, 0, 1.0);
    }
    free(ws);
}

/* PhysDerivPyrKernel: d/dxi0 = h2 d/deta0, d/dxi1 = h2 d/deta1,
 * d/dxi2 = d/deta2 + h0 d/dxi0 + h1 d/dxi1 (h0 = (1+eta0)/2, h1 = (1+eta1)/2,
 * h2 = 2/(1-eta2)); out_c = sum_a df[3c+a] d/dxi_a */
void or_physderiv_pyr(int nq0, int nq1, int nq2, const double *D0, const double *D1,
                      const double *D2, const double *z0, const double *z1, const double *z2,
                      int E, int deformed, const double *df, const double *in, double *out0,
                      double *out1, double *out2)
{
    const int Q = nq0 * nq1 * nq2;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
    for (int e = 0; e < E; ++e) {
        double *o0 = out0 + (size_t)e * Q, *o1 = out1 + (size_t)e * Q, *o2 = out2 + (size_t)e * Q;
        or_deriv3_elmt(nq0, nq1, nq2, D0, D1, D2, in + (size_t)e * Q, o0, o1, o2);
        double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (!deformed)
            for (int a = 0; a < 9; ++a) f[a] = df[a * np + e];
        for (int k = 0, c = 0; k < nq2; ++k) {
            const double x2 = 2.0 / (1.0 - z2[k]);
            for (int j = 0; j < nq1; ++j) {
                const double x1 = 0.5 * (1 + z1[j]);
                for (int i = 0; i < nq0; ++i, ++c) {
                    const double d0 = o0[c] * x2, d1 = o1[c] * x2;
                    double d2 = o2[c];
                    const double x0 = 0.5 * (1 + z0[i]);
                    d2 = fma(x0, d0, d2);
                    d2 = fma(x1, d1, d2);
                    if (deformed)
                        for (int a = 0; a < 9; ++a) f[a] = df[a * np + (size_t)e * Q + c];
                    double o;
                    o = d0 * f[0]; o = fma(d1, f[1], o); o = fma(d2, f[2], o); o0[c] = o;
                    o = d0 * f[3]; o = fma(d1, f[4], o); o = fma(d2, f[5], o); o1[c] = o;
                    o = d0 * f[6]; o = fma(d1, f[7], o); o = fma(d2, f[8], o); o2[c] = o;
                }
            }
        }
    }
}

/* HelmholtzPyrImpl: lambda * mass + the weak Laplacian with the metric formed per point
 * (Helmholtz.h:1478-1510 for h0, h1, h2) */
void or_helmholtz_pyr(int nm, int nq0, int nq1, int nq2, const double *b0, const double *db0,
                      const double *b1, const double *db1, const double *b2, const double *db2,
                      const double *D0, const double *D1, const double *D2, const double *z0,
                      const double *z1, const double *z2, const double *w0, const double *w1,
                      const double *w2, int E, int deformed, const double *jac,
                      const double *df, double lambda, const double *in, double *out)
{
    const int M = nm * (nm + 1) * (2 * nm + 1) / 6, Q = nq0 * nq1 * nq2;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
    double *bwd = (double *)malloc(sizeof(double) * Q), *dv0 = (double *)malloc(sizeof(double) * Q),
           *dv1 = (double *)malloc(sizeof(double) * Q), *dv2 = (double *)malloc(sizeof(double) * Q);
    double *ws = (double *)malloc(sizeof(double) * (nq1 * nq2 + nq2 + nm * nm + nm));
    for (int e = 0; e < E; ++e) {
        const double *jp = deformed ? jac + (size_t)e * Q : jac + e;
        double *o = out + (size_t)e * M;
        double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (!deformed)
            for (int a = 0; a < 9; ++a) f[a] = df[a * np + e];
        bwd_pyr_elmt(nm, nq0, nq1, nq2, b0, b1, b2, in + (size_t)e * M, ws, ws + nm * nm, bwd);
        iprod_pyr_elmt(nm, nq0, nq1, nq2, b0, b1, b2, w0, w1, w2, deformed, jp, bwd, ws,
                       ws + nq1 * nq2, o, 1, 0, lambda);
        or_deriv3_elmt(nq0, nq1, nq2, D0, D1, D2, bwd, dv0, dv1, dv2);
        for (int k = 0, c = 0; k < nq2; ++k) {
            const double h2 = 2.0 / (1.0 - z2[k]);
            for (int j = 0; j < nq1; ++j) {
                const double h1 = 0.5 * (1 + z1[j]), h1h2 = h1 * h2;
                for (int i = 0; i < nq0; ++i, ++c) {
                    const double h0h2 = 0.5 * (1 + z0[i]) * h2;
                    if (deformed)
                        for (int a = 0; a < 9; ++a) f[a] = df[a * np + (size_t)e * Q + c];
                    const double t0 = fma(h0h2, f[2], h2 * f[0]);
                    const double t1 = fma(h0h2, f[5], h2 * f[3]);
                    const double t2 = fma(h0h2, f[8], h2 * f[6]);
                    const double t3 = fma(h1h2, f[2], h2 * f[1]);
                    const double t4 = fma(h1h2, f[5], h2 * f[4]);
                    const double t5 = fma(h1h2, f[8], h2 * f[7]);
                    double g0 = t0 * t0; g0 = fma(t1, t1, g0); g0 = fma(t2, t2, g0);
                    double g1 = t3 * t3; g1 = fma(t4, t4, g1); g1 = fma(t5, t5, g1);
                    double g2 = f[2] * f[2]; g2 = fma(f[5], f[5], g2); g2 = fma(f[8], f[8], g2);
                    double g3 = t0 * t3; g3 = fma(t1, t4, g3); g3 = fma(t2, t5, g3);
                    double g4 = f[2] * t0; g4 = fma(f[5], t1, g4); g4 = fma(f[8], t2, g4);
                    double g5 = f[2] * t3; g5 = fma(f[5], t4, g5); g5 = fma(f[8], t5, g5);
                    const double a0 = dv0[c], a1 = dv1[c], a2 = dv2[c];
                    double t;
                    t = g0 * a0; t = fma(g3, a1, t); t = fma(g4, a2, t); dv0[c] = t;
                    t = g3 * a0; t = fma(g1, a1, t); t = fma(g5, a2, t); dv1[c] = t;
                    t = g4 * a0; t = fma(g5, a1, t); t = fma(g2, a2, t); dv2[c] = t;
                }
            }
        }
        double *s1 = ws, *s2 = ws + nq1 * nq2;
        iprod_pyr_elmt(nm, nq0, nq1, nq2, db0, b1, b2, w0, w1, w2, deformed, jp, dv0, s1, s2, o, 0, 1, 1.0);
        iprod_pyr_elmt(nm, nq0, nq1, nq2, b0, db1, b2, w0, w1, w2, deformed, jp, dv1, s1, s2, o, 0, 1, 1.0);
        iprod_pyr_elmt(nm, nq0, nq1, nq2, b0, b1, db2, w0, w1, w2, deformed, jp, dv2, s1, s2, o, 0, 1, 1.0);
    }
    free(ws);
    free(bwd);
    free(dv0);
    free(dv1);
    free(dv2);
}

/* IProductWRTDerivBasePyrImpl (MatrixFreeOps/IProductWRTDerivBase.h:2056-): t_a =
 * sum_c df[3c+a] in_c, tmp0 = 2/(1-eta2) (t0 + (1+eta0)/2 t2),
 * tmp1 = 2/(1-eta2) (t1 + (1+eta1)/2 t2), tmp2 = t2 */
void or_iprodderiv_pyr(int nm, int nq0, int nq1, int nq2, const double *b0, const double *db0,
                       const double *b1, const double *db1, const double *b2, const double *db2,
                       const double *z0, const double *z1, const double *z2, const double *w0,
                       const double *w1, const double *w2, int E, int deformed,
                       const double *jac, const double *df, const double *in0,
                       const double *in1, const double *in2, double *out)
{
    const int M = nm * (nm + 1) * (2 * nm + 1) / 6, Q = nq0 * nq1 * nq2;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
    double *tmp = (double *)malloc(sizeof(double) * 3 * Q);
    double *ws = (double *)malloc(sizeof(double) * (nq1 * nq2 + nq2));
    const double *in[3] = {in0, in1, in2};
    for (int e = 0; e < E; ++e) {
        const double *jp = jac + (deformed ? (size_t)e * Q : (size_t)e);
        double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (!deformed)
            for (int a = 0; a < 9; ++a) f[a] = df[a * np + e];
        for (int k = 0, c = 0; k < nq2; ++k) {
            const double f0 = 2.0 / (1.0 - z2[k]);
            for (int j = 0; j < nq1; ++j) {
                const double hf2 = 0.5 * (1.0 + z1[j]);
                for (int i = 0; i < nq0; ++i, ++c) {
                    if (deformed)
                        for (int a = 0; a < 9; ++a) f[a] = df[a * np + (size_t)e * Q + c];
                    double x[3], t[3];
                    for (int d = 0; d < 3; ++d) x[d] = in[d][(size_t)e * Q + c];
                    for (int a = 0; a < 3; ++a) t[a] = f[a] * x[0] + f[3 + a] * x[1] + f[6 + a] * x[2];
                    const double hf1 = 0.5 * (1.0 + z0[i]);
                    t[0] = t[0] * f0;
                    t[0] = fma(hf1 * t[2], f0, t[0]);
                    t[1] = t[1] * f0;
                    t[1] = fma(hf2 * t[2], f0, t[1]);
                    for (int a = 0; a < 3; ++a) tmp[a * Q + c] = t[a];
                }
            }
        }
        double *o = out + (size_t)e * M;
        iprod_pyr_elmt(nm, nq0, nq1, nq2, db0, b1, b2, w0, w1, w2, deformed, jp, tmp, ws, ws + nq1 * nq2, o, 0, 0, 1.0);
        iprod_pyr_elmt(nm, nq0, nq1, nq2, b0, db1, b2, w0, w1, w2, deformed, jp, tmp + Q, ws, ws + nq1 * nq2, o, 0, 1, 1.0);
        iprod_pyr_elmt(nm, nq0, nq1, nq2, b0, b1, db2, w0, w1, w2, deformed, jp, tmp + 2 * Q, ws, ws + nq1 * nq2, o, 0, 1, 1.0);
    }
    free(tmp);
    free(ws);
}
