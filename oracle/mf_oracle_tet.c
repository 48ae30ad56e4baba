/*
 * CPU oracle, tetrahedra (TEST INFRASTRUCTURE; see mf_oracle.h).  Plain-C restatement of
 * the reference's collapsed-coordinate tetrahedron path:
 *   Modified_C basis           LibUtilities/Foundations/Basis.cpp:513-563 (phi^c_{pqr} =
 *                              phi^b_{p+q,r} on the Gauss-Radau alpha = 2 points)
 *   Gauss-Radau alpha = 2      Polylib.cpp:119-144 (zwgrjm), GaussPoints.cpp:123-125
 *   straight-sided TetGeom     affine map of the four vertices, GeomFactors.cpp:420-474
 *   kernels                    MatrixFreeOps/BwdTransKernels.hpp:377-484 (BwdTransTet),
 *                              IProductKernels.hpp:604-760 (IProductTet),
 *                              PhysDerivKernels.hpp:556-694 (PhysDerivTet),
 *                              Helmholtz.h:1970-2005, :2266-2455 (HelmholtzTetImpl)
 * with the loop and summation order of those kernels (element interleave width 1).
 *
 * Expansion: Modified_A along eta0 (nq0 GLL), Modified_B along eta1 (nq1 Gauss-Radau
 * alpha = 1), Modified_C along eta2 (nq2 Gauss-Radau alpha = 2), nm modes per direction.
 * Coefficients u[e*M + mode], mode (p, q, r) with q < nm - p, r < nm - p - q, r fastest,
 * then q, then p; M = nm (nm + 1) (nm + 2) / 6.  b1 has one row per (p, q) (tri packing),
 * b2 one row per mode.  Points f[e*Q + (k*nq1 + j)*nq0 + i].  w1 and w2 are the Helper's
 * scaled weights (0.5 and 0.25 times the Gauss-Radau weights).  The reference's BwdTransTet
 * takes the eta2 point count from direction 1 (its instantiations have nq1 = nq2); here
 * nq2 is used, which is the same for every size the reference instantiates.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mf_oracle.h"

void or_basis_modified_c(int nm, int nq, double *z, double *w, double *D, double *bdata,
                         double *dbdata)
{
    or_zwgrjm(z, w, nq, 2.0, 0.0);
    or_Dnodes(D, z, nq);
    const int T = nm * (nm + 1) / 2;
    double *bb = (double *)malloc(sizeof(double) * (size_t)T * nq);
    or_modified_b_tables(nm, nq, z, bb);
    /* block p of C = the Modified_B groups p, p+1, ..., nm-1 one after the other */
    size_t off = 0, boff = 0;
    for (int p = 0; p < nm; ++p) {
        const size_t n = (size_t)nq * (nm - p) * (nm - p + 1) / 2;
        memcpy(bdata + off, bb + boff, n * sizeof(double));
        boff += (size_t)nq * (nm - p);
        off += n;
    }
    free(bb);
    const int M = nm * (nm + 1) * (nm + 2) / 6;
    for (int m = 0; m < M; ++m)
        for (int i = 0; i < nq; ++i) {
            double s = 0.0;
            for (int k = 0; k < nq; ++k) s += D[k * nq + i] * bdata[m * nq + k];
            dbdata[m * nq + i] = s;
        }
}

/* x(xi) = -(1 + xi0 + xi1 + xi2)/2 v0 + (1 + xi0)/2 v1 + (1 + xi1)/2 v2 + (1 + xi2)/2 v3:
 * constant derivative deriv[a*3 + c] = dx_c / dxi_a */
static void tet_map_deriv(const double *v, double *deriv)
{
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) deriv[a * 3 + c] = 0.5 * (v[(a + 1) * 3 + c] - v[c]);
}

void or_geom_tet(const double *verts, int E, int nq0, int nq1, int nq2, int deformed,
                 double *jac, double *df)
{
    const int Q = nq0 * nq1 * nq2;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
    for (int e = 0; e < E; ++e) {
        double deriv[9], J, fac[9];
        tet_map_deriv(verts + (size_t)e * 12, deriv);
        or_deriv_factors(3, deriv, &J, fac);
        const int nqq = deformed ? Q : 1;
        for (int q = 0; q < nqq; ++q) {
            const size_t idx = deformed ? (size_t)e * Q + q : (size_t)e;
            jac[idx] = J;
            for (int d = 0; d < 9; ++d) df[d * np + idx] = fac[d];
        }
    }
}

void or_coords_tet(const double *verts, int E, int nq0, int nq1, int nq2, const double *z0,
                   const double *z1, const double *z2, double *xyz)
{
    const int Q = nq0 * nq1 * nq2;
    for (int e = 0; e < E; ++e) {
        const double *v = verts + (size_t)e * 12;
        for (int q = 0; q < Q; ++q) {
            const int i = q % nq0, j = (q / nq0) % nq1, k = q / (nq0 * nq1);
            /* collapsed (eta0, eta1, eta2) -> xi */
            const double x2 = z2[k];
            const double x1 = 0.5 * (1.0 + z1[j]) * (1.0 - x2) - 1.0;
            const double x0 = 0.5 * (1.0 + z0[i]) * (-x1 - x2) - 1.0;
            const double N[4] = {-0.5 * (1.0 + x0 + x1 + x2), 0.5 * (1.0 + x0), 0.5 * (1.0 + x1),
                                 0.5 * (1.0 + x2)};
            for (int c = 0; c < 3; ++c) {
                double s = 0.0;
                for (int n = 0; n < 4; ++n) s += N[n] * v[n * 3 + c];
                xyz[(size_t)c * E * Q + (size_t)e * Q + q] = s;
            }
        }
    }
}

/* BwdTransTetKernel with the CORRECT terms: the top vertex (mode 1), the bottom vertex of
 * the collapsed edge (mode nm) and the singular edge (modes nm + r) also carry the
 * phi_1(eta0) / phi_0-phi_1 products of the lower directions */
static void bwd_tet_elmt(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                         const double *b2, const double *in, double *fpq, double *fp,
                         double *out)
{
    for (int k = 0, cnt = 0; k < nq2; ++k) {
        int cnt_pq = 0, mode = 0;
        for (int p = 0; p < nm; ++p)
            for (int q = 0; q < nm - p; ++q, ++cnt_pq) {
                double prod = in[mode] * b2[k + nq2 * mode];
                ++mode;
                for (int r = 1; r < nm - p - q; ++r, ++mode) prod = fma(in[mode], b2[k + nq2 * mode], prod);
                fpq[cnt_pq] = prod;
            }
        for (int j = 0; j < nq1; ++j) {
            mode = cnt_pq = 0;
            for (int p = 0; p < nm; ++p) {
                double prod = fpq[cnt_pq] * b1[mode * nq1 + j];
                ++cnt_pq;
                for (int q = 1; q < nm - p; ++q, ++cnt_pq)
                    prod = fma(fpq[cnt_pq], b1[(mode + q) * nq1 + j], prod);
                fp[p] = prod;
                mode += nm - p;
            }
            for (int i = 0; i < nq0; ++i, ++cnt) {
                double tmp = b0[i] * fp[0];
                for (int p = 1; p < nm; ++p) tmp = fma(b0[p * nq0 + i], fp[p], tmp);
                double t1 = b0[i] * b1[nq1 + j];
                t1 = fma(b0[nq0 + i], b1[j], t1);
                t1 = fma(b0[nq0 + i], b1[nq1 + j], t1);
                t1 = t1 * b2[nq2 + k];
                tmp = fma(t1, in[1], tmp);
                t1 = b0[nq0 + i] * b1[nq1 + j];
                t1 = t1 * b2[k];
                tmp = fma(in[nm], t1, tmp);
                for (int r = 1; r < nm - 1; ++r) {
                    t1 = b1[nq1 + j] * b0[nq0 + i];
                    t1 = t1 * b2[(r + 1) * nq2 + k];
                    tmp = fma(in[nm + r], t1, tmp);
                }
                out[cnt] = tmp;
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

/* IProductTetKernel: the (p, q, r) sums, then the CORRECT contributions to modes 1, nm and
 * nm + r, each stored with the same scale / append rule (the corrections always append) */
static void iprod_tet_elmt(int nm, int nq0, int nq1, int nq2, const double *b0,
                           const double *b1, const double *b2, const double *w0,
                           const double *w1, const double *w2, int deformed, const double *jac,
                           const double *in, double *f, double *fb, double *out, int scale,
                           int append, double sc)
{
    for (int p = 0, mode = 0, mode2 = 0, cnt_pqr = 0; p < nm; ++p) {
        int cnt_kji = 0;
        for (int k = 0, cnt_kj = 0; k < nq2; ++k)
            for (int j = 0; j < nq1; ++j, ++cnt_kj) {
                double jv = deformed ? jac[nq0 * nq1 * k + nq0 * j] : jac[0];
                double f_kj = in[cnt_kji] * b0[nq0 * p] * jv * w0[0];
                ++cnt_kji;
                for (int i = 1; i < nq0; ++i, ++cnt_kji) {
                    jv = deformed ? jac[nq0 * nq1 * k + nq0 * j + i] : jac[0];
                    const double inxmm = in[cnt_kji] * b0[i + nq0 * p] * jv;
                    f_kj = fma(inxmm, w0[i], f_kj);
                }
                f[cnt_kj] = f_kj;
            }
        for (int q = 0; q < nm - p; ++q, ++mode) {
            int cnt_kj = 0;
            for (int k = 0; k < nq2; ++k) {
                double f_k = b1[mode * nq1] * f[cnt_kj] * w1[0];
                ++cnt_kj;
                for (int j = 1; j < nq1; ++j, ++cnt_kj) {
                    const double t2 = b1[mode * nq1 + j] * f[cnt_kj];
                    f_k = fma(t2, w1[j], f_k);
                }
                fb[k] = f_k;
            }
            for (int r = 0; r < nm - p - q; ++r, ++mode2, ++cnt_pqr) {
                double tmp = fb[0] * b2[mode2 * nq2] * w2[0];
                for (int k = 1; k < nq2; ++k) {
                    const double t2 = fb[k] * b2[mode2 * nq2 + k];
                    tmp = fma(t2, w2[k], tmp);
                }
                store(&out[cnt_pqr], tmp, scale, append, sc);
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
                t = b0[nq0 + i] * b1[nq1 + j] * b2[k] * tin;
                store(&out[nm], t * tq, scale, 1, sc);
                for (int r = 1; r < nm - 1; ++r) {
                    t = b2[(r + 1) * nq2 + k] * b1[nq1 + j] * b0[nq0 + i] * tin;
                    store(&out[nm + r], t * tq, scale, 1, sc);
                }
            }
        }
    }
}

void or_bwdtrans_tet(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                     const double *b2, int E, const double *in, double *out)
{
    const int M = nm * (nm + 1) * (nm + 2) / 6, Q = nq0 * nq1 * nq2;
    double *fpq = (double *)malloc(sizeof(double) * nm * nm);
    double *fp = (double *)malloc(sizeof(double) * nm);
    for (int e = 0; e < E; ++e)
        bwd_tet_elmt(nm, nq0, nq1, nq2, b0, b1, b2, in + (size_t)e * M, fpq, fp,
                     out + (size_t)e * Q);
    free(fpq);
    free(fp);
}

void or_iproduct_tet(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                     const double *b2, const double *w0, const double *w1, const double *w2,
                     int E, int deformed, const double *jac, const double *in, double *out)
{
    const int M = nm * (nm + 1) * (nm + 2) / 6, Q = nq0 * nq1 * nq2;
    double *ws = (double *)malloc(sizeof(double) * (nq1 * nq2 + nq2));
    for (int e = 0; e < E; ++e) {
        const double *jp = deformed ? jac + (size_t)e * Q : jac + e;
        iprod_tet_elmt(nm, nq0, nq1, nq2, b0, b1, b2, w0, w1, w2, deformed, jp,
                       in + (size_t)e * Q, ws, ws + nq1 * nq2, out + (size_t)e * M, 0, 0, 1.0);
    }
    free(ws);
}

/* the collapsed chain rule of PhysDerivTetKernel: d/dxi0 = h2 h3 d/deta0,
 * d/dxi1 = h0 d/dxi0 + h3 d/deta1, d/dxi2 = h0 d/dxi0 + h1 h3 d/deta1 + d/deta2 with
 * h0 = (1+eta0)/2, h1 = (1+eta1)/2, h2 = 2/(1-eta1), h3 = 2/(1-eta2); then
 * out_c = sum_a df[3c+a] d/dxi_a */
void or_physderiv_tet(int nq0, int nq1, int nq2, const double *D0, const double *D1,
                      const double *D2, const double *z0, const double *z1, const double *z2,
                      int E, int deformed, const double *df, const double *in, double *out0,
                      double *out1, double *out2)
{
    const int Q = nq0 * nq1 * nq2;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
    double *g0 = (double *)malloc(sizeof(double) * Q), *g1 = (double *)malloc(sizeof(double) * Q),
           *g2 = (double *)malloc(sizeof(double) * Q);
    for (int e = 0; e < E; ++e) {
        double *o0 = out0 + (size_t)e * Q, *o1 = out1 + (size_t)e * Q, *o2 = out2 + (size_t)e * Q;
        or_deriv3_elmt(nq0, nq1, nq2, D0, D1, D2, in + (size_t)e * Q, g0, g1, g2);
        for (int k = 0, c = 0; k < nq2; ++k) {
            const double x2 = 2.0 / (1.0 - z2[k]);
            for (int j = 0; j < nq1; ++j) {
                const double x = (2.0 / (1.0 - z1[j])) * x2;
                for (int i = 0; i < nq0; ++i, ++c) {
                    const double d0 = x * g0[c];
                    o0[c] = d0;
                    g0[c] = d0;
                }
            }
        }
        for (int k = 0, c = 0; k < nq2; ++k) {
            const double x2 = 2.0 / (1.0 - z2[k]);
            for (int j = 0; j < nq1; ++j)
                for (int i = 0; i < nq0; ++i, ++c) {
                    const double a0 = 0.5 * (1.0 + z0[i]) * g0[c];
                    g0[c] = a0;
                    const double d1 = x2 * g1[c];
                    o1[c] = a0 + d1;
                    g1[c] = d1;
                }
        }
        for (int k = 0, c = 0; k < nq2; ++k)
            for (int j = 0; j < nq1; ++j) {
                const double x1 = 0.5 * (1.0 + z1[j]);
                for (int i = 0; i < nq0; ++i, ++c) {
                    double o = fma(g1[c], x1, g0[c]);
                    o2[c] = o + g2[c];
                }
            }
        double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (!deformed)
            for (int a = 0; a < 9; ++a) f[a] = df[a * np + e];
        for (int c = 0; c < Q; ++c) {
            if (deformed)
                for (int a = 0; a < 9; ++a) f[a] = df[a * np + (size_t)e * Q + c];
            const double d0 = o0[c], d1 = o1[c], d2 = o2[c];
            double o;
            o = d0 * f[0]; o = fma(d1, f[1], o); o = fma(d2, f[2], o); o0[c] = o;
            o = d0 * f[3]; o = fma(d1, f[4], o); o = fma(d2, f[5], o); o1[c] = o;
            o = d0 * f[6]; o = fma(d1, f[7], o); o = fma(d2, f[8], o); o2[c] = o;
        }
    }
    free(g0);
    free(g1);
    free(g2);
}

/* HelmholtzTetImpl: lambda * mass + the weak Laplacian, the metric formed per point from
 * the derivative factors and the collapse factors h0..h3 (Helmholtz.h:1990-2004) */
void or_helmholtz_tet(int nm, int nq0, int nq1, int nq2, const double *b0, const double *db0,
                      const double *b1, const double *db1, const double *b2, const double *db2,
                      const double *D0, const double *D1, const double *D2, const double *z0,
                      const double *z1, const double *z2, const double *w0, const double *w1,
                      const double *w2, int E, int deformed, const double *jac,
                      const double *df, double lambda, const double *in, double *out)
{
    const int M = nm * (nm + 1) * (nm + 2) / 6, Q = nq0 * nq1 * nq2;
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
        bwd_tet_elmt(nm, nq0, nq1, nq2, b0, b1, b2, in + (size_t)e * M, ws, ws + nm * nm, bwd);
        iprod_tet_elmt(nm, nq0, nq1, nq2, b0, b1, b2, w0, w1, w2, deformed, jp, bwd, ws,
                       ws + nq1 * nq2, o, 1, 0, lambda);
        or_deriv3_elmt(nq0, nq1, nq2, D0, D1, D2, bwd, dv0, dv1, dv2);
        for (int k = 0, c = 0; k < nq2; ++k) {
            const double h3 = 2.0 / (1.0 - z2[k]);
            for (int j = 0; j < nq1; ++j) {
                const double h1 = 0.5 * (1.0 + z1[j]), h2 = 2.0 / (1.0 - z1[j]);
                const double h2h3 = h2 * h3, h1h3 = h1 * h3;
                for (int i = 0; i < nq0; ++i, ++c) {
                    if (deformed)
                        for (int a = 0; a < 9; ++a) f[a] = df[a * np + (size_t)e * Q + c];
                    const double h0h2h3 = 0.5 * (1.0 + z0[i]) * h2h3;
                    const double t1 = fma(f[0], h2h3, h0h2h3 * (f[1] + f[2]));
                    const double t2 = fma(f[3], h2h3, h0h2h3 * (f[4] + f[5]));
                    const double t3 = fma(f[6], h2h3, h0h2h3 * (f[7] + f[8]));
                    double g0 = t1 * t1; g0 = fma(t2, t2, g0); g0 = fma(t3, t3, g0);
                    double g4 = f[2] * t1; g4 = fma(f[5], t2, g4); g4 = fma(f[8], t3, g4);
                    const double t4 = fma(f[2], h1h3, f[1] * h3);
                    const double t5 = fma(f[5], h1h3, f[4] * h3);
                    const double t6 = fma(f[8], h1h3, f[7] * h3);
                    double g3 = t1 * t4; g3 = fma(t2, t5, g3); g3 = fma(t3, t6, g3);
                    double g1 = t4 * t4; g1 = fma(t5, t5, g1); g1 = fma(t6, t6, g1);
                    double g5 = f[2] * t4; g5 = fma(f[5], t5, g5); g5 = fma(f[8], t6, g5);
                    double g2 = f[2] * f[2]; g2 = fma(f[5], f[5], g2); g2 = fma(f[8], f[8], g2);
                    const double a0 = dv0[c], a1 = dv1[c], a2 = dv2[c];
                    double t;
                    t = g0 * a0; t = fma(g3, a1, t); t = fma(g4, a2, t); dv0[c] = t;
                    t = g3 * a0; t = fma(g1, a1, t); t = fma(g5, a2, t); dv1[c] = t;
                    t = g4 * a0; t = fma(g5, a1, t); t = fma(g2, a2, t); dv2[c] = t;
                }
            }
        }
        double *s1 = ws, *s2 = ws + nq1 * nq2;
        iprod_tet_elmt(nm, nq0, nq1, nq2, db0, b1, b2, w0, w1, w2, deformed, jp, dv0, s1, s2, o, 0, 1, 1.0);
        iprod_tet_elmt(nm, nq0, nq1, nq2, b0, db1, b2, w0, w1, w2, deformed, jp, dv1, s1, s2, o, 0, 1, 1.0);
        iprod_tet_elmt(nm, nq0, nq1, nq2, b0, b1, db2, w0, w1, w2, deformed, jp, dv2, s1, s2, o, 0, 1, 1.0);
    }
    free(ws);
    free(bwd);
    free(dv0);
    free(dv1);
    free(dv2);
}

/* IProductWRTDerivBaseTetImpl (MatrixFreeOps/IProductWRTDerivBase.h:2484-): t_a =
 * sum_c df[3c+a] in_c, tmp0 = 4/((1-eta1)(1-eta2)) (t0 + (1+eta0)/2 (t1 + t2)),
 * tmp1 = 2/(1-eta2) (t1 + (1+eta1)/2 t2), tmp2 = t2 */
void or_iprodderiv_tet(int nm, int nq0, int nq1, int nq2, const double *b0, const double *db0,
                       const double *b1, const double *db1, const double *b2, const double *db2,
                       const double *z0, const double *z1, const double *z2, const double *w0,
                       const double *w1, const double *w2, int E, int deformed,
                       const double *jac, const double *df, const double *in0,
                       const double *in1, const double *in2, double *out)
{
    const int M = nm * (nm + 1) * (nm + 2) / 6, Q = nq0 * nq1 * nq2;
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
            const double f2 = 2.0 / (1.0 - z2[k]);
            for (int j = 0; j < nq1; ++j) {
                const double f3 = 0.5 * (1.0 + z1[j]), f0 = 2.0 * f2 / (1.0 - z1[j]);
                for (int i = 0; i < nq0; ++i, ++c) {
                    if (deformed)
                        for (int a = 0; a < 9; ++a) f[a] = df[a * np + (size_t)e * Q + c];
                    double x[3], t[3];
                    for (int d = 0; d < 3; ++d) x[d] = in[d][(size_t)e * Q + c];
                    for (int a = 0; a < 3; ++a) t[a] = f[a] * x[0] + f[3 + a] * x[1] + f[6 + a] * x[2];
                    const double f1 = 0.5 * (1.0 + z0[i]);
                    t[0] = fma(t[1] + t[2], f1, t[0]);
                    t[0] = t[0] * f0;
                    t[1] = fma(t[2], f3, t[1]);
                    t[1] = t[1] * f2;
                    for (int a = 0; a < 3; ++a) tmp[a * Q + c] = t[a];
                }
            }
        }
        double *o = out + (size_t)e * M;
        iprod_tet_elmt(nm, nq0, nq1, nq2, db0, b1, b2, w0, w1, w2, deformed, jp, tmp, ws, ws + nq1 * nq2, o, 0, 0, 1.0);
        iprod_tet_elmt(nm, nq0, nq1, nq2, b0, db1, b2, w0, w1, w2, deformed, jp, tmp + Q, ws, ws + nq1 * nq2, o, 0, 1, 1.0);
        iprod_tet_elmt(nm, nq0, nq1, nq2, b0, b1, db2, w0, w1, w2, deformed, jp, tmp + 2 * Q, ws, ws + nq1 * nq2, o, 0, 1, 1.0);
    }
    free(tmp);
    free(ws);
}
