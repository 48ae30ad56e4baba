/*
 * CPU oracle, prisms (TEST INFRASTRUCTURE; see mf_oracle.h).  Plain-C restatement of the
 * reference's collapsed-coordinate prism path:
 *   straight-sided PrismGeom   vertex order of StdPrismExp (0..3 the xi2 = -1 quad, 4, 5
 *                              the collapsed top edge), GeomFactors.cpp:420-474
 *   kernels                    MatrixFreeOps/BwdTransKernels.hpp:224-300 (BwdTransPrism),
 *                              IProductKernels.hpp:315-449 (IProductPrism),
 *                              PhysDerivKernels.hpp:219-294, :377-460 (PhysDerivPrism),
 *                              Helmholtz.h:1291-1460 (HelmholtzPrismImpl)
 * with the loop and summation order of those kernels (element interleave width 1).
 *
 * Expansion: Modified_A along eta0 (nq0 GLL) and xi1 (nq1 GLL), Modified_B along eta2 (nq2
 * Gauss-Radau, alpha = 1) with the (p, r) modes of direction 2; nm modes per direction.
 * Coefficients u[e*M + mode], mode = sum_{p'<p} nm (nm - p') + q (nm - p) + r (r fastest),
 * M = nm * nm (nm + 1) / 2; points f[e*Q + (k*nq1 + j)*nq0 + i].  w2 is the Helper's
 * scaled weight (0.5 times the Gauss-Radau alpha = 1 weight).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mf_oracle.h"

/* x(xi) = sum_v N_v x_v: triangle (xi0, xi2) shape functions la = -(xi0+xi2)/2,
 * lb = (1+xi0)/2, lc = (1+xi2)/2 times the segment ones of xi1 */
static void prism_map(const double *v, const double *xi, double *deriv, double *x)
{
    const double la = -0.5 * (xi[0] + xi[2]), lb = 0.5 * (1 + xi[0]), lc = 0.5 * (1 + xi[2]);
    const double sl = 0.5 * (1 - xi[1]), sh = 0.5 * (1 + xi[1]);
    const double N[6] = {la * sl, lb * sl, lb * sh, la * sh, lc * sl, lc * sh};
    /* dN/dxi_a */
    const double dN[6][3] = {{-0.5 * sl, -0.5 * la, -0.5 * sl}, {0.5 * sl, -0.5 * lb, 0.0},
                             {0.5 * sh, 0.5 * lb, 0.0},         {-0.5 * sh, 0.5 * la, -0.5 * sh},
                             {0.0, -0.5 * lc, 0.5 * sl},        {0.0, 0.5 * lc, 0.5 * sh}};
    if (deriv) memset(deriv, 0, 9 * sizeof(double));
    if (x) memset(x, 0, 3 * sizeof(double));
    for (int n = 0; n < 6; ++n)
        for (int c = 0; c < 3; ++c) {
            if (x) x[c] += N[n] * v[n * 3 + c];
            if (deriv)
                for (int a = 0; a < 3; ++a) deriv[a * 3 + c] += dN[n][a] * v[n * 3 + c];
        }
}

static void prism_xi(const double *z0, const double *z1, const double *z2, int i, int j, int k,
                     double *xi)
{
    xi[0] = 0.5 * (1.0 + z0[i]) * (1.0 - z2[k]) - 1.0;
    xi[1] = z1[j];
    xi[2] = z2[k];
}

void or_geom_prism(const double *verts, int E, int nq0, int nq1, int nq2, const double *z0,
                   const double *z1, const double *z2, int deformed, double *jac, double *df)
{
    const int Q = nq0 * nq1 * nq2;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
    for (int e = 0; e < E; ++e) {
        const int nqq = deformed ? Q : 1;
        for (int q = 0; q < nqq; ++q) {
            const int i = q % nq0, j = (q / nq0) % nq1, k = q / (nq0 * nq1);
            double xi[3], deriv[9], J, fac[9];
            prism_xi(z0, z1, z2, i, j, k, xi);
            prism_map(verts + (size_t)e * 18, xi, deriv, NULL);
            or_deriv_factors(3, deriv, &J, fac);
            const size_t idx = deformed ? (size_t)e * Q + q : (size_t)e;
            jac[idx] = J;
            for (int d = 0; d < 9; ++d) df[d * np + idx] = fac[d];
        }
    }
}

void or_coords_prism(const double *verts, int E, int nq0, int nq1, int nq2, const double *z0,
                     const double *z1, const double *z2, double *xyz)
{
    const int Q = nq0 * nq1 * nq2;
    for (int e = 0; e < E; ++e)
        for (int q = 0; q < Q; ++q) {
            const int i = q % nq0, j = (q / nq0) % nq1, k = q / (nq0 * nq1);
            double xi[3], x[3];
            prism_xi(z0, z1, z2, i, j, k, xi);
            prism_map(verts + (size_t)e * 18, xi, NULL, x);
            for (int c = 0; c < 3; ++c) xyz[c * (size_t)E * Q + (size_t)e * Q + q] = x[c];
        }
}

/* BwdTransPrismKernel */
static void bwd_prism_elmt(int nm, int nq0, int nq1, int nq2, const double *b0,
                           const double *b1, const double *b2, const double *in, double *fpq,
                           double *fp, double *out)
{
    for (int k = 0, cnt = 0; k < nq2; ++k) {
        int mode_pqr = 0, mode_pq = 0, mode_pr = 0;
        for (int p = 0; p < nm; ++p) {
            for (int q = 0; q < nm; ++q, ++mode_pq) {
                double prod = 0.0;
                for (int r = 0; r < nm - p; ++r, ++mode_pqr)
                    prod = fma(in[mode_pqr], b2[(mode_pr + r) * nq2 + k], prod);
                fpq[mode_pq] = prod;
            }
            mode_pr += nm - p;
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
                const double basis_2 = b2[nq2 + k], basis_0 = b0[nq0 + i];
                for (int q = 0; q < nm; ++q)
                    v = fma(basis_2 * b1[q * nq1 + j], basis_0 * in[q * nm + 1], v);
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

/* IProductPrismKernel (CORRECT: the singular-edge modes (0, q, 1) also get phi_1(eta0)) */
static void iprod_prism_elmt(int nm, int nq0, int nq1, int nq2, const double *b0,
                             const double *b1, const double *b2, const double *w0,
                             const double *w1, const double *w2, int deformed,
                             const double *jac, const double *in, double *skj, double *sk,
                             double *corr, double *out, int scale, int append, double sc)
{
    int mode_pr = 0, mode_pqr = 0;
    for (int p = 0; p < nm; ++p) {
        int cnt = 0, cnt_kj = 0;
        for (int k = 0; k < nq2; ++k)
            for (int j = 0; j < nq1; ++j, ++cnt_kj) {
                double s = 0.0;
                for (int i = 0; i < nq0; ++i, ++cnt) {
                    const double jv = deformed ? jac[nq0 * nq1 * k + nq0 * j + i] : jac[0];
                    const double prod = b0[nq0 * p + i] * jv * w0[i];
                    s = fma(prod, in[cnt], s);
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
            for (int r = 0; r < nm - p; ++r, ++mode_pqr) {
                double s = 0.0;
                for (int k = 0; k < nq2; ++k) s = fma(b2[(mode_pr + r) * nq2 + k] * w2[k], sk[k], s);
                store(&out[mode_pqr], s, scale, append, sc);
            }
        }
        mode_pr += nm - p;
    }
    for (int q = 0; q < nm; ++q) corr[q] = 0.0;
    for (int k = 0, cnt = 0; k < nq2; ++k) {
        double kw = w2[k];
        if (!deformed) kw = kw * jac[0];
        for (int j = 0; j < nq1; ++j) {
            const double kjw = kw * w1[j];
            for (int i = 0; i < nq0; ++i, ++cnt) {
                double prod = kjw * w0[i] * in[cnt];
                if (deformed) prod *= jac[k * nq1 * nq0 + j * nq0 + i];
                const double basis_2 = b2[nq2 + k], basis_0 = b0[nq0 + i];
                for (int q = 0; q < nm; ++q) corr[q] = fma(basis_2 * b1[q * nq1 + j], basis_0 * prod, corr[q]);
            }
        }
    }
    for (int q = 0; q < nm; ++q) store(&out[nm * q + 1], corr[q], scale, 1, sc);
}

/* PhysDerivTensor3DKernel for nq0 x nq1 x nq2 points */
void or_deriv3_elmt(int nq0, int nq1, int nq2, const double *D0, const double *D1,
                    const double *D2, const double *in, double *d0, double *d1, double *d2)
{
    for (int i = 0; i < nq0; ++i)
        for (int jk = 0; jk < nq1 * nq2; ++jk) {
            double s = 0.0;
            for (int k = 0; k < nq0; ++k) s = fma(D0[k * nq0 + i], in[jk * nq0 + k], s);
            d0[jk * nq0 + i] = s;
        }
    for (int blk = 0; blk < nq2; ++blk) {
        const int st = blk * nq0 * nq1;
        for (int i = 0; i < nq0; ++i)
            for (int j = 0; j < nq1; ++j) {
                double s = 0.0;
                for (int k = 0; k < nq1; ++k) s = fma(in[st + k * nq0 + i], D1[k * nq1 + j], s);
                d1[st + j * nq0 + i] = s;
            }
    }
    for (int i = 0; i < nq0 * nq1; ++i)
        for (int j = 0; j < nq2; ++j) {
            double s = 0.0;
            for (int k = 0; k < nq2; ++k) s = fma(in[k * nq0 * nq1 + i], D2[k * nq2 + j], s);
            d2[j * nq0 * nq1 + i] = s;
        }
}

void or_bwdtrans_prism(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                       const double *b2, int E, const double *in, double *out)
{
    const int M = nm * nm * (nm + 1) / 2, Q = nq0 * nq1 * nq2;
#pragma omp parallel
    {
        double *fpq = (double *)malloc(sizeof(double) * (nm * nm + nm));
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e)
            bwd_prism_elmt(nm, nq0, nq1, nq2, b0, b1, b2, in + (size_t)e * M, fpq, fpq + nm * nm,
                           out + (size_t)e * Q);
        free(fpq);
    }
}

void or_iproduct_prism(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                       const double *b2, const double *w0, const double *w1, const double *w2,
                       int E, int deformed, const double *jac, const double *in, double *out)
{
    const int M = nm * nm * (nm + 1) / 2, Q = nq0 * nq1 * nq2;
#pragma omp parallel
    {
        double *ws = (double *)malloc(sizeof(double) * (nq1 * nq2 + nq2 + nm));
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e)
            iprod_prism_elmt(nm, nq0, nq1, nq2, b0, b1, b2, w0, w1, w2, deformed,
                             jac + (deformed ? (size_t)e * Q : (size_t)e), in + (size_t)e * Q,
                             ws, ws + nq1 * nq2, ws + nq1 * nq2 + nq2, out + (size_t)e * M, 0, 0,
                             1.0);
        free(ws);
    }
}

void or_physderiv_prism(int nq0, int nq1, int nq2, const double *D0, const double *D1,
                        const double *D2, const double *z0, const double *z2, int E,
                        int deformed, const double *df, const double *in, double *out0,
                        double *out1, double *out2)
{
    const int Q = nq0 * nq1 * nq2;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        double *d0 = out0 + (size_t)e * Q, *d1 = out1 + (size_t)e * Q, *d2 = out2 + (size_t)e * Q;
        or_deriv3_elmt(nq0, nq1, nq2, D0, D1, D2, in + (size_t)e * Q, d0, d1, d2);
        double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (!deformed)
            for (int a = 0; a < 9; ++a) f[a] = df[a * np + e];
        for (int k = 0, c = 0; k < nq2; ++k) {
            const double x2 = 2.0 / (1.0 - z2[k]);
            for (int j = 0; j < nq1; ++j)
                for (int i = 0; i < nq0; ++i, ++c) {
                    if (deformed)
                        for (int a = 0; a < 9; ++a) f[a] = df[a * np + (size_t)e * Q + c];
                    const double a0 = d0[c] * x2;
                    const double a2 = fma(0.5 * (1.0 + z0[i]), a0, d2[c]);
                    const double a1 = d1[c];
                    double o;
                    o = a0 * f[0]; o = fma(a1, f[1], o); o = fma(a2, f[2], o); d0[c] = o;
                    o = a0 * f[3]; o = fma(a1, f[4], o); o = fma(a2, f[5], o); d1[c] = o;
                    o = a0 * f[6]; o = fma(a1, f[7], o); o = fma(a2, f[8], o); d2[c] = o;
                }
        }
    }
}

void or_helmholtz_prism(int nm, int nq0, int nq1, int nq2, const double *b0, const double *db0,
                        const double *b1, const double *db1, const double *b2,
                        const double *db2, const double *D0, const double *D1,
                        const double *D2, const double *z0, const double *z2,
                        const double *w0, const double *w1, const double *w2, int E,
                        int deformed, const double *jac, const double *df, double lambda,
                        const double *in, double *out)
{
    const int M = nm * nm * (nm + 1) / 2, Q = nq0 * nq1 * nq2;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
#pragma omp parallel
    {
        double *bwd = (double *)malloc(sizeof(double) * 4 * Q);
        double *dv0 = bwd + Q, *dv1 = dv0 + Q, *dv2 = dv1 + Q;
        const int nw = nq1 * nq2 + nq2 + nm + nm * nm + nm;
        double *ws = (double *)malloc(sizeof(double) * nw);
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e) {
            const double *jp = jac + (deformed ? (size_t)e * Q : (size_t)e);
            double *o = out + (size_t)e * M;
            double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (!deformed)
                for (int a = 0; a < 9; ++a) f[a] = df[a * np + e];
            bwd_prism_elmt(nm, nq0, nq1, nq2, b0, b1, b2, in + (size_t)e * M, ws, ws + nm * nm, bwd);
            iprod_prism_elmt(nm, nq0, nq1, nq2, b0, b1, b2, w0, w1, w2, deformed, jp, bwd, ws,
                             ws + nq1 * nq2, ws + nq1 * nq2 + nq2, o, 1, 0, lambda);
            or_deriv3_elmt(nq0, nq1, nq2, D0, D1, D2, bwd, dv0, dv1, dv2);
            for (int k = 0, c = 0; k < nq2; ++k) {
                const double h1 = 2.0 / (1.0 - z2[k]);
                for (int j = 0; j < nq1; ++j)
                    for (int i = 0; i < nq0; ++i, ++c) {
                        const double h0 = 0.5 * (1.0 + z0[i]);
                        if (deformed)
                            for (int a = 0; a < 9; ++a) f[a] = df[a * np + (size_t)e * Q + c];
                        const double t1 = h1 * (h0 * f[2] + f[0]);
                        const double t2 = h1 * (h0 * f[5] + f[3]);
                        const double t3 = h1 * (h0 * f[8] + f[6]);
                        double g0 = t1 * t1; g0 = fma(t2, t2, g0); g0 = fma(t3, t3, g0);
                        double g3 = f[1] * t1; g3 = fma(f[4], t2, g3); g3 = fma(f[7], t3, g3);
                        double g4 = f[2] * t1; g4 = fma(f[5], t2, g4); g4 = fma(f[8], t3, g4);
                        double g1 = f[1] * f[1]; g1 = fma(f[4], f[4], g1); g1 = fma(f[7], f[7], g1);
                        double g2 = f[2] * f[2]; g2 = fma(f[5], f[5], g2); g2 = fma(f[8], f[8], g2);
                        double g5 = f[1] * f[2]; g5 = fma(f[4], f[5], g5); g5 = fma(f[7], f[8], g5);
                        const double a0 = dv0[c], a1 = dv1[c], a2 = dv2[c];
                        double t;
                        t = g0 * a0; t = fma(g3, a1, t); t = fma(g4, a2, t); dv0[c] = t;
                        t = g3 * a0; t = fma(g1, a1, t); t = fma(g5, a2, t); dv1[c] = t;
                        t = g4 * a0; t = fma(g5, a1, t); t = fma(g2, a2, t); dv2[c] = t;
                    }
            }
            double *s1 = ws, *s2 = ws + nq1 * nq2, *s3 = s2 + nq2;
            iprod_prism_elmt(nm, nq0, nq1, nq2, db0, b1, b2, w0, w1, w2, deformed, jp, dv0, s1, s2, s3, o, 0, 1, 1.0);
            iprod_prism_elmt(nm, nq0, nq1, nq2, b0, db1, b2, w0, w1, w2, deformed, jp, dv1, s1, s2, s3, o, 0, 1, 1.0);
            iprod_prism_elmt(nm, nq0, nq1, nq2, b0, b1, db2, w0, w1, w2, deformed, jp, dv2, s1, s2, s3, o, 0, 1, 1.0);
        }
        free(ws);
        free(bwd);
    }
}

/* IProductWRTDerivBasePrismImpl (MatrixFreeOps/IProductWRTDerivBase.h:1630-): t_a =
 * sum_c df[3c+a] in_c, tmp0 = 2/(1-eta2) (t0 + (1+eta0)/2 t2), tmp1 = t1, tmp2 = t2, and
 * the inner products with (dB0, B1, B2), (B0, dB1, B2), (B0, B1, dB2) */
void or_iprodderiv_prism(int nm, int nq0, int nq1, int nq2, const double *b0,
                         const double *db0, const double *b1, const double *db1,
                         const double *b2, const double *db2, const double *z0,
                         const double *z2, const double *w0, const double *w1, const double *w2,
                         int E, int deformed, const double *jac, const double *df,
                         const double *in0, const double *in1, const double *in2, double *out)
{
    const int M = nm * nm * (nm + 1) / 2, Q = nq0 * nq1 * nq2;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
    double *tmp = (double *)malloc(sizeof(double) * 3 * Q);
    double *ws = (double *)malloc(sizeof(double) * (nq1 * nq2 + nq2 + nm));
    const double *in[3] = {in0, in1, in2};
    for (int e = 0; e < E; ++e) {
        const double *jp = jac + (deformed ? (size_t)e * Q : (size_t)e);
        double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (!deformed)
            for (int a = 0; a < 9; ++a) f[a] = df[a * np + e];
        for (int k = 0, c = 0; k < nq2; ++k) {
            const double f0 = 2.0 / (1.0 - z2[k]);
            for (int j = 0; j < nq1; ++j)
                for (int i = 0; i < nq0; ++i, ++c) {
                    if (deformed)
                        for (int a = 0; a < 9; ++a) f[a] = df[a * np + (size_t)e * Q + c];
                    double x[3], t[3];
                    for (int d = 0; d < 3; ++d) x[d] = in[d][(size_t)e * Q + c];
                    for (int a = 0; a < 3; ++a) t[a] = f[a] * x[0] + f[3 + a] * x[1] + f[6 + a] * x[2];
                    const double hf1 = 0.5 * (1.0 + z0[i]);
                    t[0] = t[0] * f0;
                    t[0] = fma(hf1 * t[2], f0, t[0]);
                    for (int a = 0; a < 3; ++a) tmp[a * Q + c] = t[a];
                }
        }
        double *o = out + (size_t)e * M, *s1 = ws, *s2 = ws + nq1 * nq2, *s3 = s2 + nq2;
        iprod_prism_elmt(nm, nq0, nq1, nq2, db0, b1, b2, w0, w1, w2, deformed, jp, tmp, s1, s2, s3, o, 0, 0, 1.0);
        iprod_prism_elmt(nm, nq0, nq1, nq2, b0, db1, b2, w0, w1, w2, deformed, jp, tmp + Q, s1, s2, s3, o, 0, 1, 1.0);
        iprod_prism_elmt(nm, nq0, nq1, nq2, b0, b1, db2, w0, w1, w2, deformed, jp, tmp + 2 * Q, s1, s2, s3, o, 0, 1, 1.0);
    }
    free(tmp);
    free(ws);
}
