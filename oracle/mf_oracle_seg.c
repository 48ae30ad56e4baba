/* Oracle (test infrastructure): scalar C restatement of the reference's MatrixFree segment
 * operators, checked by the tests against dense one-dimensional matrices and closed forms.
 * A segment carries nm Modified_A modes on nq GLL points; ncoord (1 to 3) is the coordinate
 * dimension of the mesh it lies in.  Layouts as the GPU path: coefficients u[e*nm + p],
 * points f[e*nq + i], jac [E*nq] deformed / [E] regular, df [ncoord][E*nq] / [ncoord][E]. */
#include <math.h>
#include <stddef.h>

#include "mf_oracle.h"

/* GeomFactors::ComputeDerivFactors (SpatialDomains/GeomFactors.cpp:399-478) for one
 * expansion direction: g = t.t with t = dx/dxi = (x1 - x0)/2 for a straight segment; the
 * adjoint of a 1 x 1 matrix is 1, so J = sqrt(g) and df[c] = t_c / g. */
void or_geom_seg(const double *verts, int E, int ncoord, int nq, int deformed, double *jac,
                 double *df)
{
    const size_t np = deformed ? (size_t)E * nq : (size_t)E;
    for (int e = 0; e < E; ++e) {
        const double *x0 = verts + (size_t)e * 2 * ncoord, *x1 = x0 + ncoord;
        double t[3] = {0, 0, 0}, g = 0.0;
        for (int c = 0; c < ncoord; ++c) {
            t[c] = 0.5 * (x1[c] - x0[c]);
            g += t[c] * t[c];
        }
        const int n = deformed ? nq : 1;
        for (int i = 0; i < n; ++i) {
            const size_t k = deformed ? (size_t)e * nq + i : (size_t)e;
            jac[k] = sqrt(g);
            for (int c = 0; c < ncoord; ++c) df[c * np + k] = t[c] / g;
        }
    }
}

/* BwdTransSeg (MatrixFreeOps/BwdTrans.h:16-): u_i = sum_p B[p][i] c_p */
void or_bwdtrans_seg(int nm, int nq, const double *b, int E, const double *in, double *out)
{
    for (int e = 0; e < E; ++e)
        for (int i = 0; i < nq; ++i) {
            double s = 0.0;
            for (int p = 0; p < nm; ++p) s += b[p * nq + i] * in[(size_t)e * nm + p];
            out[(size_t)e * nq + i] = s;
        }
}

/* IProductSeg (MatrixFreeOps/IProduct.h:17-; IProductSegKernel): out_p = sum_i B[p][i] w_i
 * J_i f_i */
void or_iproduct_seg(int nm, int nq, const double *b, const double *w, int E, int deformed,
                     const double *jac, const double *in, double *out)
{
    for (int e = 0; e < E; ++e)
        for (int p = 0; p < nm; ++p) {
            double s = 0.0;
            for (int i = 0; i < nq; ++i) {
                const double j = jac[deformed ? (size_t)e * nq + i : (size_t)e];
                s += b[p * nq + i] * w[i] * j * in[(size_t)e * nq + i];
            }
            out[(size_t)e * nm + p] = s;
        }
}

/* PhysDerivSeg (MatrixFreeOps/PhysDeriv.h:17-): du/dxi at the points by the collocation
 * matrix (D[j][i] = l_j'(z_i)), then out_c = df[c] du/dxi for each coordinate c */
void or_physderiv_seg(int nq, const double *D, int E, int deformed, int ncoord,
                      const double *df, const double *in, double *out)
{
    const size_t np = deformed ? (size_t)E * nq : (size_t)E;
    for (int e = 0; e < E; ++e)
        for (int i = 0; i < nq; ++i) {
            double s = 0.0;
            for (int j = 0; j < nq; ++j) s += D[j * nq + i] * in[(size_t)e * nq + j];
            const size_t k = deformed ? (size_t)e * nq + i : (size_t)e;
            for (int c = 0; c < ncoord; ++c)
                out[(size_t)c * E * nq + (size_t)e * nq + i] = df[c * np + k] * s;
        }
}

/* IProductWRTDerivBaseSeg (MatrixFreeOps/IProductWRTDerivBase.h:18-,
 * IProductWRTDerivBaseSegImpl): t_i = sum_c df[c] in_c, then the inner product with the
 * derivative table dB */
void or_iprodderiv_seg(int nm, int nq, const double *db, const double *w, int E, int deformed,
                       int ncoord, const double *jac, const double *df, const double *in,
                       double *out)
{
    const size_t np = deformed ? (size_t)E * nq : (size_t)E;
    for (int e = 0; e < E; ++e)
        for (int p = 0; p < nm; ++p) {
            double s = 0.0;
            for (int i = 0; i < nq; ++i) {
                const size_t k = deformed ? (size_t)e * nq + i : (size_t)e;
                double t = 0.0;
                for (int c = 0; c < ncoord; ++c)
                    t += df[c * np + k] * in[(size_t)c * E * nq + (size_t)e * nq + i];
                s += db[p * nq + i] * w[i] * jac[k] * t;
            }
            out[(size_t)e * nm + p] = s;
        }
}
