/*
 * CPU oracle for the matrix-free spectral/hp operator path (TEST INFRASTRUCTURE).
 *
 * A plain-C restatement of the reference CPU algorithms.  It is the checker for the
 * HIP product path and the timed CPU baseline; the product never links or calls it.
 * Every function cites the reference file:line it follows (paths relative to the
 * reference repository root).  The element loops follow the reference's width-1
 * (scalar tinysimd, the default NEKTAR_ENABLE_SIMD_* = OFF build) summation order.
 */
#include "mf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void or_set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* ------------------------------------------------------------------------- */
/* Polylib                                                                     */
/* ------------------------------------------------------------------------- */

/* Jacobi polynomial P_n^{(alpha,beta)} and derivative by the three-term
 * recurrence: library/LibUtilities/Polylib/Polylib.cpp jacobfd (~:1034-1120). */
void or_jacobfd(int np, const double *z, double *poly_in, double *polyd, int n,
                double alpha, double beta)
{
    if (!np) return;
    if (n == 0) {
        for (int i = 0; i < np; ++i) {
            if (poly_in) poly_in[i] = 1.0;
            if (polyd) polyd[i] = 0.0;
        }
        return;
    }
    if (n == 1) {
        for (int i = 0; i < np; ++i) {
            if (poly_in) poly_in[i] = 0.5 * (alpha - beta + (alpha + beta + 2.0) * z[i]);
            if (polyd) polyd[i] = 0.5 * (alpha + beta + 2.0);
        }
        return;
    }
    double apb = alpha + beta;
    double *buf = (double *)malloc(3 * (size_t)np * sizeof(double));
    double *pn1 = buf, *pn2 = buf + np, *poly = poly_in ? poly_in : buf + 2 * np;
    for (int i = 0; i < np; ++i) {
        pn2[i] = 1.0;
        pn1[i] = 0.5 * (alpha - beta + (alpha + beta + 2.0) * z[i]);
    }
    for (int k = 2; k <= n; ++k) {
        double a1 = 2.0 * k * (k + apb) * (2.0 * k + apb - 2.0);
        double a2 = (2.0 * k + apb - 1.0) * (alpha * alpha - beta * beta);
        double a3 = (2.0 * k + apb - 2.0) * (2.0 * k + apb - 1.0) * (2.0 * k + apb);
        double a4 = 2.0 * (k + alpha - 1.0) * (k + beta - 1.0) * (2.0 * k + apb);
        a2 /= a1; a3 /= a1; a4 /= a1;
        for (int i = 0; i < np; ++i) {
            poly[i] = (a2 + a3 * z[i]) * pn1[i] - a4 * pn2[i];
            pn2[i] = pn1[i];
            pn1[i] = poly[i];
        }
    }
    if (polyd) {
        double a1 = n * (alpha - beta);
        double a2 = n * (2.0 * n + alpha + beta);
        double a3 = 2.0 * (n + alpha) * (n + beta);
        double a4 = (2.0 * n + alpha + beta);
        a1 /= a4; a2 /= a4; a3 /= a4;
        for (int i = 0; i < np; ++i) {
            polyd[i] = (a1 - a2 * z[i]) * poly[i] + a3 * pn2[i];
            polyd[i] /= (1.0 - z[i] * z[i]);
        }
    }
    free(buf);
}

/* Zeros of P_n^{(alpha,beta)} by Newton iteration with deflation, Chebyshev
 * starting guesses: Polylib.cpp Jacobz (selected because POLYNOMIAL_DEFLATION is
 * defined, Polylib.cpp:55-63). */
void or_jacobz(int n, double *z, double alpha, double beta)
{
    const double EPS = 100 * 2.220446049250313e-16;
    const int STOP = 30;
    double dth = M_PI / (2.0 * (double)n), rlast = 0.0;
    for (int k = 0; k < n; ++k) {
        double r = -cos((2.0 * (double)k + 1.0) * dth);
        if (k) r = 0.5 * (r + rlast);
        for (int j = 1; j < STOP; ++j) {
            double poly, pder, sum = 0.0;
            or_jacobfd(1, &r, &poly, &pder, n, alpha, beta);
            for (int i = 0; i < k; ++i) sum += 1.0 / (r - z[i]);
            double delr = -poly / (pder - sum * poly);
            r += delr;
            if (fabs(delr) < EPS) break;
        }
        z[k] = r;
        rlast = r;
    }
}

/* Gauss-Lobatto-Legendre zeros and weights: Polylib.cpp:195-232 (zwglj) with
 * alpha = beta = 0, where the gamma-function prefactor reduces to 2/(np(np-1)). */
void or_zwglj(double *z, double *w, int np)
{
    if (np == 1) { z[0] = 0.0; w[0] = 2.0; return; }
    if (np == 2) { z[0] = -1.0; z[1] = 1.0; w[0] = 1.0; w[1] = 1.0; return; }
    z[0] = -1.0;
    z[np - 1] = 1.0;
    or_jacobz(np - 2, z + 1, 1.0, 1.0);
    or_jacobfd(np, z, w, NULL, np - 1, 0.0, 0.0);
    double fac = 2.0 / ((double)np * (double)(np - 1));
    for (int i = 0; i < np; ++i) w[i] = fac / (w[i] * w[i]);
}

/* GLL collocation derivative matrix: Polylib.cpp:690-730 (Dglj), alpha=beta=0.
 * Stored flat as D[a*np + b] = l_a'(z_b), which GaussPoints.cpp:178-245 copies
 * verbatim into the column-major derivative matrix (D(b,a) = l_a'(z_b)). */
void or_Dglj(double *D, const double *z, int np)
{
    if (np <= 1) { D[0] = 0.0; return; }
    double *pd = (double *)malloc((size_t)np * sizeof(double));
    /* pd[0] = 2 (-1)^np gammaFracGammaF(np,0,np-1,0) / gammaF(2) = 2 (-1)^np (np-1) */
    pd[0] = 2.0 * pow(-1.0, np) * (double)(np - 1);
    /* jacobd(np-2, z+1, pd+1, np-2, 1, 1) = 0.5 (np+1) P_{np-3}^{(2,2)}(z) */
    if (np > 2) {
        or_jacobfd(np - 2, z + 1, pd + 1, NULL, np - 3, 2.0, 2.0);
        for (int i = 1; i < np - 1; ++i) pd[i] *= 0.5 * (double)(np + 1);
    }
    for (int i = 1; i < np - 1; ++i) pd[i] *= (1.0 - z[i] * z[i]);
    /* pd[np-1] = -2 gammaFracGammaF(np,0,np-1,0) / gammaF(2) */
    pd[np - 1] = -2.0 * (double)(np - 1);
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j) {
            if (i != j)
                D[i * np + j] = pd[j] / (pd[i] * (z[j] - z[i]));
            else if (j == 0)
                D[i * np + j] = (0.0 - (np - 1) * (np + 0.0)) / (2.0 * 2.0);
            else if (j == np - 1)
                D[i * np + j] = -(0.0 - (np - 1) * (np + 0.0)) / (2.0 * 2.0);
            else
                D[i * np + j] = 0.0 / (2.0 * (1.0 - z[j] * z[j]));
        }
    free(pd);
}

/* Modified_A basis on GLL points: Basis.cpp:392-421.
 * bdata[p*nq+i] = phi_p(z_i); dbdata = D * bdata (the Blas::Dgemm at :418). */
void or_basis_modified_a(int nm, int nq, double *z, double *w, double *D,
                         double *bdata, double *dbdata)
{
    or_zwglj(z, w, nq);
    or_Dglj(D, z, nq);
    for (int i = 0; i < nq; ++i) {
        bdata[i] = 0.5 * (1 - z[i]);
        bdata[nq + i] = 0.5 * (1 + z[i]);
    }
    for (int p = 2; p < nm; ++p) {
        double *mode = bdata + p * nq;
        or_jacobfd(nq, z, mode, NULL, p - 2, 1.0, 1.0);
        for (int i = 0; i < nq; ++i) mode[i] *= bdata[i] * bdata[nq + i];
    }
    /* column-major C(nq x nm) = D(nq x nq) B(nq x nm), D(i,k) = flat[k*nq+i] */
    for (int p = 0; p < nm; ++p)
        for (int i = 0; i < nq; ++i) {
            double s = 0.0;
            for (int k = 0; k < nq; ++k) s += D[k * nq + i] * bdata[p * nq + k];
            dbdata[p * nq + i] = s;
        }
}

/* ------------------------------------------------------------------------- */
/* Geometric factors                                                           */
/* ------------------------------------------------------------------------- */

/* 3x3 adjugate of a symmetric-layout matrix stored row-major (GeomFactors::Adjoint). */
static void adj3(const double *a, double *b)
{
    b[0] = a[4] * a[8] - a[5] * a[7];
    b[1] = a[2] * a[7] - a[1] * a[8];
    b[2] = a[1] * a[5] - a[2] * a[4];
    b[3] = a[5] * a[6] - a[3] * a[8];
    b[4] = a[0] * a[8] - a[2] * a[6];
    b[5] = a[2] * a[3] - a[0] * a[5];
    b[6] = a[3] * a[7] - a[4] * a[6];
    b[7] = a[1] * a[6] - a[0] * a[7];
    b[8] = a[0] * a[4] - a[1] * a[3];
}

/* From the map derivatives deriv[i][k] = dx_k/dxi_i at one point compute
 * jac = sqrt(det g) and factors[k*dim+j] = sum_i deriv[i][k] g^{-1}[i][j]:
 * GeomFactors.cpp:420-474 (ComputeDerivFactors).  dim = coordim. */
void or_deriv_factors(int dim, const double *deriv, double *jac, double *fac)
{
    double g[9] = {0}, gi[9];
    for (int i = 0, l = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j, ++l) {
            double s = 0.0;
            for (int k = 0; k < dim; ++k) s += deriv[i * dim + k] * deriv[j * dim + k];
            g[l] = s;
        }
    double det = 0.0;
    if (dim == 3) {
        adj3(g, gi);
        for (int i = 0; i < 3; ++i) det += g[i] * gi[i * 3];
    } else {
        gi[0] = g[3]; gi[1] = -g[1]; gi[2] = -g[2]; gi[3] = g[0];
        for (int i = 0; i < 2; ++i) det += g[i] * gi[i * 2];
    }
    for (int i = 0; i < dim * dim; ++i) gi[i] /= det;
    *jac = sqrt(det);
    for (int k = 0, l = 0; k < dim; ++k)
        for (int j = 0; j < dim; ++j, ++l) {
            double s = 0.0;
            for (int i = 0; i < dim; ++i) s += deriv[i * dim + k] * gi[dim * i + j];
            fac[l] = s;
        }
}

/* Trilinear hex map x(xi) = sum_v x_v N_v(xi) with geometric vertex order
 * (StdHexExp.cpp:1042-1082: bit0 -> xi0=+1, bit1 of (v%4)>1 -> xi1=+1, v>3 -> xi2=+1).
 * Straight-sided hexes use a 2-mode Modified_A coordinate map, whose derivatives
 * the reference interpolates exactly to the target points (GeomFactors.cpp:165-235). */
static const int HEXC[8][3] = {{0,0,0},{1,0,0},{1,1,0},{0,1,0},{0,0,1},{1,0,1},{1,1,1},{0,1,1}};
static const int QUADC[4][2] = {{0,0},{1,0},{1,1},{0,1}};

static void hex_map_deriv(const double *v, const double *xi, double *deriv, double *x)
{
    for (int a = 0; a < 9; ++a) deriv[a] = 0.0;
    if (x) x[0] = x[1] = x[2] = 0.0;
    for (int n = 0; n < 8; ++n) {
        double f[3], df[3];
        for (int d = 0; d < 3; ++d) {
            f[d] = HEXC[n][d] ? 0.5 * (1 + xi[d]) : 0.5 * (1 - xi[d]);
            df[d] = HEXC[n][d] ? 0.5 : -0.5;
        }
        double N = f[0] * f[1] * f[2];
        double dN[3] = {df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]};
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) deriv[i * 3 + k] += dN[i] * v[n * 3 + k];
        if (x)
            for (int k = 0; k < 3; ++k) x[k] += N * v[n * 3 + k];
    }
}

static void quad_map_deriv(const double *v, const double *xi, double *deriv, double *x)
{
    for (int a = 0; a < 4; ++a) deriv[a] = 0.0;
    if (x) x[0] = x[1] = 0.0;
    for (int n = 0; n < 4; ++n) {
        double f[2], df[2];
        for (int d = 0; d < 2; ++d) {
            f[d] = QUADC[n][d] ? 0.5 * (1 + xi[d]) : 0.5 * (1 - xi[d]);
            df[d] = QUADC[n][d] ? 0.5 : -0.5;
        }
        double N = f[0] * f[1];
        double dN[2] = {df[0] * f[1], f[0] * df[1]};
        for (int i = 0; i < 2; ++i)
            for (int k = 0; k < 2; ++k) deriv[i * 2 + k] += dN[i] * v[n * 2 + k];
        if (x)
            for (int k = 0; k < 2; ++k) x[k] += N * v[n * 2 + k];
    }
}

void or_geom_hex(const double *verts, int E, int nq, const double *z, int deformed,
                 double *jac, double *df)
{
    const int Q = nq * nq * nq;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        const double *v = verts + (size_t)e * 24;
        int nqq = deformed ? Q : 1;
        for (int q = 0; q < nqq; ++q) {
            int i = q % nq, j = (q / nq) % nq, k = q / (nq * nq);
            double xi[3] = {z[i], z[j], z[k]}, deriv[9], J, fac[9];
            hex_map_deriv(v, xi, deriv, NULL);
            or_deriv_factors(3, deriv, &J, fac);
            size_t idx = deformed ? (size_t)e * Q + q : (size_t)e;
            jac[idx] = J;
            for (int d = 0; d < 9; ++d) df[d * np + idx] = fac[d];
        }
    }
}

void or_geom_quad(const double *verts, int E, int nq, const double *z, int deformed,
                  double *jac, double *df)
{
    const int Q = nq * nq;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
    for (int e = 0; e < E; ++e) {
        const double *v = verts + (size_t)e * 8;
        int nqq = deformed ? Q : 1;
        for (int q = 0; q < nqq; ++q) {
            int i = q % nq, j = q / nq;
            double xi[2] = {z[i], z[j]}, deriv[4], J, fac[4];
            quad_map_deriv(v, xi, deriv, NULL);
            or_deriv_factors(2, deriv, &J, fac);
            size_t idx = deformed ? (size_t)e * Q + q : (size_t)e;
            jac[idx] = J;
            for (int d = 0; d < 4; ++d) df[d * np + idx] = fac[d];
        }
    }
}

void or_coords_hex(const double *verts, int E, int nq, const double *z, double *xyz)
{
    const int Q = nq * nq * nq;
    const size_t np = (size_t)E * Q;
    for (int e = 0; e < E; ++e)
        for (int q = 0; q < Q; ++q) {
            int i = q % nq, j = (q / nq) % nq, k = q / (nq * nq);
            double xi[3] = {z[i], z[j], z[k]}, deriv[9], x[3];
            hex_map_deriv(verts + (size_t)e * 24, xi, deriv, x);
            for (int c = 0; c < 3; ++c) xyz[c * np + (size_t)e * Q + q] = x[c];
        }
}

void or_coords_quad(const double *verts, int E, int nq, const double *z, double *xy)
{
    const int Q = nq * nq;
    const size_t np = (size_t)E * Q;
    for (int e = 0; e < E; ++e)
        for (int q = 0; q < Q; ++q) {
            double xi[2] = {z[q % nq], z[q / nq]}, deriv[4], x[2];
            quad_map_deriv(verts + (size_t)e * 8, xi, deriv, x);
            for (int c = 0; c < 2; ++c) xy[c * np + (size_t)e * Q + q] = x[c];
        }
}

/* ------------------------------------------------------------------------- */
/* Hex kernels (width 1)                                                       */
/* ------------------------------------------------------------------------- */

/* BwdTransHexKernel: MatrixFreeOps/BwdTransKernels.hpp:302-372.  The reference
 * uses nq2 = NUMQUAD1 (:314); with homogeneous points this is nq. */
static void bwd_hex_elmt(int nm, int nq, const double *B, const double *in,
                         double *sum_irq, double *sum_jir, double *out)
{
    for (int i = 0, cnt_irq = 0; i < nq; ++i)
        for (int r = 0, cnt_rqp = 0; r < nm; ++r)
            for (int q = 0; q < nm; ++q, ++cnt_irq) {
                double tmp = in[cnt_rqp] * B[i];
                ++cnt_rqp;
                for (int p = 1; p < nm; ++p, ++cnt_rqp) tmp += in[cnt_rqp] * B[p * nq + i];
                sum_irq[cnt_irq] = tmp;
            }
    for (int j = 0, cnt_jir = 0; j < nq; ++j)
        for (int i = 0, cnt_irq = 0; i < nq; ++i)
            for (int r = 0; r < nm; ++r, ++cnt_jir) {
                double tmp = sum_irq[cnt_irq] * B[j];
                ++cnt_irq;
                for (int q = 1; q < nm; ++q) tmp += sum_irq[cnt_irq++] * B[q * nq + j];
                sum_jir[cnt_jir] = tmp;
            }
    for (int k = 0, cnt_kji = 0; k < nq; ++k)
        for (int j = 0, cnt_jir = 0; j < nq; ++j)
            for (int i = 0; i < nq; ++i, ++cnt_kji) {
                double tmp = sum_jir[cnt_jir] * B[k];
                ++cnt_jir;
                for (int r = 1; r < nm; ++r) tmp += sum_jir[cnt_jir++] * B[r * nq + k];
                out[cnt_kji] = tmp;
            }
}

/* ScaleAppend + IProductHexKernel: MatrixFreeOps/IProductKernels.hpp:14-37, 236-311.
 * mode: 0 = store, 1 = store*scale, 2 = append, 3 = append*scale */
static void iprod_hex_elmt(int nm, int nq, const double *B0, const double *B1,
                           const double *B2, const double *w, int deformed,
                           const double *jac, const double *in, double *sums_kj,
                           double *sums_k, double *out, int scale, int append,
                           double sc)
{
    for (int p = 0; p < nm; ++p) {
        int cnt_kji = 0, cnt_kj = 0;
        for (int k = 0; k < nq; ++k)
            for (int j = 0; j < nq; ++j, ++cnt_kj) {
                double sum_kj = 0.0;
                for (int i = 0; i < nq; ++i, ++cnt_kji) {
                    double jv = deformed ? jac[nq * nq * k + nq * j + i] : jac[0];
                    double prod = in[cnt_kji] * B0[i + nq * p] * jv;
                    sum_kj += prod * w[i];
                }
                sums_kj[cnt_kj] = sum_kj;
            }
        for (int q = 0; q < nm; ++q) {
            cnt_kj = 0;
            for (int k = 0; k < nq; ++k) {
                double sum_k = 0.0;
                for (int j = 0; j < nq; ++j, ++cnt_kj) {
                    double prod = sums_kj[cnt_kj] * B1[q * nq + j];
                    sum_k += prod * w[j];
                }
                sums_k[k] = sum_k;
            }
            for (int r = 0; r < nm; ++r) {
                double sum = 0.0;
                for (int k = 0; k < nq; ++k) {
                    double prod = sums_k[k] * B2[r * nq + k];
                    sum += prod * w[k];
                }
                double *o = &out[r * nm * nm + q * nm + p];
                if (scale && append) *o += sum * sc;
                else if (append) *o = *o + sum;
                else if (scale) *o = sum * sc;
                else *o = sum;
            }
        }
    }
}

/* PhysDerivTensor3DKernel: MatrixFreeOps/PhysDerivKernels.hpp:219-294. */
static void deriv_hex_elmt(int nq, const double *D, const double *in, double *d0,
                           double *d1, double *d2)
{
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < nq * nq; ++j) {
            double s = 0.0;
            for (int k = 0; k < nq; ++k) s += D[k * nq + i] * in[j * nq + k];
            d0[j * nq + i] = s;
        }
    for (int block = 0; block < nq; ++block) {
        int start = block * nq * nq;
        for (int i = 0; i < nq; ++i)
            for (int j = 0; j < nq; ++j) {
                double s = 0.0;
                for (int k = 0; k < nq; ++k) s += in[start + k * nq + i] * D[k * nq + j];
                d1[start + j * nq + i] = s;
            }
    }
    for (int i = 0; i < nq * nq; ++i)
        for (int j = 0; j < nq; ++j) {
            double s = 0.0;
            for (int k = 0; k < nq; ++k) s += in[k * nq * nq + i] * D[k * nq + j];
            d2[j * nq * nq + i] = s;
        }
}

void or_bwdtrans_hex(int nm, int nq, const double *bdata, int E, const double *in,
                     double *out)
{
    const int M = nm * nm * nm, Q = nq * nq * nq;
#pragma omp parallel
    {
        double *s1 = (double *)malloc(sizeof(double) * Q);
        double *s2 = (double *)malloc(sizeof(double) * Q);
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e)
            bwd_hex_elmt(nm, nq, bdata, in + (size_t)e * M, s1, s2, out + (size_t)e * Q);
        free(s1);
        free(s2);
    }
}

void or_iproduct_hex(int nm, int nq, const double *bdata, const double *w, int E,
                     int deformed, const double *jac, const double *in, double *out)
{
    const int M = nm * nm * nm, Q = nq * nq * nq;
#pragma omp parallel
    {
        double *s1 = (double *)malloc(sizeof(double) * nq * nq);
        double *s2 = (double *)malloc(sizeof(double) * nq);
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e) {
            const double *jp = deformed ? jac + (size_t)e * Q : jac + e;
            iprod_hex_elmt(nm, nq, bdata, bdata, bdata, w, deformed, jp,
                           in + (size_t)e * Q, s1, s2, out + (size_t)e * M, 0, 0, 1.0);
        }
        free(s1);
        free(s2);
    }
}

/* Physical derivatives d/dx_c = sum_a df[c*3+a] d/dxi_a, as PhysDerivHexKernel
 * (PhysDerivKernels.hpp:296-370) applies the factors after the tensor kernel. */
void or_physderiv_hex(int nq, const double *D, int E, int deformed, const double *df,
                      const double *in, double *out0, double *out1, double *out2)
{
    const int Q = nq * nq * nq;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
#pragma omp parallel
    {
        double *d[3];
        for (int a = 0; a < 3; ++a) d[a] = (double *)malloc(sizeof(double) * Q);
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e) {
            deriv_hex_elmt(nq, D, in + (size_t)e * Q, d[0], d[1], d[2]);
            double *outs[3] = {out0 + (size_t)e * Q, out1 + (size_t)e * Q, out2 + (size_t)e * Q};
            for (int q = 0; q < Q; ++q) {
                size_t idx = deformed ? (size_t)e * Q + q : (size_t)e;
                for (int c = 0; c < 3; ++c) {
                    double s = df[(c * 3 + 0) * np + idx] * d[0][q];
                    s += df[(c * 3 + 1) * np + idx] * d[1][q];
                    s += df[(c * 3 + 2) * np + idx] * d[2][q];
                    outs[c][q] = s;
                }
            }
        }
        for (int a = 0; a < 3; ++a) free(d[a]);
    }
}

/* IProductWRTDerivBaseHexImpl: MatrixFreeOps/IProductWRTDerivBase.h:1231-1343.
 * tmp_a = sum_c df[c*3+a] in_c, then IP(dB0,B1,B2; tmp0) + IP(B0,dB1,B2; tmp1)
 * + IP(B0,B1,dB2; tmp2), each IP weighted by J w (IProductHexKernel). */
void or_iprod_deriv_hex(int nm, int nq, const double *bdata, const double *dbdata,
                        const double *w, int E, int deformed, const double *jac,
                        const double *df, const double *in0, const double *in1,
                        const double *in2, double *out)
{
    const int M = nm * nm * nm, Q = nq * nq * nq;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
#pragma omp parallel
    {
        double *t[3], *wsp1 = (double *)malloc(sizeof(double) * Q);
        double *wsp2 = (double *)malloc(sizeof(double) * Q);
        for (int a = 0; a < 3; ++a) t[a] = (double *)malloc(sizeof(double) * Q);
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e) {
            const double *jp = deformed ? jac + (size_t)e * Q : jac + e;
            const size_t o = (size_t)e * Q;
            for (int q = 0; q < Q; ++q) {
                const size_t idx = deformed ? o + q : (size_t)e;
                for (int a = 0; a < 3; ++a)
                    t[a][q] = df[(0 * 3 + a) * np + idx] * in0[o + q] +
                              df[(1 * 3 + a) * np + idx] * in1[o + q] +
                              df[(2 * 3 + a) * np + idx] * in2[o + q];
            }
            double *op = out + (size_t)e * M;
            iprod_hex_elmt(nm, nq, dbdata, bdata, bdata, w, deformed, jp, t[0], wsp1, wsp2,
                           op, 0, 0, 1.0);
            iprod_hex_elmt(nm, nq, bdata, dbdata, bdata, w, deformed, jp, t[1], wsp1, wsp2,
                           op, 0, 1, 1.0);
            iprod_hex_elmt(nm, nq, bdata, bdata, dbdata, w, deformed, jp, t[2], wsp1, wsp2,
                           op, 0, 1, 1.0);
        }
        for (int a = 0; a < 3; ++a) free(t[a]);
        free(wsp1);
        free(wsp2);
    }
}

/* HelmholtzHexImpl: MatrixFreeOps/Helmholtz.h:764-996 (both DEFORMED branches). */
void or_helmholtz_hex(int nm, int nq, const double *bdata, const double *dbdata,
                      const double *D, const double *w, int E, int deformed,
                      const double *jac, const double *df, double lambda,
                      const double *in, double *out)
{
    const int M = nm * nm * nm, Q = nq * nq * nq;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
#pragma omp parallel
    {
        double *wsp1 = (double *)malloc(sizeof(double) * Q);
        double *wsp2 = (double *)malloc(sizeof(double) * Q);
        double *bwd = (double *)malloc(sizeof(double) * Q);
        double *d0 = (double *)malloc(sizeof(double) * Q);
        double *d1 = (double *)malloc(sizeof(double) * Q);
        double *d2 = (double *)malloc(sizeof(double) * Q);
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e) {
            const double *jp = deformed ? jac + (size_t)e * Q : jac + e;
            double *o = out + (size_t)e * M;
            bwd_hex_elmt(nm, nq, bdata, in + (size_t)e * M, wsp1, wsp2, bwd);
            iprod_hex_elmt(nm, nq, bdata, bdata, bdata, w, deformed, jp, bwd, wsp1, wsp2,
                           o, 1, 0, lambda);
            deriv_hex_elmt(nq, D, bwd, d0, d1, d2);
            for (int q = 0; q < Q; ++q) {
                size_t idx = deformed ? (size_t)e * Q + q : (size_t)e;
                double f[9];
                for (int a = 0; a < 9; ++a) f[a] = df[a * np + idx];
                double m00 = f[0] * f[0]; m00 += f[3] * f[3]; m00 += f[6] * f[6];
                double m01 = f[0] * f[1]; m01 += f[3] * f[4]; m01 += f[6] * f[7];
                double m02 = f[0] * f[2]; m02 += f[3] * f[5]; m02 += f[6] * f[8];
                double m11 = f[1] * f[1]; m11 += f[4] * f[4]; m11 += f[7] * f[7];
                double m12 = f[1] * f[2]; m12 += f[4] * f[5]; m12 += f[7] * f[8];
                double m22 = f[2] * f[2]; m22 += f[5] * f[5]; m22 += f[8] * f[8];
                double a0 = d0[q], a1 = d1[q], a2 = d2[q];
                double t = m00 * a0; t += m01 * a1; t += m02 * a2; d0[q] = t;
                t = m01 * a0; t += m11 * a1; t += m12 * a2; d1[q] = t;
                t = m02 * a0; t += m12 * a1; t += m22 * a2; d2[q] = t;
            }
            iprod_hex_elmt(nm, nq, dbdata, bdata, bdata, w, deformed, jp, d0, wsp1, wsp2,
                           o, 0, 1, 1.0);
            iprod_hex_elmt(nm, nq, bdata, dbdata, bdata, w, deformed, jp, d1, wsp1, wsp2,
                           o, 0, 1, 1.0);
            iprod_hex_elmt(nm, nq, bdata, bdata, dbdata, w, deformed, jp, d2, wsp1, wsp2,
                           o, 0, 1, 1.0);
        }
        free(wsp1); free(wsp2); free(bwd); free(d0); free(d1); free(d2);
    }
}

/* ------------------------------------------------------------------------- */
/* Quad kernels (width 1)                                                      */
/* ------------------------------------------------------------------------- */

/* BwdTransQuadKernel: BwdTransKernels.hpp:35-76 */
static void bwd_quad_elmt(int nm, int nq, const double *B, const double *in,
                          double *wsp, double *out)
{
    for (int i = 0, cnt_iq = 0; i < nq; ++i)
        for (int q = 0, cnt_pq = 0; q < nm; ++q, ++cnt_iq) {
            double tmp = in[cnt_pq] * B[i];
            ++cnt_pq;
            for (int p = 1; p < nm; ++p, ++cnt_pq) tmp += in[cnt_pq] * B[p * nq + i];
            wsp[cnt_iq] = tmp;
        }
    for (int j = 0, cnt_ij = 0; j < nq; ++j)
        for (int i = 0, cnt_iq = 0; i < nq; ++i, ++cnt_ij) {
            double tmp = wsp[cnt_iq] * B[j];
            ++cnt_iq;
            for (int q = 1; q < nm; ++q, ++cnt_iq) tmp += wsp[cnt_iq] * B[q * nq + j];
            out[cnt_ij] = tmp;
        }
}

/* IProductQuadKernel: IProductKernels.hpp:76-133 (stores out[q*nm1+p]) */
static void iprod_quad_elmt(int nm, int nq, const double *B0, const double *B1,
                            const double *w, int deformed, const double *jac,
                            const double *in, double *sums_j, double *out, int scale,
                            int append, double sc)
{
    for (int p = 0; p < nm; ++p) {
        int cnt_ji = 0;
        for (int j = 0; j < nq; ++j) {
            double sum_j = 0.0;
            for (int i = 0; i < nq; ++i, ++cnt_ji) {
                double jv = deformed ? jac[j * nq + i] : jac[0];
                double prod = in[cnt_ji] * B0[p * nq + i] * jv;
                sum_j += prod * w[i];
            }
            sums_j[j] = sum_j;
        }
        for (int q = 0; q < nm; ++q) {
            double sum = 0.0;
            for (int j = 0; j < nq; ++j) {
                double prod = sums_j[j] * B1[q * nq + j];
                sum += prod * w[j];
            }
            double *o = &out[q * nm + p];
            if (scale && append) *o += sum * sc;
            else if (append) *o = *o + sum;
            else if (scale) *o = sum * sc;
            else *o = sum;
        }
    }
}

/* PhysDerivTensor2DKernel: PhysDerivKernels.hpp:39-91 */
static void deriv_quad_elmt(int nq, const double *D, const double *in, double *d0,
                            double *d1)
{
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < nq; ++j) {
            double s = 0.0;
            for (int k = 0; k < nq; ++k) s += D[k * nq + i] * in[j * nq + k];
            d0[j * nq + i] = s;
        }
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < nq; ++j) {
            double s = 0.0;
            for (int k = 0; k < nq; ++k) s += in[k * nq + i] * D[k * nq + j];
            d1[j * nq + i] = s;
        }
}

void or_bwdtrans_quad(int nm, int nq, const double *bdata, int E, const double *in,
                      double *out)
{
    const int M = nm * nm, Q = nq * nq;
#pragma omp parallel
    {
        double *wsp = (double *)malloc(sizeof(double) * nq * nm);
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e)
            bwd_quad_elmt(nm, nq, bdata, in + (size_t)e * M, wsp, out + (size_t)e * Q);
        free(wsp);
    }
}

void or_iproduct_quad(int nm, int nq, const double *bdata, const double *w, int E,
                      int deformed, const double *jac, const double *in, double *out)
{
    const int M = nm * nm, Q = nq * nq;
#pragma omp parallel
    {
        double *sj = (double *)malloc(sizeof(double) * nq);
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e) {
            const double *jp = deformed ? jac + (size_t)e * Q : jac + e;
            iprod_quad_elmt(nm, nq, bdata, bdata, w, deformed, jp, in + (size_t)e * Q, sj,
                            out + (size_t)e * M, 0, 0, 1.0);
        }
        free(sj);
    }
}

void or_physderiv_quad(int nq, const double *D, int E, int deformed, const double *df,
                       const double *in, double *out0, double *out1)
{
    const int Q = nq * nq;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
#pragma omp parallel
    {
        double *d0 = (double *)malloc(sizeof(double) * Q);
        double *d1 = (double *)malloc(sizeof(double) * Q);
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e) {
            deriv_quad_elmt(nq, D, in + (size_t)e * Q, d0, d1);
            for (int q = 0; q < Q; ++q) {
                size_t idx = deformed ? (size_t)e * Q + q : (size_t)e;
                out0[(size_t)e * Q + q] = df[0 * np + idx] * d0[q] + df[1 * np + idx] * d1[q];
                out1[(size_t)e * Q + q] = df[2 * np + idx] * d0[q] + df[3 * np + idx] * d1[q];
            }
        }
        free(d0);
        free(d1);
    }
}

/* IProductWRTDerivBaseQuad (2D coordinates): MatrixFreeOps/IProductWRTDerivBase.h:371-
 * (tmp0 = df0 in0 + df2 in1, tmp1 = df1 in0 + df3 in1; IP(dB0,B1; tmp0) + IP(B0,dB1; tmp1)). */
void or_iprod_deriv_quad(int nm, int nq, const double *bdata, const double *dbdata,
                         const double *w, int E, int deformed, const double *jac,
                         const double *df, const double *in0, const double *in1,
                         double *out)
{
    const int M = nm * nm, Q = nq * nq;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
#pragma omp parallel
    {
        double *t0 = (double *)malloc(sizeof(double) * Q);
        double *t1 = (double *)malloc(sizeof(double) * Q);
        double *sj = (double *)malloc(sizeof(double) * nq);
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e) {
            const double *jp = deformed ? jac + (size_t)e * Q : jac + e;
            const size_t o = (size_t)e * Q;
            for (int q = 0; q < Q; ++q) {
                const size_t idx = deformed ? o + q : (size_t)e;
                t0[q] = df[0 * np + idx] * in0[o + q] + df[2 * np + idx] * in1[o + q];
                t1[q] = df[1 * np + idx] * in0[o + q] + df[3 * np + idx] * in1[o + q];
            }
            double *op = out + (size_t)e * M;
            iprod_quad_elmt(nm, nq, dbdata, bdata, w, deformed, jp, t0, sj, op, 0, 0, 1.0);
            iprod_quad_elmt(nm, nq, bdata, dbdata, w, deformed, jp, t1, sj, op, 0, 1, 1.0);
        }
        free(t0);
        free(t1);
        free(sj);
    }
}

/* HelmholtzQuadImpl: MatrixFreeOps/Helmholtz.h:138-276 */
void or_helmholtz_quad(int nm, int nq, const double *bdata, const double *dbdata,
                       const double *D, const double *w, int E, int deformed,
                       const double *jac, const double *df, double lambda,
                       const double *in, double *out)
{
    const int M = nm * nm, Q = nq * nq;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
#pragma omp parallel
    {
        int wspn = nq > nq * nm ? nq : nq * nm;
        double *wsp = (double *)malloc(sizeof(double) * wspn);
        double *bwd = (double *)malloc(sizeof(double) * Q);
        double *d0 = (double *)malloc(sizeof(double) * Q);
        double *d1 = (double *)malloc(sizeof(double) * Q);
#pragma omp for schedule(static)
        for (int e = 0; e < E; ++e) {
            const double *jp = deformed ? jac + (size_t)e * Q : jac + e;
            double *o = out + (size_t)e * M;
            bwd_quad_elmt(nm, nq, bdata, in + (size_t)e * M, wsp, bwd);
            iprod_quad_elmt(nm, nq, bdata, bdata, w, deformed, jp, bwd, wsp, o, 1, 0, lambda);
            deriv_quad_elmt(nq, D, bwd, d0, d1);
            for (int q = 0; q < Q; ++q) {
                size_t idx = deformed ? (size_t)e * Q + q : (size_t)e;
                double f0 = df[0 * np + idx], f1 = df[1 * np + idx];
                double f2 = df[2 * np + idx], f3 = df[3 * np + idx];
                double m00 = f0 * f0; m00 += f2 * f2;
                double m01 = f0 * f1; m01 += f2 * f3;
                double m11 = f1 * f1; m11 += f3 * f3;
                double a0 = d0[q], a1 = d1[q];
                double t = m00 * a0; t += m01 * a1; bwd[q] = t;
                t = m01 * a0; t += m11 * a1; d0[q] = t;
            }
            iprod_quad_elmt(nm, nq, dbdata, bdata, w, deformed, jp, bwd, wsp, o, 0, 1, 1.0);
            iprod_quad_elmt(nm, nq, bdata, dbdata, w, deformed, jp, d0, wsp, o, 0, 1, 1.0);
        }
        free(wsp); free(bwd); free(d0); free(d1);
    }
}
