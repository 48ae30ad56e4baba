// Prism operators in collapsed coordinates (FP64, runtime sizes).
//
// Restates the reference's MatrixFree prism kernels: BwdTransPrismKernel
// (MatrixFreeOps/BwdTransKernels.hpp:224-300), IProductPrismKernel (IProductKernels.hpp:
// 315-449), PhysDerivPrismKernel (PhysDerivKernels.hpp:219-294, :377-460) and
// HelmholtzPrismImpl (Helmholtz.h:1291-1460).  The expansion is Modified_A along eta0 and
// xi1 (GLL points) times Modified_B along eta2 (Gauss-Radau, alpha = 1), collapsing eta0 on
// the top edge xi2 = 1: mode (p, q, r) with p, q < nm, r < nm - p, stored at
// nm * sum_{p'<p} (nm - p') + q (nm - p) + r.  The modes (0, q, 1) run along the collapsed
// edge and also carry phi_1(eta0) (the CORRECT terms).  The weights along eta2 include the
// collapse factor 1/2 (Helper, Operator.hpp:244-258); the metric of the Laplacian is
// formed per point from the nine derivative factors with h0 = (1 + eta0)/2,
// h1 = 2/(1 - eta2).
//
// Layout on the GPU: one workgroup per element at a time (a grid-stride loop over the
// elements), the basis tables staged once per workgroup in LDS and the element's
// coefficients, point arrays and sum-factorisation intermediates in LDS.  Every phase is
// a loop of the workgroup's threads over the phase's independent outputs:
//   BwdTrans   t1[p][q][k] = sum_r B2 c     (the collapsed-edge modes added into p = 1),
//              t2[p][j][k] = sum_q B1 t1,   u[k][j][i] = sum_p B0 t2
//   IProduct   s1[p][j][k] = sum_i B0 g,    s2[p][q][k] = sum_j B1 s1,
//              c(p, q, r) = sum_k B2 s2     (+ sum_k B2_1 s2[1][q][k] for r = 1, p = 0)
// with g the point values times w0 w1 w2 J.
#include <hip/hip_runtime.h>

#include "imf_common.hpp"
#include "mf_prism.hpp"

namespace imf {

namespace {

struct PrismTab {
    const double *B0, *dB0, *D0, *w0, *z0, *B1, *dB1, *D1, *w1, *z1, *B2, *dB2, *D2, *w2, *z2;
};

__device__ inline PrismTab prism_tab(const double *t, int nm, int nq0, int nq1, int nq2)
{
    PrismTab r;
    r.B0 = t;
    r.dB0 = r.B0 + nm * nq0;
    r.D0 = r.dB0 + nm * nq0;
    r.w0 = r.D0 + nq0 * nq0;
    r.z0 = r.w0 + nq0;
    r.B1 = r.z0 + nq0;
    r.dB1 = r.B1 + nm * nq1;
    r.D1 = r.dB1 + nm * nq1;
    r.w1 = r.D1 + nq1 * nq1;
    r.z1 = r.w1 + nq1;
    r.B2 = r.z1 + nq1;
    r.dB2 = r.B2 + tri_modes(nm) * nq2;
    r.D2 = r.dB2 + tri_modes(nm) * nq2;
    r.w2 = r.D2 + nq2 * nq2;
    r.z2 = r.w2 + nq2;
    return r;
}

// first (p, r) index of group p in the packed Modified_B table
__device__ inline int pr_start(int p, int nm) { return p * nm - p * (p - 1) / 2; }

struct Dims {
    int nm, nq0, nq1, nq2;
};

// u = BwdTrans(c) through t1 (nm*nm*nq2) and t2 (nm*nq1*nq2).
__device__ void prism_bwd(const PrismTab &T, Dims d, const double *c, double *t1, double *t2,
                          double *u, int tid, int nt)
{
    const int nm = d.nm, nq0 = d.nq0, nq1 = d.nq1, nq2 = d.nq2;
    for (int t = tid; t < nm * nm * nq2; t += nt) {
        const int k = t % nq2, pq = t / nq2, q = pq % nm, p = pq / nm;
        const int ps = pr_start(p, nm), m0 = nm * ps + q * (nm - p);
        double s = 0.0;
        for (int r = 0; r < nm - p; ++r) s = fma(T.B2[(ps + r) * nq2 + k], c[m0 + r], s);
        // mode (0, q, 1) = phi_0(eta0) phi_q(xi1) psi_01(eta2) + phi_1(eta0) phi_q psi_01
        if (p == 1) s = fma(T.B2[nq2 + k], c[q * nm + 1], s);
        t1[t] = s;
    }
    __syncthreads();
    for (int t = tid; t < nm * nq1 * nq2; t += nt) {
        const int k = t % nq2, pj = t / nq2, j = pj % nq1, p = pj / nq1;
        double s = 0.0;
        for (int q = 0; q < nm; ++q) s = fma(T.B1[q * nq1 + j], t1[(p * nm + q) * nq2 + k], s);
        t2[t] = s;
    }
    __syncthreads();
    for (int t = tid; t < nq0 * nq1 * nq2; t += nt) {
        const int i = t % nq0, kj = t / nq0, j = kj % nq1, k = kj / nq1;
        double s = 0.0;
        for (int p = 0; p < nm; ++p) s = fma(T.B0[p * nq0 + i], t2[(p * nq1 + j) * nq2 + k], s);
        u[t] = s;
    }
    __syncthreads();
}

// Inner products of the (already weighted) point array g with the tables b0 (eta0), b1
// (xi1) and b2 (eta2) through s1 (nm*nq1*nq2) and s2 (nm*nm*nq2).  Stores out = s * sc
// (scale), out += s (append) or out = s.
__device__ void prism_iprod(Dims d, const double *b0, const double *b1, const double *b2,
                            const double *g, double *s1, double *s2, double *out, bool scale,
                            bool append, double sc, int tid, int nt)
{
    const int nm = d.nm, nq0 = d.nq0, nq1 = d.nq1, nq2 = d.nq2;
    for (int t = tid; t < nm * nq1 * nq2; t += nt) {
        const int k = t % nq2, pj = t / nq2, j = pj % nq1, p = pj / nq1;
        const double *gr = g + (k * nq1 + j) * nq0;
        double s = 0.0;
        for (int i = 0; i < nq0; ++i) s = fma(b0[p * nq0 + i], gr[i], s);
        s1[t] = s;
    }
    __syncthreads();
    for (int t = tid; t < nm * nm * nq2; t += nt) {
        const int k = t % nq2, pq = t / nq2, q = pq % nm, p = pq / nm;
        double s = 0.0;
        for (int j = 0; j < nq1; ++j) s = fma(b1[q * nq1 + j], s1[(p * nq1 + j) * nq2 + k], s);
        s2[t] = s;
    }
    __syncthreads();
    for (int t = tid; t < prism_modes(nm); t += nt) {
        int p = 0;
        while (nm * pr_start(p + 1, nm) <= t) ++p;
        const int rem = t - nm * pr_start(p, nm), q = rem / (nm - p), r = rem - q * (nm - p);
        const double *row = b2 + (pr_start(p, nm) + r) * nq2, *col = s2 + (p * nm + q) * nq2;
        double s = 0.0;
        for (int k = 0; k < nq2; ++k) s = fma(row[k], col[k], s);
        if (p == 0 && r == 1) {  // the phi_1(eta0) part of the collapsed-edge mode
            const double *c1 = s2 + (nm + q) * nq2;
            for (int k = 0; k < nq2; ++k) s = fma(b2[nq2 + k], c1[k], s);
        }
        out[t] = scale && append ? fma(s, sc, out[t]) : append ? out[t] + s : scale ? s * sc : s;
    }
    __syncthreads();
}

// Reference-space derivatives along eta0, xi1 and eta2 (PhysDerivTensor3DKernel).
__device__ void prism_deriv(const PrismTab &T, Dims d, const double *u, double *d0, double *d1,
                            double *d2, int tid, int nt)
{
    const int nq0 = d.nq0, nq1 = d.nq1, nq2 = d.nq2, s2 = nq0 * nq1;
    for (int t = tid; t < nq0 * nq1 * nq2; t += nt) {
        const int i = t % nq0, kj = t / nq0, j = kj % nq1, k = kj / nq1;
        const double *ui = u + (k * nq1 + j) * nq0, *uj = u + k * s2 + i, *uk = u + j * nq0 + i;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int a = 0; a < nq0; ++a) a0 = fma(T.D0[a * nq0 + i], ui[a], a0);
        for (int a = 0; a < nq1; ++a) a1 = fma(T.D1[a * nq1 + j], uj[a * nq0], a1);
        for (int a = 0; a < nq2; ++a) a2 = fma(T.D2[a * nq2 + k], uk[a * s2], a2);
        d0[t] = a0;
        d1[t] = a1;
        d2[t] = a2;
    }
    __syncthreads();
}

__device__ inline double point_weight(const PrismTab &T, Dims d, int t)
{
    const int i = t % d.nq0, kj = t / d.nq0, j = kj % d.nq1, k = kj / d.nq1;
    return T.w0[i] * T.w1[j] * T.w2[k];
}

__global__ __launch_bounds__(256) void prism_kernel(PrismArgs a)
{
    extern __shared__ double sm[];
    const Dims d{a.nm, a.nq0, a.nq1, a.nq2};
    const int nm = a.nm, M = prism_modes(nm), Q = a.nq0 * a.nq1 * a.nq2;
    const int tl = prism_table_len(nm, a.nq0, a.nq1, a.nq2);
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int t = tid; t < tl; t += nt) sm[t] = a.tab[t];
    double *cin = sm + tl, *cout = cin + M, *u = cout + M, *d0 = u + Q, *d1 = d0 + Q,
           *d2 = d1 + Q, *sa = d2 + Q, *sb = sa + nm * nm * a.nq2;
    const bool deformed = a.geo == kDeformed;
    const PrismTab T = prism_tab(sm, nm, a.nq0, a.nq1, a.nq2);
    const bool coef_in = (a.kind == kBwdTrans && !a.transpose) ||
                         (a.kind == kIProduct && a.transpose) || a.kind == kHelmholtz;
    const long np = deformed ? (long)a.nElmt * Q : a.nElmt;
    __syncthreads();

    for (long e = blockIdx.x; e < a.nElmt; e += gridDim.x) {
        const double *jac = a.jac ? a.jac + (deformed ? e * Q : e) : nullptr;
        if (coef_in)
            for (int t = tid; t < M; t += nt) cin[t] = a.in[e * M + t];
        else if (a.kind != kIProdDeriv)
            for (int t = tid; t < Q; t += nt) u[t] = a.in[e * Q + t];
        __syncthreads();

        if (a.kind == kBwdTrans && !a.transpose) {
            prism_bwd(T, d, cin, sa, sb, u, tid, nt);
            for (int t = tid; t < Q; t += nt) a.out[e * Q + t] = u[t];
        } else if (a.kind == kIProduct && a.transpose) {
            // the transpose of IProduct = B^T W J is W J B
            prism_bwd(T, d, cin, sa, sb, u, tid, nt);
            for (int t = tid; t < Q; t += nt)
                a.out[e * Q + t] = u[t] * point_weight(T, d, t) * (deformed ? jac[t] : jac[0]);
        } else if (a.kind == kBwdTrans || a.kind == kIProduct) {
            if (a.kind == kIProduct) {  // BwdTrans^T: unit weights, no Jacobian
                for (int t = tid; t < Q; t += nt)
                    u[t] = u[t] * point_weight(T, d, t) * (deformed ? jac[t] : jac[0]);
                __syncthreads();
            }
            prism_iprod(d, T.B0, T.B1, T.B2, u, sb, sa, cout, false, false, 1.0, tid, nt);
            for (int t = tid; t < M; t += nt) a.out[e * M + t] = cout[t];
        } else if (a.kind == kIProdDeriv) {
            // IProductWRTDerivBase (IProductWRTDerivBase.h:1630-): per point t_a = sum_c df[3c+a] in_c,
            // the transpose of PhysDeriv's collapsed chain rule, the weights w J, then the
            // inner products with (dB0, B1, B2), (B0, dB1, B2) and (B0, B1, dB2)
            const long nQ = (long)a.nElmt * Q;
            for (int t = tid; t < Q; t += nt) {
                const long g = deformed ? e * Q + t : e;
                const double x0 = a.in[e * Q + t], x1 = a.in[nQ + e * Q + t],
                             x2 = a.in[2 * nQ + e * Q + t];
                double v[3];
                for (int c = 0; c < 3; ++c) {
                    double o = x0 * a.df[c * np + g];
                    o = fma(x1, a.df[(3 + c) * np + g], o);
                    v[c] = fma(x2, a.df[(6 + c) * np + g], o);
                }
                double v0 = v[0], v1 = v[1];
                const double v2 = v[2];
                // adjoint of (h1 D0, D1, h0 h1 D0 + D2): (h1 (t0 + h0 t2), t1, t2)
                const int i = t % d.nq0, k = t / (d.nq0 * d.nq1);
                v0 = (2.0 / (1.0 - T.z2[k])) * fma(0.5 * (1.0 + T.z0[i]), v2, v0);
                const double wj = point_weight(T, d, t) * (deformed ? jac[t] : jac[0]);
                d0[t] = wj * v0;
                d1[t] = wj * v1;
                d2[t] = wj * v2;
            }
            __syncthreads();
            prism_iprod(d, T.dB0, T.B1, T.B2, d0, sb, sa, cout, false, false, 1.0, tid, nt);
            prism_iprod(d, T.B0, T.dB1, T.B2, d1, sb, sa, cout, false, true, 1.0, tid, nt);
            prism_iprod(d, T.B0, T.B1, T.dB2, d2, sb, sa, cout, false, true, 1.0, tid, nt);
            for (int t = tid; t < M; t += nt) a.out[e * M + t] = cout[t];
        } else if (a.kind == kPhysDeriv) {
            prism_deriv(T, d, u, d0, d1, d2, tid, nt);
            const long nQ = (long)a.nElmt * Q;
            for (int t = tid; t < Q; t += nt) {
                const int i = t % d.nq0, k = t / (d.nq0 * d.nq1);
                const long g = deformed ? e * Q + t : e;
                const double a0 = d0[t] * (2.0 / (1.0 - T.z2[k]));
                const double a2 = fma(0.5 * (1.0 + T.z0[i]), a0, d2[t]);
                const double a1 = d1[t];
                for (int c = 0; c < 3; ++c) {
                    double o = a0 * a.df[(3 * c) * np + g];
                    o = fma(a1, a.df[(3 * c + 1) * np + g], o);
                    o = fma(a2, a.df[(3 * c + 2) * np + g], o);
                    a.out[c * nQ + e * Q + t] = o;
                }
            }
        } else {  // Helmholtz
            prism_bwd(T, d, cin, sa, sb, u, tid, nt);
            prism_deriv(T, d, u, d0, d1, d2, tid, nt);
            for (int t = tid; t < Q; t += nt) {
                const int i = t % d.nq0, k = t / (d.nq0 * d.nq1);
                const long g = deformed ? e * Q + t : e;
                const double wj = point_weight(T, d, t) * (deformed ? jac[t] : jac[0]);
                u[t] *= wj;
                double f[9];
                for (int c = 0; c < 9; ++c) f[c] = a.df[c * np + g];
                const double h1 = 2.0 / (1.0 - T.z2[k]), h0 = 0.5 * (1.0 + T.z0[i]);
                // gradient coefficients of (d0, d1, d2) per physical direction c:
                // (h1 (f[3c] + h0 f[3c+2]), f[3c+1], f[3c+2])
                const double t0 = h1 * fma(h0, f[2], f[0]), t1 = h1 * fma(h0, f[5], f[3]),
                             t2 = h1 * fma(h0, f[8], f[6]);
                const double g00 = fma(t2, t2, fma(t1, t1, t0 * t0));
                const double g01 = fma(f[7], t2, fma(f[4], t1, f[1] * t0));
                const double g02 = fma(f[8], t2, fma(f[5], t1, f[2] * t0));
                const double g11 = fma(f[7], f[7], fma(f[4], f[4], f[1] * f[1]));
                const double g22 = fma(f[8], f[8], fma(f[5], f[5], f[2] * f[2]));
                const double g12 = fma(f[7], f[8], fma(f[4], f[5], f[1] * f[2]));
                const double a0 = d0[t], a1 = d1[t], a2 = d2[t];
                d0[t] = wj * fma(g02, a2, fma(g01, a1, g00 * a0));
                d1[t] = wj * fma(g12, a2, fma(g11, a1, g01 * a0));
                d2[t] = wj * fma(g22, a2, fma(g12, a1, g02 * a0));
            }
            __syncthreads();
            prism_iprod(d, T.B0, T.B1, T.B2, u, sb, sa, cout, true, false, a.lambda, tid, nt);
            prism_iprod(d, T.dB0, T.B1, T.B2, d0, sb, sa, cout, false, true, 1.0, tid, nt);
            prism_iprod(d, T.B0, T.dB1, T.B2, d1, sb, sa, cout, false, true, 1.0, tid, nt);
            prism_iprod(d, T.B0, T.B1, T.dB2, d2, sb, sa, cout, false, true, 1.0, tid, nt);
            for (int t = tid; t < M; t += nt) a.out[e * M + t] = cout[t];
        }
        __syncthreads();  // the next element overwrites the arrays
    }
}

}  // namespace

hipError_t launch_prism(const PrismArgs &a, hipStream_t s)
{
    const size_t lds = sizeof(double) * prism_lds_len(a.nm, a.nq0, a.nq1, a.nq2);
    hipError_t e = hipFuncSetAttribute((const void *)&prism_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const int Q = a.nq0 * a.nq1 * a.nq2;
    const int nt = Q <= 64 ? 64 : Q <= 128 ? 128 : 256;
    // enough workgroups to fill every CU several times over, each looping over elements
    const unsigned grid = (unsigned)(a.nElmt < 8192 ? a.nElmt : 8192);
    prism_kernel<<<grid, nt, lds, s>>>(a);
    return hipGetLastError();
}

}  // namespace imf
