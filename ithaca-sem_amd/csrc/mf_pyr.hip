// Pyramid operators in collapsed coordinates (FP64, runtime sizes).
//
// Restates the reference's MatrixFree pyramid kernels: BwdTransPyrKernel
// (MatrixFreeOps/BwdTransKernels.hpp:131-222), IProductPyrKernel (IProductKernels.hpp:
// 455-598), PhysDerivPyrKernel (PhysDerivKernels.hpp:464-550) and HelmholtzPyrImpl
// (Helmholtz.h:1478-1510, :1771-1935).  The expansion is Modified_A along eta0 and eta1
// (GLL points) times ModifiedPyr_C along eta2 (Gauss-Radau, alpha = 2), collapsing the top
// face to the apex: mode (p, q, r) with p, q < nm and r < nm - max(p, q), r fastest, then
// q, then p; B2 has a row per mode.  The top vertex (mode 1 = (0, 0, 1)) also carries
// B2_1 (B0_0 B1_1 + B0_1 B1_0 + B0_1 B1_1) (the CORRECT term).  The weights along eta2
// include the collapse factor 1/4 (Helper, Operator.hpp:244-258); the metric of the
// Laplacian is formed per point from the nine derivative factors with h0 = (1 + eta0)/2,
// h1 = (1 + eta1)/2, h2 = 2/(1 - eta2).
//
// Layout on the GPU: one workgroup per element at a time (a grid-stride loop over the
// elements), the tables staged once per workgroup in LDS, the element's arrays in LDS.
// Every phase is a loop of the workgroup's threads over the phase's independent outputs:
//   BwdTrans   t1[p][q][k] = sum_r B2 c  (+ c_1 B2_1 for (p, q) = (0,1), (1,0), (1,1)),
//              t2[p][j][k] = sum_q B1 t1,   u[k][j][i] = sum_p B0 t2
//   IProduct   s1[p][j][k] = sum_i B0 g,    s2[p][q][k] = sum_j B1 s1,
//              c(p, q, r) = sum_k B2 s2     (+ sum_k B2_1 (s2[0][1] + s2[1][0] + s2[1][1])
//              for mode 1)
// with g the point values times w0 w1 w2 J.
#include <hip/hip_runtime.h>

#include "imf_common.hpp"
#include "mf_pyr.hpp"

namespace imf {

namespace {

struct PyrTab {
    const double *B0, *dB0, *D0, *w0, *z0, *B1, *dB1, *D1, *w1, *z1, *B2, *dB2, *D2, *w2, *z2;
};

__device__ inline PyrTab pyr_tab(const double *t, int nm, int nq0, int nq1, int nq2)
{
    PyrTab r;
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
    r.dB2 = r.B2 + pyr_modes(nm) * nq2;
    r.D2 = r.dB2 + pyr_modes(nm) * nq2;
    r.w2 = r.D2 + nq2 * nq2;
    r.z2 = r.w2 + nq2;
    return r;
}

struct Dims {
    int nm, nq0, nq1, nq2;
};

__device__ inline int imax(int a, int b) { return a > b ? a : b; }

// first mode of (p, q): every (p', q') before it holds nm - max(p', q') modes
__device__ inline int pq_mode0(int p, int q, int nm)
{
    int m = 0;
    for (int pp = 0; pp < p; ++pp)
        for (int qq = 0; qq < nm; ++qq) m += nm - imax(pp, qq);
    for (int qq = 0; qq < q; ++qq) m += nm - imax(p, qq);
    return m;
}

// u = BwdTrans(c) through t1 (nm*nm*nq2) and t2 (nm*nq1*nq2).
__device__ void pyr_bwd(const PyrTab &T, Dims d, const double *c, double *t1, double *t2,
                        double *u, int tid, int nt)
{
    const int nm = d.nm, nq0 = d.nq0, nq1 = d.nq1, nq2 = d.nq2;
    for (int t = tid; t < nm * nm * nq2; t += nt) {
        const int k = t % nq2, pq = t / nq2, q = pq % nm, p = pq / nm;
        const int m0 = pq_mode0(p, q, nm);
        double s = 0.0;
        for (int r = 0; r < nm - imax(p, q); ++r) s = fma(T.B2[(m0 + r) * nq2 + k], c[m0 + r], s);
        // the top vertex's B0_0 B1_1, B0_1 B1_0 and B0_1 B1_1 parts
        if (p + q > 0 && p <= 1 && q <= 1) s = fma(c[1], T.B2[nq2 + k], s);
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
// (eta1) and b2 (eta2) through s1 (nm*nq1*nq2) and s2 (nm*nm*nq2).  Stores out = s * sc
// (scale), out += s (append) or out = s.
__device__ void pyr_iprod(Dims d, const double *b0, const double *b1, const double *b2,
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
    for (int t = tid; t < pyr_modes(nm); t += nt) {
        int p = 0, q = 0, m0 = 0;
        for (;;) {  // the (p, q) holding mode t
            const int n = nm - imax(p, q);
            if (t < m0 + n) break;
            m0 += n;
            if (++q == nm) {
                q = 0;
                ++p;
            }
        }
        const double *row = b2 + t * nq2, *col = s2 + (p * nm + q) * nq2;
        double s = 0.0;
        for (int k = 0; k < nq2; ++k) s = fma(row[k], col[k], s);
        if (t == 1) {  // top vertex: (p, q) = (0,1), (1,0), (1,1)
            const double *c01 = s2 + nq2, *c10 = s2 + nm * nq2, *c11 = s2 + (nm + 1) * nq2;
            for (int k = 0; k < nq2; ++k) s = fma(b2[nq2 + k], c01[k] + c10[k] + c11[k], s);
        }
        out[t] = scale && append ? fma(s, sc, out[t]) : append ? out[t] + s : scale ? s * sc : s;
    }
    __syncthreads();
}

// Reference-space derivatives along eta0, eta1 and eta2 (PhysDerivTensor3DKernel).
__device__ void pyr_deriv(const PyrTab &T, Dims d, const double *u, double *d0, double *d1,
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

__device__ inline double point_weight(const PyrTab &T, Dims d, int t)
{
    const int i = t % d.nq0, kj = t / d.nq0, j = kj % d.nq1, k = kj / d.nq1;
    return T.w0[i] * T.w1[j] * T.w2[k];
}

__global__ __launch_bounds__(256) void pyr_kernel(PyrArgs a)
{
    extern __shared__ double sm[];
    const Dims d{a.nm, a.nq0, a.nq1, a.nq2};
    const int nm = a.nm, M = pyr_modes(nm), Q = a.nq0 * a.nq1 * a.nq2;
    const int tl = pyr_table_len(nm, a.nq0, a.nq1, a.nq2);
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int t = tid; t < tl; t += nt) sm[t] = a.tab[t];
    double *cin = sm + tl, *cout = cin + M, *u = cout + M, *d0 = u + Q, *d1 = d0 + Q,
           *d2 = d1 + Q, *sa = d2 + Q, *sb = sa + nm * nm * a.nq2;
    const bool deformed = a.geo == kDeformed;
    const PyrTab T = pyr_tab(sm, nm, a.nq0, a.nq1, a.nq2);
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
            pyr_bwd(T, d, cin, sa, sb, u, tid, nt);
            for (int t = tid; t < Q; t += nt) a.out[e * Q + t] = u[t];
        } else if (a.kind == kIProduct && a.transpose) {
            // the transpose of IProduct = B^T W J is W J B
            pyr_bwd(T, d, cin, sa, sb, u, tid, nt);
            for (int t = tid; t < Q; t += nt)
                a.out[e * Q + t] = u[t] * point_weight(T, d, t) * (deformed ? jac[t] : jac[0]);
        } else if (a.kind == kBwdTrans || a.kind == kIProduct) {
            if (a.kind == kIProduct) {  // BwdTrans^T: unit weights, no Jacobian
                for (int t = tid; t < Q; t += nt)
                    u[t] = u[t] * point_weight(T, d, t) * (deformed ? jac[t] : jac[0]);
                __syncthreads();
            }
            pyr_iprod(d, T.B0, T.B1, T.B2, u, sb, sa, cout, false, false, 1.0, tid, nt);
            for (int t = tid; t < M; t += nt) a.out[e * M + t] = cout[t];
        } else if (a.kind == kIProdDeriv) {
            // IProductWRTDerivBase (IProductWRTDerivBase.h:2484-): per point t_a = sum_c df[3c+a] in_c,
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
                // adjoint of (h2 D0, h2 D1, h0 h2 D0 + h1 h2 D1 + D2):
                // (h2 (t0 + h0 t2), h2 (t1 + h1 t2), t2)
                const int i = t % d.nq0, kj = t / d.nq0, j = kj % d.nq1, k = kj / d.nq1;
                const double h2 = 2.0 / (1.0 - T.z2[k]);
                v0 = h2 * fma(0.5 * (1.0 + T.z0[i]), v2, v0);
                v1 = h2 * fma(0.5 * (1.0 + T.z1[j]), v2, v1);
                const double wj = point_weight(T, d, t) * (deformed ? jac[t] : jac[0]);
                d0[t] = wj * v0;
                d1[t] = wj * v1;
                d2[t] = wj * v2;
            }
            __syncthreads();
            pyr_iprod(d, T.dB0, T.B1, T.B2, d0, sb, sa, cout, false, false, 1.0, tid, nt);
            pyr_iprod(d, T.B0, T.dB1, T.B2, d1, sb, sa, cout, false, true, 1.0, tid, nt);
            pyr_iprod(d, T.B0, T.B1, T.dB2, d2, sb, sa, cout, false, true, 1.0, tid, nt);
            for (int t = tid; t < M; t += nt) a.out[e * M + t] = cout[t];
        } else if (a.kind == kPhysDeriv) {
            pyr_deriv(T, d, u, d0, d1, d2, tid, nt);
            const long nQ = (long)a.nElmt * Q;
            for (int t = tid; t < Q; t += nt) {
                const int i = t % d.nq0, kj = t / d.nq0, j = kj % d.nq1, k = kj / d.nq1;
                const long g = deformed ? e * Q + t : e;
                const double h2 = 2.0 / (1.0 - T.z2[k]);
                const double x0 = d0[t] * h2, x1 = d1[t] * h2;
                const double x2 = fma(0.5 * (1.0 + T.z1[j]), x1, fma(0.5 * (1.0 + T.z0[i]), x0, d2[t]));
                for (int c = 0; c < 3; ++c) {
                    double o = x0 * a.df[(3 * c) * np + g];
                    o = fma(x1, a.df[(3 * c + 1) * np + g], o);
                    o = fma(x2, a.df[(3 * c + 2) * np + g], o);
                    a.out[c * nQ + e * Q + t] = o;
                }
            }
        } else {  // Helmholtz
            pyr_bwd(T, d, cin, sa, sb, u, tid, nt);
            pyr_deriv(T, d, u, d0, d1, d2, tid, nt);
            for (int t = tid; t < Q; t += nt) {
                const int i = t % d.nq0, kj = t / d.nq0, j = kj % d.nq1, k = kj / d.nq1;
                const long g = deformed ? e * Q + t : e;
                const double wj = point_weight(T, d, t) * (deformed ? jac[t] : jac[0]);
                u[t] *= wj;
                double f[9];
                for (int c = 0; c < 9; ++c) f[c] = a.df[c * np + g];
                const double h2 = 2.0 / (1.0 - T.z2[k]);
                const double h0h2 = 0.5 * (1.0 + T.z0[i]) * h2, h1h2 = 0.5 * (1.0 + T.z1[j]) * h2;
                // gradient coefficients of (d0, d1, d2) per physical direction c:
                // (h2 f[3c] + h0h2 f[3c+2], h2 f[3c+1] + h1h2 f[3c+2], f[3c+2])
                const double t0 = fma(h0h2, f[2], h2 * f[0]), t1 = fma(h0h2, f[5], h2 * f[3]),
                             t2 = fma(h0h2, f[8], h2 * f[6]);
                const double t3 = fma(h1h2, f[2], h2 * f[1]), t4 = fma(h1h2, f[5], h2 * f[4]),
                             t5 = fma(h1h2, f[8], h2 * f[7]);
                const double g00 = fma(t2, t2, fma(t1, t1, t0 * t0));
                const double g11 = fma(t5, t5, fma(t4, t4, t3 * t3));
                const double g22 = fma(f[8], f[8], fma(f[5], f[5], f[2] * f[2]));
                const double g01 = fma(t2, t5, fma(t1, t4, t0 * t3));
                const double g02 = fma(f[8], t2, fma(f[5], t1, f[2] * t0));
                const double g12 = fma(f[8], t5, fma(f[5], t4, f[2] * t3));
                const double a0 = d0[t], a1 = d1[t], a2 = d2[t];
                d0[t] = wj * fma(g02, a2, fma(g01, a1, g00 * a0));
                d1[t] = wj * fma(g12, a2, fma(g11, a1, g01 * a0));
                d2[t] = wj * fma(g22, a2, fma(g12, a1, g02 * a0));
            }
            __syncthreads();
            pyr_iprod(d, T.B0, T.B1, T.B2, u, sb, sa, cout, true, false, a.lambda, tid, nt);
            pyr_iprod(d, T.dB0, T.B1, T.B2, d0, sb, sa, cout, false, true, 1.0, tid, nt);
            pyr_iprod(d, T.B0, T.dB1, T.B2, d1, sb, sa, cout, false, true, 1.0, tid, nt);
            pyr_iprod(d, T.B0, T.B1, T.dB2, d2, sb, sa, cout, false, true, 1.0, tid, nt);
            for (int t = tid; t < M; t += nt) a.out[e * M + t] = cout[t];
        }
        __syncthreads();  // the next element overwrites the arrays
    }
}

}  // namespace

hipError_t launch_pyr(const PyrArgs &a, hipStream_t s)
{
    const size_t lds = sizeof(double) * pyr_lds_len(a.nm, a.nq0, a.nq1, a.nq2);
    hipError_t e = hipFuncSetAttribute((const void *)&pyr_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const int Q = a.nq0 * a.nq1 * a.nq2;
    const int nt = Q <= 64 ? 64 : Q <= 128 ? 128 : 256;
    // enough workgroups to fill every CU several times over, each looping over elements
    const unsigned grid = (unsigned)(a.nElmt < 8192 ? a.nElmt : 8192);
    pyr_kernel<<<grid, nt, lds, s>>>(a);
    return hipGetLastError();
}

}  // namespace imf
