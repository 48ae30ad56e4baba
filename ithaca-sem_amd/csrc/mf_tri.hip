// Triangle operators in collapsed coordinates (FP64, runtime sizes).
//
// Restates the reference's MatrixFree triangle kernels: BwdTransTriKernel
// (MatrixFreeOps/BwdTransKernels.hpp:77-126), IProductTriKernel (IProductKernels.hpp:
// 134-236), PhysDerivTriKernel (PhysDerivKernels.hpp:150-216) and HelmholtzTriImpl
// (Helmholtz.h:504-620).  The expansion is Modified_A along eta0 (nq0 GLL points) times
// Modified_B along eta1 (nq1 Gauss-Radau points, alpha = 1), with the modes packed p-major
// (mode = sum_{p'<p} (nm - p') + q) and mode 1 the collapsed vertex, which also carries
// phi_1(eta0) (the CORRECT terms).  The weights along eta1 include the collapse factor 1/2
// (Helper, Operator.hpp:244-258); the metric of the Laplacian is formed per point from
// the derivative factors with h0 = (1 + eta0)/2, h1 = 2/(1 - eta1).
//
// Layout on the GPU: one wave per element, four elements per 256-thread workgroup; the
// basis tables are staged once per workgroup in LDS, and each element's coefficients and
// point arrays live in LDS while its wave runs the phases of the operator, every phase a
// loop of the wave's 64 lanes over the phase's independent outputs.  Triangles are small
// (at most 36 modes and 240 points), so the cost is dominated by the passes over LDS,
// not by memory traffic.
#include <hip/hip_runtime.h>

#include "imf_common.hpp"
#include "mf_tri.hpp"

namespace imf {

namespace {

constexpr int TB = 256;             // threads per workgroup
constexpr int EPB = TB / 64;        // elements per workgroup (one wave each)

struct TriTab {
    const double *B0, *dB0, *D0, *w0, *z0, *B1, *dB1, *D1, *w1, *z1;
};

__device__ inline TriTab tri_tab(const double *t, int nm, int nq0, int nq1)
{
    TriTab r;
    r.B0 = t;
    r.dB0 = r.B0 + nm * nq0;
    r.D0 = r.dB0 + nm * nq0;
    r.w0 = r.D0 + nq0 * nq0;
    r.z0 = r.w0 + nq0;
    r.B1 = r.z0 + nq0;
    r.dB1 = r.B1 + tri_modes(nm) * nq1;
    r.D1 = r.dB1 + tri_modes(nm) * nq1;
    r.w1 = r.D1 + nq1 * nq1;
    r.z1 = r.w1 + nq1;
    return r;
}

__device__ inline int mode_start(int p, int nm) { return p * nm - p * (p - 1) / 2; }

// u = BwdTrans(c): p-sums along eta1 for every (p, j), then the sum over p at every point
// with the collapsed-vertex correction.  ps: nm*nq1 scratch.
__device__ void tri_bwd(const TriTab &T, int nm, int nq0, int nq1, const double *c, double *ps,
                        double *u, int lane)
{
    for (int t = lane; t < nm * nq1; t += 64) {
        const int p = t / nq1, j = t - p * nq1, m0 = mode_start(p, nm);
        double s = 0.0;
        for (int q = 0; q < nm - p; ++q) s = fma(T.B1[(m0 + q) * nq1 + j], c[m0 + q], s);
        ps[p * nq1 + j] = s;
    }
    __syncthreads();
    for (int t = lane; t < nq0 * nq1; t += 64) {
        const int j = t / nq0, i = t - j * nq0;
        double s = 0.0;
        for (int p = 0; p < nm; ++p) s = fma(ps[p * nq1 + j], T.B0[p * nq0 + i], s);
        u[t] = fma(c[1] * T.B0[nq0 + i], T.B1[nq1 + j], s);
    }
    __syncthreads();
}

// IProduct of the point array f with tables b0 (eta0) and b1 (eta1): out[mode] =
// sum_j b1[mode](j) w1_j sum_i b0[p](i) w0_i J f, then the correction of mode 1.  With
// unit, weights and Jacobian are one (the transpose of BwdTrans).  Stores out = s * sc
// (scale), out += s (append) or out = s.
__device__ void tri_iprod(const TriTab &T, const double *b0, const double *b1, int nm,
                          int nq0, int nq1, bool unit, bool deformed, const double *jac,
                          const double *f, double *ps, double *out, bool scale, bool append,
                          double sc, int lane)
{
    const int M = tri_modes(nm);
    for (int t = lane; t < nm * nq1; t += 64) {
        const int p = t / nq1, j = t - p * nq1;
        double s = 0.0;
        for (int i = 0; i < nq0; ++i) {
            const double jv = unit ? 1.0 : deformed ? jac[j * nq0 + i] : jac[0];
            s = fma(f[j * nq0 + i] * b0[p * nq0 + i] * jv, unit ? 1.0 : T.w0[i], s);
        }
        ps[p * nq1 + j] = s;
    }
    __syncthreads();
    for (int t = lane; t < M; t += 64) {
        int p = 0;
        while (mode_start(p + 1, nm) <= t) ++p;
        double s = 0.0;
        for (int j = 0; j < nq1; ++j) s = fma(ps[p * nq1 + j] * b1[t * nq1 + j], unit ? 1.0 : T.w1[j], s);
        double v = scale && append ? fma(s, sc, out[t]) : append ? out[t] + s : scale ? s * sc : s;
        if (t == 1) {
            double c01 = 0.0;
            for (int j = 0; j < nq1; ++j) {
                const double pre = unit ? b1[nq1 + j]
                                 : deformed ? T.w1[j] * b1[nq1 + j] : T.w1[j] * jac[0] * b1[nq1 + j];
                for (int i = 0; i < nq0; ++i) {
                    double prod = f[j * nq0 + i] * pre * (unit ? 1.0 : T.w0[i]);
                    if (!unit && deformed) prod = prod * jac[j * nq0 + i];
                    c01 = fma(prod, b0[nq0 + i], c01);
                }
            }
            v = scale ? fma(c01, sc, v) : v + c01;
        }
        out[t] = v;
    }
    __syncthreads();
}

// Reference-space derivatives along eta0 and eta1 (PhysDerivTensor2DKernel).
__device__ void tri_deriv(const TriTab &T, int nq0, int nq1, const double *u, double *d0,
                          double *d1, int lane)
{
    for (int t = lane; t < nq0 * nq1; t += 64) {
        const int j = t / nq0, i = t - j * nq0;
        double s0 = 0.0, s1 = 0.0;
        for (int k = 0; k < nq0; ++k) s0 = fma(T.D0[k * nq0 + i], u[j * nq0 + k], s0);
        for (int k = 0; k < nq1; ++k) s1 = fma(u[k * nq0 + i], T.D1[k * nq1 + j], s1);
        d0[t] = s0;
        d1[t] = s1;
    }
    __syncthreads();
}

__global__ __launch_bounds__(TB) void tri_kernel(TriArgs a)
{
    extern __shared__ double sm[];
    const int nm = a.nm, nq0 = a.nq0, nq1 = a.nq1, M = tri_modes(nm), Q = nq0 * nq1;
    const int tl = tri_table_len(nm, nq0, nq1);
    for (int t = threadIdx.x; t < tl; t += TB) sm[t] = a.tab[t];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int per = 2 * M + 3 * Q + nm * nq1;
    double *cin = sm + tl + wv * per, *cout = cin + M, *u = cout + M, *d0 = u + Q, *d1 = d0 + Q,
           *ps = d1 + Q;
    const long e = (long)blockIdx.x * EPB + wv;
    const bool valid = e < a.nElmt;
    const bool deformed = a.geo == kDeformed;
    // waves past the last element compute on element 0's geometry (in bounds) and store nothing
    const long ej = valid ? e : 0;
    const double *jac = a.jac ? a.jac + (deformed ? ej * Q : ej) : nullptr;
    __syncthreads();
    const TriTab T = tri_tab(sm, nm, nq0, nq1);

    const bool coef_in = (a.kind == kBwdTrans && !a.transpose) || (a.kind == kIProduct && a.transpose) ||
                         a.kind == kHelmholtz;
    if (coef_in) {
        for (int t = lane; t < M; t += 64) cin[t] = valid ? a.in[e * M + t] : 0.0;
    } else if (a.kind != kIProdDeriv) {
        for (int t = lane; t < Q; t += 64) u[t] = valid ? a.in[e * Q + t] : 0.0;
    }
    __syncthreads();

    if (a.kind == kBwdTrans && !a.transpose) {
        tri_bwd(T, nm, nq0, nq1, cin, ps, u, lane);
        if (valid)
            for (int t = lane; t < Q; t += 64) a.out[e * Q + t] = u[t];
    } else if (a.kind == kIProduct && a.transpose) {
        // the transpose of IProduct = B^T W J is W J B: out = w0 w1 J (B c)
        tri_bwd(T, nm, nq0, nq1, cin, ps, u, lane);
        if (valid)
            for (int t = lane; t < Q; t += 64) {
                const int j = t / nq0, i = t - j * nq0;
                a.out[e * Q + t] = u[t] * (T.w0[i] * T.w1[j]) * (deformed ? jac[t] : jac[0]);
            }
    } else if (a.kind == kBwdTrans || a.kind == kIProduct) {
        const bool unit = a.kind == kBwdTrans;  // BwdTrans^T: unit weights, no Jacobian
        tri_iprod(T, T.B0, T.B1, nm, nq0, nq1, unit, deformed, jac, u, ps, cout, false, false,
                  1.0, lane);
        if (valid)
            for (int t = lane; t < M; t += 64) a.out[e * M + t] = cout[t];
    } else if (a.kind == kIProdDeriv) {
        // IProductWRTDerivBase: per point t_a = sum_c df[2c+a] in_c over the ncoord fields (a
        // triangle in 3D space has three, IProductWRTDerivBase.h:906-), the transpose of the
        // collapsed chain rule of PhysDeriv (tmp0 = h1 (t0 + h0 t1), tmp1 = t1), then the
        // inner products with (dB0, B1) and (B0, dB1)
        const long np = deformed ? (long)a.nElmt * Q : a.nElmt, nQ = (long)a.nElmt * Q;
        for (int t = lane; t < Q; t += 64) {
            const int j = t / nq0, i = t - j * nq0;
            const long k = deformed ? ej * Q + t : ej;
            const double f0 = a.df[k], f1 = a.df[np + k], f2 = a.df[2 * np + k],
                         f3 = a.df[3 * np + k];
            const double x0 = valid ? a.in[e * Q + t] : 0.0, x1 = valid ? a.in[nQ + e * Q + t] : 0.0;
            double t0 = fma(f2, x1, f0 * x0), t1 = fma(f3, x1, f1 * x0);
            if (a.ncoord == 3) {
                const double x2 = valid ? a.in[2 * nQ + e * Q + t] : 0.0;
                t0 = fma(a.df[4 * np + k], x2, t0);
                t1 = fma(a.df[5 * np + k], x2, t1);
            }
            d0[t] = (2.0 / (1.0 - T.z1[j])) * fma(0.5 * (1.0 + T.z0[i]), t1, t0);
            d1[t] = t1;
        }
        __syncthreads();
        tri_iprod(T, T.dB0, T.B1, nm, nq0, nq1, false, deformed, jac, d0, ps, cout, false, false,
                  1.0, lane);
        tri_iprod(T, T.B0, T.dB1, nm, nq0, nq1, false, deformed, jac, d1, ps, cout, false, true,
                  1.0, lane);
        if (valid)
            for (int t = lane; t < M; t += 64) a.out[e * M + t] = cout[t];
    } else if (a.kind == kPhysDeriv) {
        tri_deriv(T, nq0, nq1, u, d0, d1, lane);
        if (valid) {
            const long np = deformed ? (long)a.nElmt * Q : a.nElmt;
            for (int t = lane; t < Q; t += 64) {
                const int j = t / nq0, i = t - j * nq0;
                const long k = deformed ? e * Q + t : e;
                const double f0 = a.df[k], f1 = a.df[np + k], f2 = a.df[2 * np + k],
                             f3 = a.df[3 * np + k];
                const double x0 = 2.0 / (1.0 - T.z1[j]);
                const double g0 = x0 * d0[t];
                const double g1 = fma(g0, 0.5 * (1.0 + T.z0[i]), d1[t]);
                a.out[e * Q + t] = fma(g1, f1, g0 * f0);
                a.out[(long)a.nElmt * Q + e * Q + t] = fma(g1, f3, g0 * f2);
            }
        }
    } else {  // Helmholtz
        tri_bwd(T, nm, nq0, nq1, cin, ps, u, lane);
        tri_iprod(T, T.B0, T.B1, nm, nq0, nq1, false, deformed, jac, u, ps, cout, true, false,
                  a.lambda, lane);
        tri_deriv(T, nq0, nq1, u, d0, d1, lane);
        if (valid) {
            const long np = deformed ? (long)a.nElmt * Q : a.nElmt;
            for (int t = lane; t < Q; t += 64) {
                const int j = t / nq0, i = t - j * nq0;
                const long k = deformed ? e * Q + t : e;
                const double f0 = a.df[k], f1 = a.df[np + k], f2 = a.df[2 * np + k],
                             f3 = a.df[3 * np + k];
                const double h1 = 2.0 / (1.0 - T.z1[j]), h0 = 0.5 * (1.0 + T.z0[i]);
                double m00 = h1 * (f0 + h0 * f1);
                double m01 = m00 * f1;
                m00 = m00 * m00;
                const double tt = h1 * (f2 + h0 * f3);
                m01 = fma(tt, f3, m01);
                m00 = fma(tt, tt, m00);
                const double m11 = fma(f3, f3, f1 * f1);
                const double a0 = d0[t], a1 = d1[t];
                d0[t] = fma(m01, a1, m00 * a0);
                d1[t] = fma(m11, a1, m01 * a0);
            }
        }
        __syncthreads();
        tri_iprod(T, T.dB0, T.B1, nm, nq0, nq1, false, deformed, jac, d0, ps, cout, false, true,
                  1.0, lane);
        tri_iprod(T, T.B0, T.dB1, nm, nq0, nq1, false, deformed, jac, d1, ps, cout, false, true,
                  1.0, lane);
        if (valid)
            for (int t = lane; t < M; t += 64) a.out[e * M + t] = cout[t];
    }
}

}  // namespace

hipError_t launch_tri(const TriArgs &a, hipStream_t s)
{
    const int M = tri_modes(a.nm), Q = a.nq0 * a.nq1;
    const size_t lds = sizeof(double) * (tri_table_len(a.nm, a.nq0, a.nq1) +
                                         (size_t)EPB * (2 * M + 3 * Q + a.nm * a.nq1));
    hipError_t e = hipFuncSetAttribute((const void *)&tri_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const unsigned grid = (unsigned)((a.nElmt + EPB - 1) / EPB);
    tri_kernel<<<grid, TB, lds, s>>>(a);
    return hipGetLastError();
}

}  // namespace imf
