// Tetrahedron operators in collapsed coordinates (FP64, runtime sizes).
//
// Restates the reference's MatrixFree tetrahedron kernels: BwdTransTetKernel
// (MatrixFreeOps/BwdTransKernels.hpp:377-484), IProductTetKernel (IProductKernels.hpp:
// 604-760), PhysDerivTetKernel (PhysDerivKernels.hpp:556-694) and HelmholtzTetImpl
// (Helmholtz.h:1970-2005, :2266-2455).  The expansion is Modified_A along eta0 (GLL) times
// Modified_B along eta1 (Gauss-Radau alpha = 1) times Modified_C along eta2 (Gauss-Radau
// alpha = 2): mode (p, q, r) with p + q + r < nm, r fastest, then q, then p; B1 has a row
// per (p, q) pair, B2 a row per mode.  The top vertex (mode 1), the bottom vertex of the
// collapsed edge (mode nm, i.e. (0, 1, 0)) and the singular edge (modes nm + r) carry
// extra products of the lower directions (the CORRECT terms):
//   mode 1      + B2_1 (B0_0 B1_01 + B0_1 B1_00 + B0_1 B1_01)
//   mode nm     + B2_0 B0_1 B1_01
//   mode nm + r + B2_{r+1} B0_1 B1_01      (1 <= r <= nm - 2)
// The weights along eta1 and eta2 include the collapse factors 1/2 and 1/4 (Helper,
// Operator.hpp:244-258).
//
// Layout on the GPU: one workgroup per element at a time (a grid-stride loop over the
// elements), the tables staged once per workgroup in LDS, the element's arrays in LDS.
// The sum factorisation runs in three phases each way; the corrections enter it exactly:
//   BwdTrans   t1[pq][k] = sum_r B2 c  (+ c_1 B2_1 into pair (0,1)),  X0 = c_1 B2_1,
//              X = c_1 B2_1 + c_nm B2_0 + sum_r c_{nm+r} B2_{r+1};
//              t2[p][j][k] = sum_q B1 t1 (+ B1_00 X0 + B1_01 X for p = 1);
//              u = sum_p B0 t2
//   IProduct   s1[p][j][k] = sum_i B0 g;  s2[pq][k] = sum_j B1 s1,  Y0 = sum_j B1_00 s1[1],
//              Y = sum_j B1_01 s1[1];  c(p,q,r) = sum_k B2 s2 (+ the corrections from
//              s2[(0,1)], Y0 and Y)
// with g the point values times w0 w1 w2 J.
#include <hip/hip_runtime.h>

#include "imf_common.hpp"
#include "mf_tet.hpp"

namespace imf {

namespace {

struct TetTab {
    const double *B0, *dB0, *D0, *w0, *z0, *B1, *dB1, *D1, *w1, *z1, *B2, *dB2, *D2, *w2, *z2;
};

__device__ inline TetTab tet_tab(const double *t, int nm, int nq0, int nq1, int nq2)
{
    TetTab r;
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
    r.B2 = r.z1 + nq1;
    r.dB2 = r.B2 + tet_modes(nm) * nq2;
    r.D2 = r.dB2 + tet_modes(nm) * nq2;
    r.w2 = r.D2 + nq2 * nq2;
    r.z2 = r.w2 + nq2;
    return r;
}

struct Dims {
    int nm, nq0, nq1, nq2;
};

// first (p, q) pair of group p
__device__ inline int pair_start(int p, int nm) { return p * nm - p * (p - 1) / 2; }

// (p, q) of pair index pq
__device__ inline void pair_pq(int pq, int nm, int &p, int &q)
{
    p = 0;
    while (pair_start(p + 1, nm) <= pq) ++p;
    q = pq - pair_start(p, nm);
}

// first mode of pair (p, q): the pairs before it hold nm - p' - q' modes each
__device__ inline int pair_mode0(int p, int q, int nm)
{
    int m = 0;
    for (int pp = 0; pp < p; ++pp) m += (nm - pp) * (nm - pp + 1) / 2;
    for (int qq = 0; qq < q; ++qq) m += nm - p - qq;
    return m;
}

// u = BwdTrans(c) through t1 (T*nq2, then X0 and X: nq2 each) and t2 (nm*nq1*nq2).
__device__ void tet_bwd(const TetTab &T, Dims d, const double *c, double *t1, double *t2,
                        double *u, int tid, int nt)
{
    const int nm = d.nm, nq0 = d.nq0, nq1 = d.nq1, nq2 = d.nq2, np = tri_modes(nm);
    double *X0 = t1 + np * nq2, *X = X0 + nq2;
    for (int t = tid; t < (np + 2) * nq2; t += nt) {
        const int k = t % nq2, pq = t / nq2;
        double s = 0.0;
        if (pq < np) {
            int p, q;
            pair_pq(pq, nm, p, q);
            const int m0 = pair_mode0(p, q, nm);
            for (int r = 0; r < nm - p - q; ++r) s = fma(T.B2[(m0 + r) * nq2 + k], c[m0 + r], s);
            if (pq == 1) s = fma(c[1], T.B2[nq2 + k], s);  // B0_0 B1_01 B2_1 of mode 1
        } else if (pq == np) {
            s = c[1] * T.B2[nq2 + k];
        } else {
            s = c[1] * T.B2[nq2 + k];
            s = fma(c[nm], T.B2[k], s);
            for (int r = 1; r < nm - 1; ++r) s = fma(c[nm + r], T.B2[(r + 1) * nq2 + k], s);
        }
        t1[t] = s;
    }
    __syncthreads();
    for (int t = tid; t < nm * nq1 * nq2; t += nt) {
        const int k = t % nq2, pj = t / nq2, j = pj % nq1, p = pj / nq1;
        const int ps = pair_start(p, nm);
        double s = 0.0;
        for (int q = 0; q < nm - p; ++q) s = fma(T.B1[(ps + q) * nq1 + j], t1[(ps + q) * nq2 + k], s);
        if (p == 1) {
            s = fma(T.B1[j], X0[k], s);
            s = fma(T.B1[nq1 + j], X[k], s);
        }
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
// (eta1) and b2 (eta2) through s1 (nm*nq1*nq2) and s2 (T*nq2, then Y0 and Y).  Stores
// out = s * sc (scale), out += s (append) or out = s.
__device__ void tet_iprod(Dims d, const double *b0, const double *b1, const double *b2,
                          const double *g, double *s1, double *s2, double *out, bool scale,
                          bool append, double sc, int tid, int nt)
{
    const int nm = d.nm, nq0 = d.nq0, nq1 = d.nq1, nq2 = d.nq2, np = tri_modes(nm);
    for (int t = tid; t < nm * nq1 * nq2; t += nt) {
        const int k = t % nq2, pj = t / nq2, j = pj % nq1, p = pj / nq1;
        const double *gr = g + (k * nq1 + j) * nq0;
        double s = 0.0;
        for (int i = 0; i < nq0; ++i) s = fma(b0[p * nq0 + i], gr[i], s);
        s1[t] = s;
    }
    __syncthreads();
    for (int t = tid; t < (np + 2) * nq2; t += nt) {
        const int k = t % nq2, pq = t / nq2;
        int p, row;
        if (pq < np) {
            int q;
            pair_pq(pq, nm, p, q);
            row = pq;
        } else {  // Y0 (pair (0,0)'s table) and Y (pair (0,1)'s) against s1[1]
            p = 1;
            row = pq - np;
        }
        double s = 0.0;
        for (int j = 0; j < nq1; ++j) s = fma(b1[row * nq1 + j], s1[(p * nq1 + j) * nq2 + k], s);
        s2[t] = s;
    }
    __syncthreads();
    const double *Y0 = s2 + np * nq2, *Y = Y0 + nq2;
    for (int t = tid; t < tet_modes(nm); t += nt) {
        int p = 0, q = 0, m0 = 0;
        for (;;) {  // the pair (p, q) holding mode t
            const int n = nm - p - q;
            if (t < m0 + n) break;
            m0 += n;
            if (++q == nm - p) {
                q = 0;
                ++p;
            }
        }
        const double *row = b2 + t * nq2, *col = s2 + (pair_start(p, nm) + q) * nq2;
        double s = 0.0;
        for (int k = 0; k < nq2; ++k) s = fma(row[k], col[k], s);
        if (t == 1) {  // top vertex
            const double *c01 = s2 + nq2;
            for (int k = 0; k < nq2; ++k) s = fma(b2[nq2 + k], c01[k] + Y0[k] + Y[k], s);
        } else if (t == nm) {  // bottom vertex of the collapsed edge
            for (int k = 0; k < nq2; ++k) s = fma(b2[k], Y[k], s);
        } else if (t > nm && t < 2 * nm - 1) {  // singular edge, r = t - nm
            const double *rw = b2 + (t - nm + 1) * nq2;
            for (int k = 0; k < nq2; ++k) s = fma(rw[k], Y[k], s);
        }
        out[t] = scale && append ? fma(s, sc, out[t]) : append ? out[t] + s : scale ? s * sc : s;
    }
    __syncthreads();
}

// Reference-space derivatives along eta0, eta1 and eta2 (PhysDerivTensor3DKernel).
__device__ void tet_deriv(const TetTab &T, Dims d, const double *u, double *d0, double *d1,
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

__device__ inline double point_weight(const TetTab &T, Dims d, int t)
{
    const int i = t % d.nq0, kj = t / d.nq0, j = kj % d.nq1, k = kj / d.nq1;
    return T.w0[i] * T.w1[j] * T.w2[k];
}

__global__ __launch_bounds__(256) void tet_kernel(TetArgs a)
{
    extern __shared__ double sm[];
    const Dims d{a.nm, a.nq0, a.nq1, a.nq2};
    const int nm = a.nm, M = tet_modes(nm), Q = a.nq0 * a.nq1 * a.nq2;
    const int tl = tet_table_len(nm, a.nq0, a.nq1, a.nq2);
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int t = tid; t < tl; t += nt) sm[t] = a.tab[t];
    double *cin = sm + tl, *cout = cin + M, *u = cout + M, *d0 = u + Q, *d1 = d0 + Q,
           *d2 = d1 + Q, *sa = d2 + Q, *sb = sa + (tri_modes(nm) + 2) * a.nq2;
    const bool deformed = a.geo == kDeformed;
    const TetTab T = tet_tab(sm, nm, a.nq0, a.nq1, a.nq2);
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
            tet_bwd(T, d, cin, sa, sb, u, tid, nt);
            for (int t = tid; t < Q; t += nt) a.out[e * Q + t] = u[t];
        } else if (a.kind == kIProduct && a.transpose) {
            // the transpose of IProduct = B^T W J is W J B
            tet_bwd(T, d, cin, sa, sb, u, tid, nt);
            for (int t = tid; t < Q; t += nt)
                a.out[e * Q + t] = u[t] * point_weight(T, d, t) * (deformed ? jac[t] : jac[0]);
        } else if (a.kind == kBwdTrans || a.kind == kIProduct) {
            if (a.kind == kIProduct) {  // BwdTrans^T: unit weights, no Jacobian
                for (int t = tid; t < Q; t += nt)
                    u[t] = u[t] * point_weight(T, d, t) * (deformed ? jac[t] : jac[0]);
                __syncthreads();
            }
            tet_iprod(d, T.B0, T.B1, T.B2, u, sb, sa, cout, false, false, 1.0, tid, nt);
            for (int t = tid; t < M; t += nt) a.out[e * M + t] = cout[t];
        } else if (a.kind == kIProdDeriv) {
            // IProductWRTDerivBase (IProductWRTDerivBase.h:2056-): per point t_a = sum_c df[3c+a] in_c,
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
                // adjoint of the tetrahedral chain rule (h2 h3 D0, h0 h2 h3 D0 + h3 D1,
                // h0 h2 h3 D0 + h1 h3 D1 + D2): (h2 h3 (t0 + h0 (t1 + t2)), h3 (t1 + h1 t2), t2)
                const int i = t % d.nq0, kj = t / d.nq0, j = kj % d.nq1, k = kj / d.nq1;
                const double h3 = 2.0 / (1.0 - T.z2[k]);
                v0 = (2.0 / (1.0 - T.z1[j])) * h3 * fma(0.5 * (1.0 + T.z0[i]), v1 + v2, v0);
                v1 = h3 * fma(0.5 * (1.0 + T.z1[j]), v2, v1);
                const double wj = point_weight(T, d, t) * (deformed ? jac[t] : jac[0]);
                d0[t] = wj * v0;
                d1[t] = wj * v1;
                d2[t] = wj * v2;
            }
            __syncthreads();
            tet_iprod(d, T.dB0, T.B1, T.B2, d0, sb, sa, cout, false, false, 1.0, tid, nt);
            tet_iprod(d, T.B0, T.dB1, T.B2, d1, sb, sa, cout, false, true, 1.0, tid, nt);
            tet_iprod(d, T.B0, T.B1, T.dB2, d2, sb, sa, cout, false, true, 1.0, tid, nt);
            for (int t = tid; t < M; t += nt) a.out[e * M + t] = cout[t];
        } else if (a.kind == kPhysDeriv) {
            tet_deriv(T, d, u, d0, d1, d2, tid, nt);
            const long nQ = (long)a.nElmt * Q;
            for (int t = tid; t < Q; t += nt) {
                const int i = t % d.nq0, kj = t / d.nq0, j = kj % d.nq1, k = kj / d.nq1;
                const long g = deformed ? e * Q + t : e;
                const double h3 = 2.0 / (1.0 - T.z2[k]);
                const double x0 = (2.0 / (1.0 - T.z1[j])) * h3 * d0[t];
                const double h0x0 = 0.5 * (1.0 + T.z0[i]) * x0, x1g = h3 * d1[t];
                const double x1 = h0x0 + x1g;
                const double x2 = fma(x1g, 0.5 * (1.0 + T.z1[j]), h0x0) + d2[t];
                for (int c = 0; c < 3; ++c) {
                    double o = x0 * a.df[(3 * c) * np + g];
                    o = fma(x1, a.df[(3 * c + 1) * np + g], o);
                    o = fma(x2, a.df[(3 * c + 2) * np + g], o);
                    a.out[c * nQ + e * Q + t] = o;
                }
            }
        } else {  // Helmholtz
            tet_bwd(T, d, cin, sa, sb, u, tid, nt);
            tet_deriv(T, d, u, d0, d1, d2, tid, nt);
            for (int t = tid; t < Q; t += nt) {
                const int i = t % d.nq0, kj = t / d.nq0, j = kj % d.nq1, k = kj / d.nq1;
                const long g = deformed ? e * Q + t : e;
                const double wj = point_weight(T, d, t) * (deformed ? jac[t] : jac[0]);
                u[t] *= wj;
                double f[9];
                for (int c = 0; c < 9; ++c) f[c] = a.df[c * np + g];
                const double h3 = 2.0 / (1.0 - T.z2[k]);
                const double h1 = 0.5 * (1.0 + T.z1[j]), h2 = 2.0 / (1.0 - T.z1[j]);
                const double h2h3 = h2 * h3, h1h3 = h1 * h3;
                const double h0h2h3 = 0.5 * (1.0 + T.z0[i]) * h2h3;
                // gradient coefficients of (d0, d1, d2) per physical direction c:
                // (h2h3 f[3c] + h0h2h3 (f[3c+1] + f[3c+2]), h3 f[3c+1] + h1h3 f[3c+2], f[3c+2])
                const double t1 = fma(f[0], h2h3, h0h2h3 * (f[1] + f[2]));
                const double t2 = fma(f[3], h2h3, h0h2h3 * (f[4] + f[5]));
                const double t3 = fma(f[6], h2h3, h0h2h3 * (f[7] + f[8]));
                const double t4 = fma(f[2], h1h3, f[1] * h3);
                const double t5 = fma(f[5], h1h3, f[4] * h3);
                const double t6 = fma(f[8], h1h3, f[7] * h3);
                const double g00 = fma(t3, t3, fma(t2, t2, t1 * t1));
                const double g02 = fma(f[8], t3, fma(f[5], t2, f[2] * t1));
                const double g01 = fma(t3, t6, fma(t2, t5, t1 * t4));
                const double g11 = fma(t6, t6, fma(t5, t5, t4 * t4));
                const double g12 = fma(f[8], t6, fma(f[5], t5, f[2] * t4));
                const double g22 = fma(f[8], f[8], fma(f[5], f[5], f[2] * f[2]));
                const double a0 = d0[t], a1 = d1[t], a2 = d2[t];
                d0[t] = wj * fma(g02, a2, fma(g01, a1, g00 * a0));
                d1[t] = wj * fma(g12, a2, fma(g11, a1, g01 * a0));
                d2[t] = wj * fma(g22, a2, fma(g12, a1, g02 * a0));
            }
            __syncthreads();
            tet_iprod(d, T.B0, T.B1, T.B2, u, sb, sa, cout, true, false, a.lambda, tid, nt);
            tet_iprod(d, T.dB0, T.B1, T.B2, d0, sb, sa, cout, false, true, 1.0, tid, nt);
            tet_iprod(d, T.B0, T.dB1, T.B2, d1, sb, sa, cout, false, true, 1.0, tid, nt);
            tet_iprod(d, T.B0, T.B1, T.dB2, d2, sb, sa, cout, false, true, 1.0, tid, nt);
            for (int t = tid; t < M; t += nt) a.out[e * M + t] = cout[t];
        }
        __syncthreads();  // the next element overwrites the arrays
    }
}

}  // namespace

hipError_t launch_tet(const TetArgs &a, hipStream_t s)
{
    const size_t lds = sizeof(double) * tet_lds_len(a.nm, a.nq0, a.nq1, a.nq2);
    hipError_t e = hipFuncSetAttribute((const void *)&tet_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const int Q = a.nq0 * a.nq1 * a.nq2;
    const int nt = Q <= 64 ? 64 : Q <= 128 ? 128 : 256;
    // enough workgroups to fill every CU several times over, each looping over elements
    const unsigned grid = (unsigned)(a.nElmt < 8192 ? a.nElmt : 8192);
    tet_kernel<<<grid, nt, lds, s>>>(a);
    return hipGetLastError();
}

}  // namespace imf
