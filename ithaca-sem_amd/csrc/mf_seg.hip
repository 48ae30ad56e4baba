// Segment operators (FP64, runtime sizes).
//
// Restates the reference's MatrixFree segment operators: BwdTransSeg (MatrixFreeOps/
// BwdTrans.h:16-), IProductSeg (IProduct.h:17-), PhysDerivSeg (PhysDeriv.h:17-) and
// IProductWRTDerivBaseSeg (IProductWRTDerivBase.h:18-, IProductWRTDerivBaseSegImpl): a
// Modified_A expansion of nm modes on nq points in one direction, in a space of ncoord
// coordinates (a segment of a 2D or 3D mesh has ncoord derivative factors per point;
// PhysDeriv gives ncoord derivatives and IProductWRTDerivBase takes ncoord fields).
//
// Layout on the GPU: one thread per output value, grid-stride; the tables are staged in
// LDS once per workgroup.  Consecutive threads take consecutive outputs of an element, so
// the element's inputs are read by neighbouring lanes (coalesced through the caches).
// Segments are small (nq values per element), so these kernels are bound by memory.
#include <hip/hip_runtime.h>

#include "imf_common.hpp"
#include "mf_seg.hpp"

namespace imf {

namespace {

constexpr int TB = 256;

__global__ __launch_bounds__(TB) void seg_kernel(SegArgs a)
{
    extern __shared__ double sm[];
    const int nm = a.nm, nq = a.nq, tl = seg_table_len(nm, nq);
    for (int t = threadIdx.x; t < tl; t += TB) sm[t] = a.tab[t];
    __syncthreads();
    const double *B = sm, *dB = B + nm * nq, *D = dB + nm * nq, *w = D + nq * nq;
    const bool deformed = a.geo == kDeformed;
    const long E = a.nElmt, np = deformed ? E * nq : E;
    auto jv = [&](long e, int i) { return a.jac[deformed ? e * nq + i : e]; };
    auto dfv = [&](int c, long e, int i) { return a.df[c * np + (deformed ? e * nq + i : e)]; };
    const bool to_modes = (a.kind == kIProduct && !a.transpose) ||
                          (a.kind == kBwdTrans && a.transpose) || a.kind == kIProdDeriv;
    const long nout = a.kind == kPhysDeriv ? (long)a.ncoord * E * nq : to_modes ? E * nm : E * nq;
    for (long g = (long)blockIdx.x * TB + threadIdx.x; g < nout; g += (long)gridDim.x * TB) {
        if (to_modes) {
            const long e = g / nm;
            const int p = (int)(g - e * nm);
            const double *f = a.in + e * nq;
            double s = 0.0;
            if (a.kind == kBwdTrans) {  // BwdTrans^T: unit weights, no Jacobian
                for (int i = 0; i < nq; ++i) s = fma(B[p * nq + i], f[i], s);
            } else if (a.kind == kIProduct) {
                for (int i = 0; i < nq; ++i) s = fma(B[p * nq + i], w[i] * jv(e, i) * f[i], s);
            } else {  // IProductWRTDerivBase: sum over the ncoord fields of df_c in_c
                for (int i = 0; i < nq; ++i) {
                    double t = 0.0;
                    for (int c = 0; c < a.ncoord; ++c) t = fma(dfv(c, e, i), a.in[c * E * nq + e * nq + i], t);
                    s = fma(dB[p * nq + i], w[i] * jv(e, i) * t, s);
                }
            }
            a.out[g] = s;
        } else if (a.kind == kPhysDeriv) {
            const int c = (int)(g / (E * nq));
            const long r = g - (long)c * E * nq, e = r / nq;
            const int i = (int)(r - e * nq);
            const double *u = a.in + e * nq;
            double s = 0.0;
            for (int j = 0; j < nq; ++j) s = fma(D[j * nq + i], u[j], s);
            a.out[g] = dfv(c, e, i) * s;
        } else {  // BwdTrans, or IProduct^T = W J B
            const long e = g / nq;
            const int i = (int)(g - e * nq);
            const double *cf = a.in + e * nm;
            double s = 0.0;
            for (int p = 0; p < nm; ++p) s = fma(B[p * nq + i], cf[p], s);
            a.out[g] = a.kind == kIProduct ? s * w[i] * jv(e, i) : s;
        }
    }
}

__global__ void geom_seg_kernel(const double *__restrict__ v, int nElmt, int ncoord, int nq,
                                int deformed, double *__restrict__ jac, double *__restrict__ df)
{
    const long npt = deformed ? (long)nElmt * nq : nElmt;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npt) return;
    const long e = deformed ? g / nq : g;
    const double *x0 = v + e * 2 * ncoord, *x1 = x0 + ncoord;
    double t[3] = {0.0, 0.0, 0.0}, tt = 0.0;
    for (int c = 0; c < ncoord; ++c) {
        t[c] = 0.5 * (x1[c] - x0[c]);
        tt = fma(t[c], t[c], tt);
    }
    jac[g] = sqrt(tt);
    for (int c = 0; c < ncoord; ++c) df[c * npt + g] = t[c] / tt;
}

}  // namespace

hipError_t launch_seg(const SegArgs &a, hipStream_t s)
{
    const size_t lds = sizeof(double) * seg_table_len(a.nm, a.nq);
    const bool to_modes = (a.kind == kIProduct && !a.transpose) ||
                          (a.kind == kBwdTrans && a.transpose) || a.kind == kIProdDeriv;
    const long nout = a.kind == kPhysDeriv ? (long)a.ncoord * a.nElmt * a.nq
                    : (long)a.nElmt * (to_modes ? a.nm : a.nq);
    const long blocks = (nout + TB - 1) / TB;
    const unsigned grid = (unsigned)(blocks < 16384 ? (blocks > 0 ? blocks : 1) : 16384);
    seg_kernel<<<grid, TB, lds, s>>>(a);
    return hipGetLastError();
}

hipError_t geom_seg(const double *d_verts, int nElmt, int ncoord, int nq, int deformed,
                    double *d_jac, double *d_df, hipStream_t s)
{
    const long npt = deformed ? (long)nElmt * nq : nElmt;
    geom_seg_kernel<<<(unsigned)((npt + 255) / 256), 256, 0, s>>>(d_verts, nElmt, ncoord, nq,
                                                                   deformed, d_jac, d_df);
    return hipGetLastError();
}

}  // namespace imf
