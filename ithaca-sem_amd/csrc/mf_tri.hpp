// Arguments of the triangle operators (collapsed coordinates; mf_tri.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace imf {

struct TriArgs {
    int kind;       // kBwdTrans, kIProduct, kHelmholtz, kPhysDeriv or kIProdDeriv
    int transpose;  // Euclidean transpose (BwdTrans, IProduct)
    int geo;        // kUnit, kRegular (per element) or kDeformed (per point)
    int nm, nq0, nq1, nElmt;
    int ncoord;     // IProductWRTDerivBase: fields and factor pairs, 2 or 3 (surface in 3D)
    double lambda;
    // device: B0 dB0 (nm*nq0, [p*nq0+i]) D0 (nq0^2, [a*nq0+b] = l_a'(z_b)) w0 z0 (nq0),
    //         B1 dB1 (M*nq1, [mode*nq1+j]) D1 (nq1^2) w1 (0.5 x Gauss-Radau) z1 (nq1)
    const double *tab;
    const double *in;
    double *out;
    const double *jac;  // GetJac layout: [nElmt*Q] deformed, [nElmt] regular
    const double *df;   // GetDerivFactors layout: [4][nElmt*Q] deformed, [4][nElmt] regular
                        // ([6][...] for IProductWRTDerivBase with ncoord = 3)
};

__host__ __device__ inline int tri_modes(int nm) { return nm * (nm + 1) / 2; }

__host__ __device__ inline int tri_table_len(int nm, int nq0, int nq1)
{
    return 2 * nm * nq0 + nq0 * nq0 + 2 * nq0 + 2 * tri_modes(nm) * nq1 + nq1 * nq1 + 2 * nq1;
}

hipError_t launch_tri(const TriArgs &a, hipStream_t s);

}  // namespace imf
