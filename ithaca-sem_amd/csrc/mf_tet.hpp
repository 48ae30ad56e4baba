// Arguments of the tetrahedron operators (collapsed coordinates; mf_tet.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mf_tri.hpp"

namespace imf {

struct TetArgs {
    int kind;       // kBwdTrans, kIProduct, kHelmholtz, kPhysDeriv or kIProdDeriv
    int transpose;  // Euclidean transpose (BwdTrans, IProduct)
    int geo;        // kUnit, kRegular (per element) or kDeformed (per point)
    int nm, nq0, nq1, nq2, nElmt;
    double lambda;
    // device: B0 dB0 (nm*nq0, [p*nq0+i]), D0 (nq0^2, [a*nq0+b] = l_a'(z_b)), w0, z0 (GLL);
    // B1 dB1 (nm(nm+1)/2 * nq1, [pq*nq1+j]), D1, w1 (0.5 x Gauss-Radau alpha = 1), z1;
    // B2 dB2 (M * nq2, [mode*nq2+k]), D2, w2 (0.25 x Gauss-Radau alpha = 2), z2
    const double *tab;
    const double *in;
    double *out;
    const double *jac;  // GetJac layout: [nElmt*Q] deformed, [nElmt] regular
    const double *df;   // GetDerivFactors layout: [9][nElmt*Q] deformed, [9][nElmt] regular
};

// Coefficients of one tetrahedron: (p, q, r) with p + q + r < nm.
__host__ __device__ inline int tet_modes(int nm) { return nm * (nm + 1) * (nm + 2) / 6; }

__host__ __device__ inline int tet_table_len(int nm, int nq0, int nq1, int nq2)
{
    return 2 * nm * nq0 + nq0 * nq0 + 2 * nq0 + 2 * tri_modes(nm) * nq1 + nq1 * nq1 + 2 * nq1 +
           2 * tet_modes(nm) * nq2 + nq2 * nq2 + 2 * nq2;
}

// Doubles of LDS one workgroup uses (tables, then one element's arrays).
__host__ __device__ inline int tet_lds_len(int nm, int nq0, int nq1, int nq2)
{
    const int Q = nq0 * nq1 * nq2;
    return tet_table_len(nm, nq0, nq1, nq2) + 2 * tet_modes(nm) + 4 * Q +
           (tri_modes(nm) + 2) * nq2 + nm * nq1 * nq2;
}

hipError_t launch_tet(const TetArgs &a, hipStream_t s);

}  // namespace imf
