// Arguments of the pyramid operators (collapsed coordinates; mf_pyr.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace imf {

struct PyrArgs {
    int kind;       // kBwdTrans, kIProduct, kHelmholtz, kPhysDeriv or kIProdDeriv
    int transpose;  // Euclidean transpose (BwdTrans, IProduct)
    int geo;        // kUnit, kRegular (per element) or kDeformed (per point)
    int nm, nq0, nq1, nq2, nElmt;
    double lambda;
    // device: for d = 0, 1 (Modified_A, GLL): B_d dB_d (nm*nq_d, [p*nq_d+i]), D_d (nq_d^2,
    // [a*nq_d+b] = l_a'(z_b)), w_d, z_d; for d = 2 (ModifiedPyr_C, Gauss-Radau alpha = 2):
    // B2 dB2 (M*nq2, [mode*nq2+k]), D2, w2 (0.25 x Gauss-Radau), z2
    const double *tab;
    const double *in;
    double *out;
    const double *jac;  // GetJac layout: [nElmt*Q] deformed, [nElmt] regular
    const double *df;   // GetDerivFactors layout: [9][nElmt*Q] deformed, [9][nElmt] regular
};

// Coefficients of one pyramid: (p, q, r) with p, q < nm and r < nm - max(p, q).
__host__ __device__ inline int pyr_modes(int nm) { return nm * (nm + 1) * (2 * nm + 1) / 6; }

__host__ __device__ inline int pyr_table_len(int nm, int nq0, int nq1, int nq2)
{
    return 2 * nm * nq0 + nq0 * nq0 + 2 * nq0 + 2 * nm * nq1 + nq1 * nq1 + 2 * nq1 +
           2 * pyr_modes(nm) * nq2 + nq2 * nq2 + 2 * nq2;
}

// Doubles of LDS one workgroup uses (tables, then one element's arrays).
__host__ __device__ inline int pyr_lds_len(int nm, int nq0, int nq1, int nq2)
{
    const int Q = nq0 * nq1 * nq2;
    return pyr_table_len(nm, nq0, nq1, nq2) + 2 * pyr_modes(nm) + 4 * Q + nm * nm * nq2 +
           nm * nq1 * nq2;
}

hipError_t launch_pyr(const PyrArgs &a, hipStream_t s);

}  // namespace imf
