// Arguments of the prism operators (collapsed coordinates; mf_prism.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mf_tri.hpp"

namespace imf {

struct PrismArgs {
    int kind;       // kBwdTrans, kIProduct, kHelmholtz, kPhysDeriv or kIProdDeriv
    int transpose;  // Euclidean transpose (BwdTrans, IProduct)
    int geo;        // kUnit, kRegular (per element) or kDeformed (per point)
    int nm, nq0, nq1, nq2, nElmt;
    double lambda;
    // device: for d = 0, 1 (Modified_A, GLL): B_d dB_d (nm*nq_d, [p*nq_d+i]), D_d (nq_d^2,
    // [a*nq_d+b] = l_a'(z_b)), w_d, z_d (nq_d); for d = 2 (Modified_B, Gauss-Radau):
    // B2 dB2 (nm(nm+1)/2 * nq2, [mode_pr*nq2+k]), D2 (nq2^2), w2 (0.5 x Gauss-Radau), z2
    const double *tab;
    const double *in;
    double *out;
    const double *jac;  // GetJac layout: [nElmt*Q] deformed, [nElmt] regular
    const double *df;   // GetDerivFactors layout: [9][nElmt*Q] deformed, [9][nElmt] regular
};

// Coefficients of one prism: (p, q, r) with p, q < nm and r < nm - p.
__host__ __device__ inline int prism_modes(int nm) { return nm * tri_modes(nm); }

__host__ __device__ inline int prism_table_len(int nm, int nq0, int nq1, int nq2)
{
    return 2 * nm * nq0 + nq0 * nq0 + 2 * nq0 + 2 * nm * nq1 + nq1 * nq1 + 2 * nq1 +
           2 * tri_modes(nm) * nq2 + nq2 * nq2 + 2 * nq2;
}

// Doubles of LDS one workgroup uses (tables, then one element's arrays).
__host__ __device__ inline int prism_lds_len(int nm, int nq0, int nq1, int nq2)
{
    const int Q = nq0 * nq1 * nq2;
    return prism_table_len(nm, nq0, nq1, nq2) + 2 * prism_modes(nm) + 4 * Q +
           nm * nm * nq2 + nm * nq1 * nq2;
}

hipError_t launch_prism(const PrismArgs &a, hipStream_t s);

}  // namespace imf
