// Arguments of the hexahedral operators with per-direction modes and points
// (mf_hex_aniso.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace imf {

struct AnisoArgs {
    int kind;          // kBwdTrans, kIProduct, kMass or kHelmholtz
    int geo;           // kUnit, kDeformed (per point) or kRegular (per element)
    int nm[3], nq[3];  // modes / points per direction
    int nElmt;
    int bufLen;        // doubles per LDS work buffer (largest intermediate array)
    double lambda;
    const double *tab;  // device: B0 B1 B2 (nm_d*nq_d, B[p*nq+i]), D0 D1 D2 (nq_d^2,
                        // D[l*nq+x] = l_l'(z_x)), w0 w1 w2
    const double *in;
    double *out;
    const double *jac;  // GetJac layout: [nElmt*Q] deformed, [nElmt] regular
    const double *df;   // GetDerivFactors layout: [9][nElmt*Q] deformed, [9][nElmt] regular
    int nvec = 1;
    long ldin = 0, ldout = 0;
};

__host__ __device__ inline int aniso_table_len(const int *nm, const int *nq)
{
    int n = 0;
    for (int d = 0; d < 3; ++d) n += nm[d] * nq[d] + nq[d] * nq[d] + nq[d];
    return n;
}

hipError_t launch_aniso(const AnisoArgs &a, hipStream_t s);
hipError_t launch_aniso_diag(const AnisoArgs &a, double *diag, hipStream_t s);

}  // namespace imf
