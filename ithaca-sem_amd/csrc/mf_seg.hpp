// Arguments of the segment operators (one-dimensional expansions; mf_seg.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace imf {

struct SegArgs {
    int kind;       // kBwdTrans, kIProduct, kPhysDeriv or kIProdDeriv
    int transpose;  // Euclidean transpose (BwdTrans, IProduct)
    int geo;        // kUnit, kRegular (per element) or kDeformed (per point)
    int nm, nq, nElmt;
    int ncoord;     // coordinate dimension (1, 2 or 3): derivative factors per point
    // device: B, dB (nm*nq, [p*nq+i]), D (nq^2, [a*nq+b] = l_a'(z_b)), w (nq)
    const double *tab;
    const double *in;
    double *out;
    const double *jac;  // GetJac layout: [nElmt*nq] deformed, [nElmt] regular
    const double *df;   // GetDerivFactors layout: [ncoord][nElmt*nq] deformed, [ncoord][nElmt]
};

__host__ __device__ inline int seg_table_len(int nm, int nq) { return 2 * nm * nq + nq * nq + nq; }

hipError_t launch_seg(const SegArgs &a, hipStream_t s);

// Straight segments: d_verts nElmt x 2 x ncoord; J = |t|, df[c] = t_c / |t|^2 with
// t = (x1 - x0) / 2 (GeomFactors::ComputeDerivFactors for one expansion direction).
hipError_t geom_seg(const double *d_verts, int nElmt, int ncoord, int nq, int deformed,
                    double *d_jac, double *d_df, hipStream_t s);

}  // namespace imf
