// Geometry preparation entry points (implemented in mf_geom.hip).
#pragma once

#include "imf_common.hpp"

namespace imf {

// Pack (jac, df) in the un-interleaved CoalescedGeomData layouts into the per-point
// (deformed) or per-element (regular) streamed form the operator kernels read.
hipError_t pack_geometry(int kind, int shape, int deformed, int prec, long nElmt, int nq,
                         const double *w, const double *d_jac, const double *d_df,
                         void *d_out, hipStream_t s);

// Geometric factors of straight-sided hexes (shape = kHex) or quads (kQuad).
hipError_t geom_from_vertices(int shape, const double *d_verts, long nElmt, int nq,
                              const double *z, int deformed, double *d_jac, double *d_df,
                              hipStream_t s);

// The same for hexes with nq[d] points z[d] along direction d (point (i, j, k) at index
// (k*nq1 + j)*nq0 + i).
hipError_t geom_hex_aniso(const double *d_verts, long nElmt, const int *nq,
                          const double *const *z, int deformed, double *d_jac, double *d_df,
                          hipStream_t s);

// Straight-sided triangles (vertices nElmt x 3 x 2), Q points per element when deformed.
hipError_t geom_tri(const double *d_verts, long nElmt, int Q, int deformed, double *d_jac,
                    double *d_df, hipStream_t s);

// Straight-sided tetrahedra (vertices nElmt x 4 x 3), Q points per element when deformed.
hipError_t geom_tet(const double *d_verts, long nElmt, int Q, int deformed, double *d_jac,
                    double *d_df, hipStream_t s);

// Straight-sided pyramids (vertices nElmt x 5 x 3), nq[d] points z[d] along eta0 / eta1 /
// eta2 (point (i, j, k) at index (k*nq1 + j)*nq0 + i).
hipError_t geom_pyr(const double *d_verts, long nElmt, const int *nq, const double *const *z,
                    int deformed, double *d_jac, double *d_df, hipStream_t s);

// Straight-sided prisms (vertices nElmt x 6 x 3), nq[d] points z[d] along eta0 / xi1 / eta2
// (point (i, j, k) at index (k*nq1 + j)*nq0 + i).
hipError_t geom_prism(const double *d_verts, long nElmt, const int *nq, const double *const *z,
                      int deformed, double *d_jac, double *d_df, hipStream_t s);

}  // namespace imf
