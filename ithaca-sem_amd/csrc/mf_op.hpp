// Internal definition of an operator handle, shared by the C-ABI translation units.
#pragma once

#include <cstddef>
#include <vector>

#include "imf_common.hpp"
#include "mf_aniso.hpp"
#include "mf_prism.hpp"
#include "mf_pyr.hpp"
#include "mf_tet.hpp"
#include "mf_tri.hpp"

struct imf_op_s {
    int kind = 0, shape = 0, deformed = 0, prec = 0, dim = 0, nm = 0, nq = 0;
    int ncoord = 0;  // IProductWRTDerivBase of triangles / quads in 3D space: 3 (0: the shape's dim)
    // quads with ncoord = 3: the 6 raw factors, and the two contracted fields that the packed
    // kernel (given identity factors) reads
    double *d_df3 = nullptr, *d_tmp3 = nullptr;
    size_t n_df3 = 0;
    int nElmt = 0;
    size_t M = 0, Q = 0;
    std::vector<double> w;                 // direction-0 weights (host)
    std::vector<double> bdata, dbdata;     // direction-0 tables (host, nm*nq)
    std::vector<unsigned char> tab;        // Tab<NM,NQ,T> image
    double lambda = 1.0;                   // Helmholtz default m_lambda(1.0)
    double *d_jac = nullptr, *d_df = nullptr;
    size_t n_jac = 0, n_df = 0;
    void *d_geom = nullptr;
    bool geom_ready = false;
    unsigned long geom_version = 0;        // process-wide unique id of the current geometry pack
    void *d_in = nullptr, *d_out = nullptr;  // staging for host applies
    // pipelined host applies: two pinned buffers each way, a copy and a compute stream
    void *pin_in[2] = {nullptr, nullptr}, *pin_out[2] = {nullptr, nullptr};
    size_t pin_bytes = 0;
    hipStream_t hs[2] = {nullptr, nullptr};
    hipEvent_t ev_h2d[2] = {nullptr, nullptr}, ev_kern[2] = {nullptr, nullptr},
               ev_d2h[2] = {nullptr, nullptr};
    imf::LaunchFn fwd = nullptr, tr = nullptr;
    // hexahedra with different modes / points per direction (mf_hex_aniso.hip, FP64):
    // per-direction sizes, the device table image (B, D, w of the three directions) and
    // the LDS work-buffer length; the geometry is read unpacked from d_jac / d_df
    bool aniso = false;
    int nmd[3] = {0, 0, 0}, nqd[3] = {0, 0, 0};
    double *d_atab = nullptr;
    int bufLen = 0;
    // collapsed-coordinate shapes (FP64): triangles (mf_tri.hip; nqd[0] / nqd[1] points along
    // eta0 / eta1), prisms (mf_prism.hip; nqd[0..2] along eta0 / xi1 / eta2), tetrahedra
    // (mf_tet.hip) and pyramids (mf_pyr.hip; nqd[0..2] along eta0 / eta1 / eta2), nmd[0]
    // modes per direction, the table image in d_atab; geometry read unpacked from d_jac / d_df
    bool collapsed = false;
};

// Kernel arguments of an anisotropic operator's apply (in, out, nvec vectors at ld strides).
imf::AnisoArgs aniso_args(const imf_op_s *op, const void *in, void *out, int nvec, long ldin,
                          long ldout);

// FP64 hex Helmholtz with its input gathered from a global vector: out = H(local), where
// local[i] = gsign[i] * glob[gmap[i]] (GlobalToLocal fused into the kernel's load).
// Library-internal (used by the assembled solver); defined in mf_capi.hip.
int imf_apply_gather_device(imf_op_s *op, const double *glob, const int *gmap,
                            const signed char *gsign, double *out, hipStream_t stream);
