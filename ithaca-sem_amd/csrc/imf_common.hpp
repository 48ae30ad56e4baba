// Shared definitions of the gfx950 matrix-free operator library.
//
// Basis tables travel to the kernels by value as kernel arguments: every contraction
// loop is fully unrolled, so each table entry is addressed with a compile-time index,
// lands in the scalar cache and feeds v_fma_f64 as an SGPR operand.  No LDS bandwidth
// and no vector loads are spent on the 1D tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace imf {

// Operator kinds and element shapes handled by the library.
enum Kind : int {
    kBwdTrans = 0,
    kIProduct = 1,
    kHelmholtz = 2,
    kMass = 3,
    kPhysDeriv = 4,   // quadrature values -> dim physical derivatives
    kIProdDeriv = 5   // IProductWRTDerivBase: dim quadrature fields -> coefficients
};
enum Shape : int { kQuad = 2, kHex = 3, kTri = 12, kPrism = 13, kTet = 14, kPyr = 15, kSeg = 16 };  // collapsed shapes from kTri; kSeg: one-dimensional

// Geometry modes.  kUnit: no geometry (the Euclidean transposes).  kDeformed:
// per-point packed geometry.  kRegular: per-element packed geometry.
enum GeoMode : int { kUnit = 0, kDeformed = 1, kRegular = 2 };

// Direction of the apply: forward operator or its Euclidean transpose.
enum Dir : int { kForward = 0, kTranspose = 1 };

// 1D tables of one direction (the reference requires homogeneous modes/points).
//   B[p*NQ + i] = phi_p(z_i)        (Basis::GetBdata, Basis.cpp:392-421)
//   D[a*NQ + b] = l_a'(z_b)         (Points::GetD, GaussPoints.cpp:178-245)
//   w[i]                            (Points::GetW)
template <int NM, int NQ, typename T>
struct Tab {
    T B[NM * NQ];
    T D[NQ * NQ];
    T w[NQ];
};

// Collocation-derivative table of the points-only operators (PhysDeriv).
template <int NQ, typename T>
struct DTab {
    T D[NQ * NQ];
};

// Launch shape of the hex kernels: NQ*NQ threads own the NQ*NQ point columns of one
// element; small elements are packed several to a 64-lane wavefront.
template <int NQ>
struct HexCfg {
    static constexpr int TPE = NQ * NQ;
    static constexpr int EPB = (64 / TPE) > 0 ? 64 / TPE : 1;
    static constexpr int BLOCK = ((EPB * TPE + 63) / 64) * 64;
};

// Launch shape of the quad kernels: NQ threads own the NQ point lines of one element;
// 256-thread blocks hold 4 * floor(64 / NQ) elements.
template <int NQ>
struct QuadCfg {
    static constexpr int TPE = NQ;
    static constexpr int EPW = 64 / NQ;
    static constexpr int EPB = 4 * EPW;
    static constexpr int BLOCK = 256;
};

// Number of packed geometry components per point (deformed) or per element (regular).
//   Helmholtz hex : Jw, Jw*G00, Jw*G01, Jw*G02, Jw*G11, Jw*G12, Jw*G22
//   Helmholtz quad: Jw, Jw*G00, Jw*G01, Jw*G11
//   IProduct, Mass: Jw
//   PhysDeriv     : df[d]            (d = c*dim + a, the GetDerivFactors components)
//   IProductWRTDerivBase: Jw*df[d]
// with G^{ab} = sum_c df[c*dim+a] df[c*dim+b] (Helmholtz.h:858-966, :172-190) and
// Jw = J w_i w_j (w_k).  Regular elements store J instead of Jw (the kernel applies w).
constexpr int geo_comps(int kind, int shape)
{
    return kind == kHelmholtz ? (shape == kHex ? 7 : 4)
         : (kind == kPhysDeriv || kind == kIProdDeriv) ? (shape == kHex ? 9 : 4) : 1;
}

// Per-element stride (in scalars) of deformed packed geometry: the nc*Q values of an
// element padded to a whole number of 128-byte lines, so that no cache line is shared by
// two elements (which different workgroups read at different times).
__host__ __device__ constexpr long geo_stride(long nc_q, int elt_bytes)
{
    return ((nc_q * elt_bytes + 127) / 128) * 128 / elt_bytes;
}

// Launcher signature: everything an apply needs, already on the device.
struct LaunchArgs {
    const void *tab;   // host pointer to a Tab<NM,NQ,T> image
    int nElmt;
    double lambda;
    const void *geom;  // packed geometry (device)
    const void *in;    // device
    void *out;         // device
    hipStream_t stream;
    int nvec = 1;      // vectors per launch (Mass), vector v at in + v*ldin, out + v*ldout
    long ldin = 0, ldout = 0;
    // FP64 hex Helmholtz only: the input is a global (assembled) vector, gathered per local
    // mode as gsign[i] * in[gmap[i]] (GlobalToLocal fused into the operator's load)
    const int *gmap = nullptr;
    const signed char *gsign = nullptr;
};
using LaunchFn = hipError_t (*)(const LaunchArgs &);

}  // namespace imf
