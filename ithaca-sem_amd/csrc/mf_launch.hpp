// Launchers and the (modes, points) dispatch tables of the operator kernels.
//
// The reference dispatches on (modes, points) with nested switches over compile-time
// template instantiations (e.g. Helmholtz.h:669-760): nm in 2..8, nq in nm..2*nm.  The
// same table is generated here, one translation unit per nm so the build parallelises.
#pragma once

#include <cstring>
#include <type_traits>

#include "imf_common.hpp"
#include "mf_hex_kernels.hpp"
#include "mf_quad_kernels.hpp"

namespace imf {

// One kernel variant: forward/transposed operator of a kind with a geometry mode.
template <int NM, int NQ, typename T, int KIND, int DIR, int GEO>
hipError_t launch_hex(const LaunchArgs &a)
{
    using C = HexCfg<NQ>;
    Tab<NM, NQ, T> tab;
    std::memcpy(&tab, a.tab, sizeof(tab));
    const dim3 grid((unsigned)((a.nElmt + C::EPB - 1) / C::EPB)), block(C::BLOCK);
    const T *in = static_cast<const T *>(a.in);
    const T *geom = static_cast<const T *>(a.geom);
    T *out = static_cast<T *>(a.out);
    if constexpr (KIND == kHelmholtz) {
        if constexpr (std::is_same_v<T, double>) {
            if (a.gmap) {
                hex_helmholtz_kernel<NM, NQ, T, GEO, true><<<grid, block, 0, a.stream>>>(
                    tab, a.nElmt, T(a.lambda), in, geom, out, a.gmap, a.gsign);
                return hipGetLastError();
            }
        }
        if (a.gmap) return hipErrorInvalidValue;  // gathered input: FP64 only
        if (a.nvec > 1) {
            hex_helmholtz_kernel<NM, NQ, T, GEO, false, true><<<grid, block, 0, a.stream>>>(
                tab, a.nElmt, T(a.lambda), in, geom, out, nullptr, nullptr, a.nvec, a.ldin,
                a.ldout);
            return hipGetLastError();
        }
        hex_helmholtz_kernel<NM, NQ, T, GEO>
            <<<grid, block, 0, a.stream>>>(tab, a.nElmt, T(a.lambda), in, geom, out);
    } else if constexpr (KIND == kMass) {
        hex_mass_kernel<NM, NQ, T, GEO><<<grid, block, 0, a.stream>>>(
            tab, a.nElmt, a.nvec, in, a.ldin, geom, out, a.ldout);
    } else if constexpr (KIND == kIProdDeriv) {
        hex_iprod_deriv_kernel<NM, NQ, T, GEO>
            <<<grid, block, 0, a.stream>>>(tab, a.nElmt, in, geom, out);
    } else if constexpr ((KIND == kBwdTrans && DIR == kForward) ||
                         (KIND == kIProduct && DIR == kTranspose)) {
        hex_bwdtrans_kernel<NM, NQ, T, DIR, GEO>
            <<<grid, block, 0, a.stream>>>(tab, a.nElmt, in, geom, out);
    } else {
        hex_iproduct_kernel<NM, NQ, T, GEO>
            <<<grid, block, 0, a.stream>>>(tab, a.nElmt, in, geom, out);
    }
    return hipGetLastError();
}

template <int NM, int NQ, typename T, int KIND, int DIR, int GEO>
hipError_t launch_quad(const LaunchArgs &a)
{
    using C = QuadCfg<NQ>;
    Tab<NM, NQ, T> tab;
    std::memcpy(&tab, a.tab, sizeof(tab));
    const dim3 grid((unsigned)((a.nElmt + C::EPB - 1) / C::EPB)), block(C::BLOCK);
    const T *in = static_cast<const T *>(a.in);
    const T *geom = static_cast<const T *>(a.geom);
    T *out = static_cast<T *>(a.out);
    if constexpr (KIND == kHelmholtz) {
        quad_helmholtz_kernel<NM, NQ, T, GEO>
            <<<grid, block, 0, a.stream>>>(tab, a.nElmt, T(a.lambda), in, geom, out);
    } else if constexpr (KIND == kIProdDeriv) {
        quad_iprod_deriv_kernel<NM, NQ, T, GEO>
            <<<grid, block, 0, a.stream>>>(tab, a.nElmt, in, geom, out);
    } else if constexpr ((KIND == kBwdTrans && DIR == kForward) ||
                         (KIND == kIProduct && DIR == kTranspose)) {
        quad_bwdtrans_kernel<NM, NQ, T, DIR, GEO>
            <<<grid, block, 0, a.stream>>>(tab, a.nElmt, in, geom, out);
    } else {
        quad_iproduct_kernel<NM, NQ, T, GEO>
            <<<grid, block, 0, a.stream>>>(tab, a.nElmt, in, geom, out);
    }
    return hipGetLastError();
}

// Select the kernel variant for (kind, dir, geo):
//   Helmholtz            -> helmholtz<geo>                (the operator is symmetric)
//   BwdTrans  forward    -> bwdtrans<fwd>;  transpose -> iproduct<unit>
//   IProduct  forward    -> iproduct<geo>;  transpose -> bwdtrans<transpose, geo>
//   Mass (hex)           -> mass<geo>                     (the operator is symmetric)
//   IProductWRTDerivBase -> iprod_deriv<geo>              (forward only)
template <int SHAPE, int NM, int NQ, typename T>
LaunchFn pick(int kind, int dir, int geo)
{
    if (kind == kMass) {
        if constexpr (SHAPE == kHex) {
            if (geo == kDeformed) return &launch_hex<NM, NQ, T, kMass, kForward, kDeformed>;
            if (geo == kRegular) return &launch_hex<NM, NQ, T, kMass, kForward, kRegular>;
        }
        return nullptr;
    }
#define IMF_L(K, D, G)                                                        \
    (SHAPE == kHex ? (LaunchFn)&launch_hex<NM, NQ, T, K, D, G>                \
                   : (LaunchFn)&launch_quad<NM, NQ, T, K, D, G>)
    if (kind == kHelmholtz) {
        if (geo == kDeformed) return IMF_L(kHelmholtz, kForward, kDeformed);
        if (geo == kRegular) return IMF_L(kHelmholtz, kForward, kRegular);
        return nullptr;
    }
    if (kind == kIProdDeriv) {
        if (dir != kForward) return nullptr;
        if (geo == kDeformed) return IMF_L(kIProdDeriv, kForward, kDeformed);
        if (geo == kRegular) return IMF_L(kIProdDeriv, kForward, kRegular);
        return nullptr;
    }
    if (kind == kBwdTrans) {
        if (dir == kForward) return IMF_L(kBwdTrans, kForward, kUnit);
        return IMF_L(kIProduct, kForward, kUnit);
    }
    if (kind == kIProduct) {
        if (dir == kForward) {
            if (geo == kDeformed) return IMF_L(kIProduct, kForward, kDeformed);
            if (geo == kRegular) return IMF_L(kIProduct, kForward, kRegular);
            return nullptr;
        }
        if (geo == kDeformed) return IMF_L(kIProduct, kTranspose, kDeformed);
        if (geo == kRegular) return IMF_L(kIProduct, kTranspose, kRegular);
        return nullptr;
    }
    return nullptr;
#undef IMF_L
}

// Per-nm lookup: nq in nm..2*nm, FP64 (prec 0) or FP32 (prec 1).
template <int SHAPE, int NM>
LaunchFn lookup_nm(int kind, int dir, int geo, int nq, int prec)
{
    LaunchFn f = nullptr;
    auto one = [&](auto nqc) {
        constexpr int NQ = decltype(nqc)::value;
        if (nq == NQ)
            f = prec == 0 ? pick<SHAPE, NM, NQ, double>(kind, dir, geo)
                          : pick<SHAPE, NM, NQ, float>(kind, dir, geo);
    };
    [&]<int... I>(std::integer_sequence<int, I...>) {
        (one(std::integral_constant<int, NM + I>{}), ...);
    }(std::make_integer_sequence<int, NM + 1>{});
    return f;
}

// Defined in mf_inst_*.hip, one per (shape, nm).
LaunchFn lookup_hex(int nm, int kind, int dir, int geo, int nq, int prec);
LaunchFn lookup_quad(int nm, int kind, int dir, int geo, int nq, int prec);

// PhysDeriv depends on the points only: the reference dispatches on nq in 2..16
// (PhysDeriv.h:291-330, :541-572).  Defined in mf_inst_deriv.hip.
LaunchFn lookup_physderiv(int shape, int nq, int geo, int prec);

}  // namespace imf
