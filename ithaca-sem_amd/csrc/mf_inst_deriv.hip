// PhysDeriv instantiations: hex and quad, nq = 2..16, FP64 and FP32, deformed and regular.
#include "mf_launch.hpp"

namespace imf {

template <int SHAPE, int NQ, typename T, int GEO>
hipError_t launch_physderiv(const LaunchArgs &a)
{
    DTab<NQ, T> tab;
    std::memcpy(&tab, a.tab, sizeof(tab));
    const T *in = static_cast<const T *>(a.in);
    const T *df = static_cast<const T *>(a.geom);
    T *out = static_cast<T *>(a.out);
    if constexpr (SHAPE == kHex) {
        using C = HexCfg<NQ>;
        const dim3 grid((unsigned)((a.nElmt + C::EPB - 1) / C::EPB)), block(C::BLOCK);
        hex_physderiv_kernel<NQ, T, GEO><<<grid, block, 0, a.stream>>>(tab, a.nElmt, in, df, out);
    } else {
        using C = QuadCfg<NQ>;
        const dim3 grid((unsigned)((a.nElmt + C::EPB - 1) / C::EPB)), block(C::BLOCK);
        quad_physderiv_kernel<NQ, T, GEO><<<grid, block, 0, a.stream>>>(tab, a.nElmt, in, df, out);
    }
    return hipGetLastError();
}

LaunchFn lookup_physderiv(int shape, int nq, int geo, int prec)
{
    LaunchFn f = nullptr;
    auto one = [&](auto nqc) {
        constexpr int NQ = decltype(nqc)::value;
        if (nq != NQ) return;
        auto sel = [&](auto sc) -> LaunchFn {
            constexpr int S = decltype(sc)::value;
            if (geo == kDeformed)
                return prec == 0 ? &launch_physderiv<S, NQ, double, kDeformed>
                                 : &launch_physderiv<S, NQ, float, kDeformed>;
            if (geo == kRegular)
                return prec == 0 ? &launch_physderiv<S, NQ, double, kRegular>
                                 : &launch_physderiv<S, NQ, float, kRegular>;
            return nullptr;
        };
        f = shape == kHex ? sel(std::integral_constant<int, kHex>{})
                          : sel(std::integral_constant<int, kQuad>{});
    };
    [&]<int... I>(std::integer_sequence<int, I...>) {
        (one(std::integral_constant<int, 2 + I>{}), ...);
    }(std::make_integer_sequence<int, 15>{});
    return f;
}

}  // namespace imf
