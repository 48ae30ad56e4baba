// (modes, points) instantiations: hex, nm = 8, nq = 8..16, FP64 and FP32.
#include "mf_launch.hpp"
namespace imf {
LaunchFn lookup_hex_nm8(int kind, int dir, int geo, int nq, int prec)
{
    return lookup_nm<kHex, 8>(kind, dir, geo, nq, prec);
}
}  // namespace imf
