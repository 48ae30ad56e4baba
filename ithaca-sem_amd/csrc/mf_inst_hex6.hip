// (modes, points) instantiations: hex, nm = 6, nq = 6..12, FP64 and FP32.
#include "mf_launch.hpp"
namespace imf {
LaunchFn lookup_hex_nm6(int kind, int dir, int geo, int nq, int prec)
{
    return lookup_nm<kHex, 6>(kind, dir, geo, nq, prec);
}
}  // namespace imf
