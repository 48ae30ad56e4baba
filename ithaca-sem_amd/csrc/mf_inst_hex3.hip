// (modes, points) instantiations: hex, nm = 3, nq = 3..6, FP64 and FP32.
#include "mf_launch.hpp"
namespace imf {
LaunchFn lookup_hex_nm3(int kind, int dir, int geo, int nq, int prec)
{
    return lookup_nm<kHex, 3>(kind, dir, geo, nq, prec);
}
}  // namespace imf
