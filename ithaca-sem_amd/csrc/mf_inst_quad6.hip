// (modes, points) instantiations: quad, nm = 6, nq = 6..12, FP64 and FP32.
#include "mf_launch.hpp"
namespace imf {
LaunchFn lookup_quad_nm6(int kind, int dir, int geo, int nq, int prec)
{
    return lookup_nm<kQuad, 6>(kind, dir, geo, nq, prec);
}
}  // namespace imf
