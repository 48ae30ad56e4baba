// (modes, points) instantiations: quad, nm = 4, nq = 4..8, FP64 and FP32.
#include "mf_launch.hpp"
namespace imf {
LaunchFn lookup_quad_nm4(int kind, int dir, int geo, int nq, int prec)
{
    return lookup_nm<kQuad, 4>(kind, dir, geo, nq, prec);
}
}  // namespace imf
