// (modes, points) instantiations: quad, nm = 3, nq = 3..6, FP64 and FP32.
#include "mf_launch.hpp"
namespace imf {
LaunchFn lookup_quad_nm3(int kind, int dir, int geo, int nq, int prec)
{
    return lookup_nm<kQuad, 3>(kind, dir, geo, nq, prec);
}
}  // namespace imf
