// (modes, points) instantiations: quad, nm = 2, nq = 2..4, FP64 and FP32.
#include "mf_launch.hpp"
namespace imf {
LaunchFn lookup_quad_nm2(int kind, int dir, int geo, int nq, int prec)
{
    return lookup_nm<kQuad, 2>(kind, dir, geo, nq, prec);
}
}  // namespace imf
