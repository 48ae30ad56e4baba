// (modes, points) instantiations: quad, nm = 7, nq = 7..14, FP64 and FP32.
#include "mf_launch.hpp"
namespace imf {
LaunchFn lookup_quad_nm7(int kind, int dir, int geo, int nq, int prec)
{
    return lookup_nm<kQuad, 7>(kind, dir, geo, nq, prec);
}
}  // namespace imf
