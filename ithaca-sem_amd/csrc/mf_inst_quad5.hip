// (modes, points) instantiations: quad, nm = 5, nq = 5..10, FP64 and FP32.
#include "mf_launch.hpp"
namespace imf {
LaunchFn lookup_quad_nm5(int kind, int dir, int geo, int nq, int prec)
{
    return lookup_nm<kQuad, 5>(kind, dir, geo, nq, prec);
}
}  // namespace imf
