// (modes, points) dispatch over the per-nm instantiation units, mirroring the nested
// switch of the reference operators (e.g. MatrixFreeOps/Helmholtz.h:669-760).
#include "imf_common.hpp"

namespace imf {

#define IMF_DECL(S, N) LaunchFn lookup_##S##_nm##N(int, int, int, int, int);
IMF_DECL(hex, 2) IMF_DECL(hex, 3) IMF_DECL(hex, 4) IMF_DECL(hex, 5)
IMF_DECL(hex, 6) IMF_DECL(hex, 7) IMF_DECL(hex, 8)
IMF_DECL(quad, 2) IMF_DECL(quad, 3) IMF_DECL(quad, 4) IMF_DECL(quad, 5)
IMF_DECL(quad, 6) IMF_DECL(quad, 7) IMF_DECL(quad, 8)
#undef IMF_DECL

LaunchFn lookup_hex(int nm, int kind, int dir, int geo, int nq, int prec)
{
    switch (nm) {
    case 2: return lookup_hex_nm2(kind, dir, geo, nq, prec);
    case 3: return lookup_hex_nm3(kind, dir, geo, nq, prec);
    case 4: return lookup_hex_nm4(kind, dir, geo, nq, prec);
    case 5: return lookup_hex_nm5(kind, dir, geo, nq, prec);
    case 6: return lookup_hex_nm6(kind, dir, geo, nq, prec);
    case 7: return lookup_hex_nm7(kind, dir, geo, nq, prec);
    case 8: return lookup_hex_nm8(kind, dir, geo, nq, prec);
    default: return nullptr;
    }
}

LaunchFn lookup_quad(int nm, int kind, int dir, int geo, int nq, int prec)
{
    switch (nm) {
    case 2: return lookup_quad_nm2(kind, dir, geo, nq, prec);
    case 3: return lookup_quad_nm3(kind, dir, geo, nq, prec);
    case 4: return lookup_quad_nm4(kind, dir, geo, nq, prec);
    case 5: return lookup_quad_nm5(kind, dir, geo, nq, prec);
    case 6: return lookup_quad_nm6(kind, dir, geo, nq, prec);
    case 7: return lookup_quad_nm7(kind, dir, geo, nq, prec);
    case 8: return lookup_quad_nm8(kind, dir, geo, nq, prec);
    default: return nullptr;
    }
}

}  // namespace imf
