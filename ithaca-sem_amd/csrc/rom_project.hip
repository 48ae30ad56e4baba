// Galerkin projection of one affine operator term onto a reduced basis: C = V^T (A V).
//
// The reference forms its reduced operators densely, one term at a time
// (CoupledLinearNS_ROM.cpp:2788-2802: RB^T A RB per term; recombined per parameter at
// :4694-4704).  Here A is a matrix-free operator: A V is one multi-vector launch of the
// operator (geometry staged once per element, every basis vector streamed through it),
// and V^T (A V) is the split-K MFMA product of rom_gram.hip.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/ithaca_mf.h"
#include "../../include/ithaca_rom.h"
#include "mf_op.hpp"

namespace {
thread_local std::string g_perr;

int pfail(int code, const std::string &m)
{
    g_perr = m;
    return code;
}
}  // namespace

extern "C" {

const char *irom_project_last_error(void) { return g_perr.c_str(); }

int irom_galerkin_project(imf_op op, int r, const double *d_V, long ldv, double *d_W, long ldw,
                          double *d_C, long ldc, int mode, void *d_work, size_t work_bytes,
                          void *stream)
{
    if (!op || r <= 0 || !d_V || !d_W || !d_C || !d_work)
        return pfail(IMF_ERR_ARG, "irom_galerkin_project: bad argument");
    if (op->prec != IMF_FP64)
        return pfail(IMF_ERR_STATE, "irom_galerkin_project: FP64 operators only");
    if (op->kind != imf::kHelmholtz && op->kind != imf::kMass)
        return pfail(IMF_ERR_UNSUPPORTED,
                     "irom_galerkin_project: the term must map coefficients to coefficients "
                     "(Helmholtz or Mass)");
    const long N = (long)imf_in_size(op);
    if ((long)imf_out_size(op) != N || ldv < N || ldw < N || ldc < r)
        return pfail(IMF_ERR_ARG, "irom_galerkin_project: leading dimension smaller than the "
                                  "vector size");
    if (irom_gram_workspace_size(r, r, N) > work_bytes)
        return pfail(IMF_ERR_ARG, "irom_galerkin_project: workspace too small (see "
                                  "irom_gram_workspace_size(r, r, N))");
    if (int s = imf_apply_many_device(op, r, d_V, ldv, d_W, ldw, stream))
        return pfail(s, std::string("irom_galerkin_project: ") + imf_last_error());
    if (int s = irom_gram(mode, r, r, N, d_V, ldv, d_W, ldw, d_C, ldc, 0, d_work, work_bytes,
                          stream))
        return pfail(s, std::string("irom_galerkin_project: ") + irom_gram_last_error());
    return IMF_OK;
}

}  // extern "C"
