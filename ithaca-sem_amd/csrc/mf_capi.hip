// C ABI of the matrix-free operators (declared in include/ithaca_mf.h).
//
// An imf_op owns device copies of the caller's geometry (the un-interleaved GetJac /
// GetDerivFactors arrays, Collections/CoalescedGeomData.cpp:53-112, :251-313) and the
// packed streamed form derived from them, plus the 1D basis tables as a kernel-argument
// image.  Applies on device arrays allocate nothing.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ithaca_mf.h"
#include "imf_common.hpp"
#include "mf_geom.hpp"
#include "mf_launch.hpp"
#include "mf_op.hpp"
#include "mf_seg.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(IMF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

size_t elt_size(const imf_op op) { return op->prec == IMF_FP64 ? sizeof(double) : sizeof(float); }

size_t jac_len(const imf_op op) { return (size_t)op->nElmt * (op->deformed ? op->Q : 1); }
// segments: one factor per coordinate (dim is their coordinate dimension)
size_t df_len(const imf_op op)
{
    if (op->ncoord == 3) return 6 * jac_len(op);  // a triangle in 3D space: 3 x 2 factors
    return (size_t)op->dim * (op->shape == imf::kSeg ? 1 : op->dim) * jac_len(op);
}

size_t geom_len(const imf_op op)
{
    const int nc = imf::geo_comps(op->kind, op->shape);
    if ((op->kind == imf::kHelmholtz || op->kind == imf::kIProdDeriv) &&
        op->shape == imf::kHex && op->deformed)
        return (size_t)op->nElmt * imf::geo_stride((long)nc * op->Q, (int)elt_size(op));
    return (size_t)nc * jac_len(op);
}

bool needs_jac(const imf_op op)
{
    return op->kind != imf::kBwdTrans && op->kind != imf::kPhysDeriv;
}
bool needs_df(const imf_op op)
{
    return op->kind == imf::kHelmholtz || op->kind == imf::kPhysDeriv ||
           op->kind == imf::kIProdDeriv;
}

// Re-pack the streamed geometry once everything it depends on has been set.
int maybe_pack(imf_op op)
{
    op->geom_ready = false;
    if (!needs_jac(op) && !needs_df(op)) return IMF_OK;
    if (needs_jac(op) && !op->d_jac) return IMF_OK;
    if (needs_df(op) && !op->d_df) return IMF_OK;
    if (op->aniso || op->collapsed) {  // these kernels read GetJac / GetDerivFactors directly
        static std::atomic<unsigned long> next_id{1ul << 62};
        op->geom_ready = true;
        op->geom_version = next_id.fetch_add(1);
        return IMF_OK;
    }
    if (!op->d_geom) {
        hipError_t e = hipMalloc(&op->d_geom, geom_len(op) * elt_size(op));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(geometry)");
    }
    hipError_t e = imf::pack_geometry(op->kind, op->shape, op->deformed, op->prec, op->nElmt,
                                      op->nq, op->w.data(), op->d_jac, op->d_df, op->d_geom,
                                      nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return hip_fail(e, "pack_geometry");
    op->geom_ready = true;
    // process-wide id of this geometry pack: a cache keyed on it cannot mistake a new
    // operator allocated at a freed operator's address for the old one
    static std::atomic<unsigned long> next_geom_id{1};
    op->geom_version = next_geom_id.fetch_add(1);
    return IMF_OK;
}

int set_array(imf_op op, double **dst, size_t *ndst, const double *src, size_t n, size_t want,
              const char *name, hipMemcpyKind kind)
{
    if (!op || !src) return fail(IMF_ERR_ARG, std::string(name) + ": null argument");
    if (n != want)
        return fail(IMF_ERR_ARG, std::string(name) + ": expected " + std::to_string(want) +
                                     " values, got " + std::to_string(n));
    if (!*dst) {
        hipError_t e = hipMalloc((void **)dst, want * sizeof(double));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(geometry input)");
        *ndst = want;
    }
    hipError_t e = hipMemcpy(*dst, src, want * sizeof(double), kind);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(geometry input)");
    return maybe_pack(op);
}

template <typename T>
void fill_tab(imf_op op, const double *B, const double *D, const double *w)
{
    const int nm = op->nm, nq = op->nq;
    std::vector<T> img;
    if (op->kind == imf::kPhysDeriv) {  // DTab<NQ,T>: the derivative matrix only
        for (int i = 0; i < nq * nq; ++i) img.push_back(T(D[i]));
        op->tab.resize(img.size() * sizeof(T));
        std::memcpy(op->tab.data(), img.data(), op->tab.size());
        return;
    }
    img.reserve(nm * nq + nq * nq + nq);
    for (int i = 0; i < nm * nq; ++i) img.push_back(T(B[i]));
    for (int i = 0; i < nq * nq; ++i) img.push_back(T(D[i]));
    for (int i = 0; i < nq; ++i) img.push_back(T(w[i]));
    op->tab.resize(img.size() * sizeof(T));
    std::memcpy(op->tab.data(), img.data(), op->tab.size());
}

int check_op(const imf_op op)
{
    if (!op) return fail(IMF_ERR_ARG, "null operator handle");
    return IMF_OK;
}

// Stands for "the transpose" in run() for triangles and prisms, whose kernels take the
// direction as an argument instead of being separate launch functions; never called.
hipError_t tri_transpose_fn(const imf::LaunchArgs &) { return hipErrorInvalidValue; }
imf::LaunchFn tri_transpose_marker() { return &tri_transpose_fn; }

// Quadrilaterals in 3D space, IProductWRTDerivBase (IProductWRTDerivBase.h:545-, three
// inputs): y_a = sum_c df[2c+a] in_c per point, after which the packed kernel applies
// identity factors times w J.  One thread per point; df [6][nElmt*Q] deformed, [6][nElmt]
// regular.
__global__ void quad_coord3_contract(const double *__restrict__ in,
                                     const double *__restrict__ df, long nQ, int Q, int deformed,
                                     double *__restrict__ y)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nQ) return;
    const long np = deformed ? nQ : nQ / Q, k = deformed ? g : g / Q;
    const double x0 = in[g], x1 = in[nQ + g], x2 = in[2 * nQ + g];
    y[g] = fma(df[4 * np + k], x2, fma(df[2 * np + k], x1, df[k] * x0));
    y[nQ + g] = fma(df[5 * np + k], x2, fma(df[3 * np + k], x1, df[np + k] * x0));
}

int run(imf_op op, imf::LaunchFn fn, const void *in, void *out, void *stream, int nvec = 1,
        long ldin = 0, long ldout = 0, const int *gmap = nullptr,
        const signed char *gsign = nullptr)
{
    if ((needs_jac(op) || needs_df(op)) && !op->geom_ready)
        return fail(IMF_ERR_STATE, !needs_jac(op) ? "derivative factors not set"
                                   : needs_df(op) ? "Jacobian and derivative factors not set"
                                                  : "Jacobian not set");
    if (!in || !out) return fail(IMF_ERR_ARG, "null input/output array");
    if (op->collapsed && op->shape == imf::kSeg) {
        if (nvec != 1 || gmap) return fail(IMF_ERR_UNSUPPORTED, "segments: one vector per apply");
        imf::SegArgs t{};
        t.kind = op->kind;
        t.transpose = fn == tri_transpose_marker() ? 1 : 0;
        t.geo = op->kind == imf::kBwdTrans ? imf::kUnit : (op->deformed ? imf::kDeformed : imf::kRegular);
        t.nm = op->nm;
        t.nq = op->nq;
        t.nElmt = op->nElmt;
        t.ncoord = op->dim;
        t.tab = op->d_atab;
        t.in = (const double *)in;
        t.out = (double *)out;
        t.jac = op->d_jac;
        t.df = op->d_df;
        hipError_t e = imf::launch_seg(t, (hipStream_t)stream);
        if (e != hipSuccess) return hip_fail(e, "kernel launch");
        return IMF_OK;
    }
    if (op->collapsed && op->shape == imf::kPyr) {
        if (nvec != 1 || gmap) return fail(IMF_ERR_UNSUPPORTED, "pyramids: one vector per apply");
        imf::PyrArgs t{};
        t.kind = op->kind;
        t.transpose = fn == tri_transpose_marker() ? 1 : 0;
        t.geo = op->kind == imf::kBwdTrans ? imf::kUnit : (op->deformed ? imf::kDeformed : imf::kRegular);
        t.nm = op->nmd[0];
        t.nq0 = op->nqd[0];
        t.nq1 = op->nqd[1];
        t.nq2 = op->nqd[2];
        t.nElmt = op->nElmt;
        t.lambda = op->lambda;
        t.tab = op->d_atab;
        t.in = (const double *)in;
        t.out = (double *)out;
        t.jac = op->d_jac;
        t.df = op->d_df;
        hipError_t e = imf::launch_pyr(t, (hipStream_t)stream);
        if (e != hipSuccess) return hip_fail(e, "kernel launch");
        return IMF_OK;
    }
    if (op->collapsed && op->shape == imf::kTet) {
        if (nvec != 1 || gmap) return fail(IMF_ERR_UNSUPPORTED, "tetrahedra: one vector per apply");
        imf::TetArgs t{};
        t.kind = op->kind;
        t.transpose = fn == tri_transpose_marker() ? 1 : 0;
        t.geo = op->kind == imf::kBwdTrans ? imf::kUnit : (op->deformed ? imf::kDeformed : imf::kRegular);
        t.nm = op->nmd[0];
        t.nq0 = op->nqd[0];
        t.nq1 = op->nqd[1];
        t.nq2 = op->nqd[2];
        t.nElmt = op->nElmt;
        t.lambda = op->lambda;
        t.tab = op->d_atab;
        t.in = (const double *)in;
        t.out = (double *)out;
        t.jac = op->d_jac;
        t.df = op->d_df;
        hipError_t e = imf::launch_tet(t, (hipStream_t)stream);
        if (e != hipSuccess) return hip_fail(e, "kernel launch");
        return IMF_OK;
    }
    if (op->collapsed && op->shape == imf::kPrism) {
        if (nvec != 1 || gmap) return fail(IMF_ERR_UNSUPPORTED, "prisms: one vector per apply");
        imf::PrismArgs t{};
        t.kind = op->kind;
        t.transpose = fn == tri_transpose_marker() ? 1 : 0;
        t.geo = op->kind == imf::kBwdTrans ? imf::kUnit : (op->deformed ? imf::kDeformed : imf::kRegular);
        t.nm = op->nmd[0];
        t.nq0 = op->nqd[0];
        t.nq1 = op->nqd[1];
        t.nq2 = op->nqd[2];
        t.nElmt = op->nElmt;
        t.lambda = op->lambda;
        t.tab = op->d_atab;
        t.in = (const double *)in;
        t.out = (double *)out;
        t.jac = op->d_jac;
        t.df = op->d_df;
        hipError_t e = imf::launch_prism(t, (hipStream_t)stream);
        if (e != hipSuccess) return hip_fail(e, "kernel launch");
        return IMF_OK;
    }
    if (op->collapsed) {
        if (nvec != 1 || gmap) return fail(IMF_ERR_UNSUPPORTED, "triangles: one vector per apply");
        imf::TriArgs t{};
        t.kind = op->kind;
        t.transpose = fn == tri_transpose_marker() ? 1 : 0;
        t.geo = op->kind == imf::kBwdTrans ? imf::kUnit : (op->deformed ? imf::kDeformed : imf::kRegular);
        t.nm = op->nmd[0];
        t.nq0 = op->nqd[0];
        t.nq1 = op->nqd[1];
        t.nElmt = op->nElmt;
        t.ncoord = op->ncoord == 3 ? 3 : 2;
        t.lambda = op->lambda;
        t.tab = op->d_atab;
        t.in = (const double *)in;
        t.out = (double *)out;
        t.jac = op->d_jac;
        t.df = op->d_df;
        hipError_t e = imf::launch_tri(t, (hipStream_t)stream);
        if (e != hipSuccess) return hip_fail(e, "kernel launch");
        return IMF_OK;
    }
    if (op->aniso) {
        if (gmap) return fail(IMF_ERR_UNSUPPORTED, "gathered input: isotropic operators only");
        hipError_t e = imf::launch_aniso(aniso_args(op, in, out, nvec, ldin, ldout),
                                         (hipStream_t)stream);
        if (e != hipSuccess) return hip_fail(e, "kernel launch");
        return IMF_OK;
    }
    if (op->ncoord == 3) {  // quads in 3D space: contract the three fields first
        if (nvec != 1 || gmap) return fail(IMF_ERR_UNSUPPORTED, "3D coordinates: one vector per apply");
        const long nQ = (long)op->nElmt * op->Q;
        quad_coord3_contract<<<(unsigned)((nQ + 255) / 256), 256, 0, (hipStream_t)stream>>>(
            (const double *)in, op->d_df3, nQ, (int)op->Q, op->deformed, op->d_tmp3);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "kernel launch");
        in = op->d_tmp3;
    }
    imf::LaunchArgs a{op->tab.data(), op->nElmt, op->lambda, op->d_geom, in, out,
                      (hipStream_t)stream, nvec, ldin, ldout, gmap, gsign};
    hipError_t e = fn(a);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return IMF_OK;
}

// Operators whose directions differ in modes, points or tables (FP64 hexahedra): the
// runtime-size kernels of mf_hex_aniso.hip.  Each direction needs 2 <= nm <= nq <= 16.
int create_aniso(const std::string &k, int kind, int shape, int deformed, const int *nm,
                 const int *nq, const double *const *bdata, const double *const *dbdata,
                 const double *const *D, const double *const *w, int nElmt, int precision,
                 imf_op *out)
{
    if (shape != imf::kHex || precision != IMF_FP64 ||
        (kind != imf::kBwdTrans && kind != imf::kIProduct && kind != imf::kHelmholtz &&
         kind != imf::kMass))
        return fail(IMF_ERR_UNSUPPORTED, k + ": different modes / points per direction are "
                                             "supported for FP64 hexahedral BwdTrans, IProduct, "
                                             "Helmholtz and Mass operators");
    for (int d = 0; d < 3; ++d)
        if (nm[d] < 2 || nq[d] < nm[d] || nq[d] > 16 || !bdata[d] || !D[d] || !w[d])
            return fail(IMF_ERR_UNSUPPORTED, k + ": # of modes / points combo not initialised.");
    long buf = 1;
    for (int d = 0; d < 3; ++d) buf *= std::max(nm[d], nq[d]);
    const size_t lds = sizeof(double) * (imf::aniso_table_len(nm, nq) + 4 * (size_t)buf);
    if (lds > 150 * 1024)
        return fail(IMF_ERR_UNSUPPORTED, k + ": element too large for the per-direction kernels");
    std::vector<double> tab;
    for (int d = 0; d < 3; ++d) tab.insert(tab.end(), bdata[d], bdata[d] + (size_t)nm[d] * nq[d]);
    for (int d = 0; d < 3; ++d) tab.insert(tab.end(), D[d], D[d] + (size_t)nq[d] * nq[d]);
    for (int d = 0; d < 3; ++d) tab.insert(tab.end(), w[d], w[d] + nq[d]);
    imf_op op = new imf_op_s();
    op->kind = kind; op->shape = shape; op->deformed = deformed; op->prec = precision;
    op->dim = 3; op->nm = nm[0]; op->nq = nq[0]; op->nElmt = nElmt;
    op->aniso = true;
    op->M = 1; op->Q = 1;
    for (int d = 0; d < 3; ++d) {
        op->nmd[d] = nm[d];
        op->nqd[d] = nq[d];
        op->M *= nm[d];
        op->Q *= nq[d];
    }
    op->bufLen = (int)buf;
    op->w.assign(w[0], w[0] + nq[0]);
    op->bdata.assign(bdata[0], bdata[0] + (size_t)nm[0] * nq[0]);
    if (dbdata && dbdata[0]) op->dbdata.assign(dbdata[0], dbdata[0] + (size_t)nm[0] * nq[0]);
    hipError_t e = hipMalloc(&op->d_atab, tab.size() * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpy(op->d_atab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(op->d_atab);
        delete op;
        return hip_fail(e, "anisotropic basis tables");
    }
    *out = op;
    return IMF_OK;
}

// Segments: one Modified_A direction (mf_seg.hip), FP64.  bdata/dbdata nm*nq, D nq^2, w nq
// (GLL); ncoord (the dim argument) is the coordinate dimension of the mesh the segments lie
// in: PhysDeriv gives, and IProductWRTDerivBase takes, ncoord arrays of nElmt*nq values.
int create_seg(const std::string &k, int kind, int deformed, int ncoord, int m, int q,
               const double *const *bdata, const double *const *dbdata,
               const double *const *D, const double *const *w, int nElmt, int precision,
               imf_op *out)
{
    if (precision != IMF_FP64 ||
        (kind != imf::kBwdTrans && kind != imf::kIProduct && kind != imf::kPhysDeriv &&
         kind != imf::kIProdDeriv))
        return fail(IMF_ERR_UNSUPPORTED, k + ": segments are supported for FP64 BwdTrans, "
                                             "IProduct, PhysDeriv and IProductWRTDerivBase");
    if (m < 1 || m > 16 || q < 1 || q > 32 || (kind != imf::kPhysDeriv && q < m))
        return fail(IMF_ERR_UNSUPPORTED, k + ": # of modes / points combo not implemented.");
    const bool tabs = kind != imf::kPhysDeriv, deriv = kind == imf::kIProdDeriv;
    if (!D[0] || !w[0] || (tabs && !bdata[0]) || (deriv && (!dbdata || !dbdata[0])))
        return fail(IMF_ERR_ARG, k + ": missing basis tables");
    std::vector<double> tab;
    auto put = [&](const double *p, size_t n) {
        if (p) tab.insert(tab.end(), p, p + n);
        else tab.insert(tab.end(), n, 0.0);
    };
    put(tabs ? bdata[0] : nullptr, (size_t)m * q);
    put(deriv ? dbdata[0] : nullptr, (size_t)m * q);
    put(D[0], (size_t)q * q);
    put(w[0], q);
    imf_op op = new imf_op_s();
    op->kind = kind; op->shape = imf::kSeg; op->deformed = deformed; op->prec = precision;
    op->dim = ncoord; op->nm = m; op->nq = q; op->nElmt = nElmt;
    op->collapsed = true;
    op->nmd[0] = m;
    op->nqd[0] = q;
    op->M = (size_t)m;
    op->Q = (size_t)q;
    op->w.assign(w[0], w[0] + q);
    hipError_t e = hipMalloc(&op->d_atab, tab.size() * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpy(op->d_atab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(op->d_atab);
        delete op;
        return hip_fail(e, "segment basis tables");
    }
    *out = op;
    return IMF_OK;
}

// Triangles: Modified_A x Modified_B in collapsed coordinates (mf_tri.hip), FP64.
// Direction 0: bdata/dbdata nm*nq0, D nq0^2, w nq0 (GLL); direction 1: bdata/dbdata
// nm(nm+1)/2 * nq1, D nq1^2, w nq1 (the Gauss-Radau alpha = 1 weights of the basis; the
// factor 1/2 of the collapse is applied here, as the reference's Helper does).  The points
// themselves (needed by the metric and PhysDeriv) are regenerated for GLL / Gauss-Radau.
int create_tri(const std::string &k, int kind, int deformed, const int *nm, const int *nq,
               const double *const *bdata, const double *const *dbdata,
               const double *const *D, const double *const *w, int nElmt, int precision,
               imf_op *out)
{
    if (precision != IMF_FP64 ||
        (kind != imf::kBwdTrans && kind != imf::kIProduct && kind != imf::kHelmholtz &&
         kind != imf::kPhysDeriv && kind != imf::kIProdDeriv))
        return fail(IMF_ERR_UNSUPPORTED, k + ": triangles are supported for FP64 BwdTrans, "
                                             "IProduct, Helmholtz, PhysDeriv and IProductWRTDerivBase");
    const int m = nm[0], q0 = nq[0], q1 = nq[1];
    if (m < 2 || m > 8 || q0 < 2 || q1 < 1 || q0 > 16 || q1 > 16 ||
        (kind != imf::kPhysDeriv && (nm[1] != m || q0 < m || q1 < m - 1)))
        return fail(IMF_ERR_UNSUPPORTED, k + ": # of modes / points combo not implemented.");
    const bool helm = kind == imf::kHelmholtz || kind == imf::kIProdDeriv;
    if (!D[0] || !D[1] || !w[0] || !w[1] ||
        (kind != imf::kPhysDeriv && (!bdata[0] || !bdata[1])) ||
        (helm && (!dbdata || !dbdata[0] || !dbdata[1])))
        return fail(IMF_ERR_ARG, k + ": missing basis tables");
    const int M = imf::tri_modes(m);
    std::vector<double> z0(q0), z1(q1), scratch;
    {
        std::vector<double> ww(16), dd(256), bb(16 * 16 * 16), db(16 * 16 * 16);
        if (imf_basis_modified_a(2, q0, z0.data(), ww.data(), dd.data(), bb.data(), db.data()) ||
            imf_basis_modified_b(2, q1, z1.data(), ww.data(), dd.data(), bb.data(), db.data()))
            return fail(IMF_ERR_ARG, k + ": points");
    }
    std::vector<double> tab;
    tab.reserve(imf::tri_table_len(m, q0, q1));
    auto put = [&](const double *p, size_t n) {
        if (p) tab.insert(tab.end(), p, p + n);
        else tab.insert(tab.end(), n, 0.0);
    };
    const bool tabs = kind != imf::kPhysDeriv;
    put(tabs ? bdata[0] : nullptr, (size_t)m * q0);
    put(helm ? dbdata[0] : nullptr, (size_t)m * q0);
    put(D[0], (size_t)q0 * q0);
    put(w[0], q0);
    put(z0.data(), q0);
    put(tabs ? bdata[1] : nullptr, (size_t)M * q1);
    put(helm ? dbdata[1] : nullptr, (size_t)M * q1);
    put(D[1], (size_t)q1 * q1);
    for (int j = 0; j < q1; ++j) tab.push_back(0.5 * w[1][j]);
    put(z1.data(), q1);
    imf_op op = new imf_op_s();
    op->kind = kind; op->shape = imf::kTri; op->deformed = deformed; op->prec = precision;
    op->dim = 2; op->nm = m; op->nq = q0; op->nElmt = nElmt;
    op->collapsed = true;
    op->nmd[0] = op->nmd[1] = m;
    op->nqd[0] = q0;
    op->nqd[1] = q1;
    op->M = (size_t)M;
    op->Q = (size_t)q0 * q1;
    op->w.assign(w[0], w[0] + q0);
    hipError_t e = hipMalloc(&op->d_atab, tab.size() * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpy(op->d_atab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(op->d_atab);
        delete op;
        return hip_fail(e, "triangle basis tables");
    }
    *out = op;
    return IMF_OK;
}

// Prisms: Modified_A x Modified_A x Modified_B in collapsed coordinates (mf_prism.hip),
// FP64.  Directions 0 and 1: bdata/dbdata nm*nq[d], D nq[d]^2, w nq[d] (GLL); direction 2:
// bdata/dbdata nm(nm+1)/2 * nq[2], D nq[2]^2, w nq[2] (the Gauss-Radau alpha = 1 weights;
// the factor 1/2 of the collapse is applied here).  The points are regenerated for GLL /
// Gauss-Radau.  One workgroup holds an element's arrays in LDS, so the sizes are also
// bounded by the device's LDS per workgroup.
int create_prism(const std::string &k, int kind, int deformed, const int *nm, const int *nq,
                 const double *const *bdata, const double *const *dbdata,
                 const double *const *D, const double *const *w, int nElmt, int precision,
                 imf_op *out)
{
    if (precision != IMF_FP64 ||
        (kind != imf::kBwdTrans && kind != imf::kIProduct && kind != imf::kHelmholtz &&
         kind != imf::kPhysDeriv && kind != imf::kIProdDeriv))
        return fail(IMF_ERR_UNSUPPORTED, k + ": prisms are supported for FP64 BwdTrans, "
                                             "IProduct, Helmholtz, PhysDeriv and IProductWRTDerivBase");
    const int m = nm[0], q0 = nq[0], q1 = nq[1], q2 = nq[2];
    if (m < 2 || m > 8 || q0 < 2 || q1 < 2 || q2 < 1 || q0 > 16 || q1 > 16 || q2 > 16 ||
        (kind != imf::kPhysDeriv &&
         (nm[1] != m || nm[2] != m || q0 < m || q1 < m || q2 < m - 1)))
        return fail(IMF_ERR_UNSUPPORTED, k + ": # of modes / points combo not implemented.");
    int dev = 0, lds_max = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return hip_fail(e, "LDS size query");
    if (sizeof(double) * (size_t)imf::prism_lds_len(m, q0, q1, q2) > (size_t)lds_max)
        return fail(IMF_ERR_UNSUPPORTED, k + ": an element of this size does not fit the LDS "
                                             "of one workgroup");
    const bool helm = kind == imf::kHelmholtz || kind == imf::kIProdDeriv, tabs = kind != imf::kPhysDeriv;
    if (!D[0] || !D[1] || !D[2] || !w[0] || !w[1] || !w[2] ||
        (tabs && (!bdata[0] || !bdata[1] || !bdata[2])) ||
        (helm && (!dbdata || !dbdata[0] || !dbdata[1] || !dbdata[2])))
        return fail(IMF_ERR_ARG, k + ": missing basis tables");
    const int T = imf::tri_modes(m);
    std::vector<double> z0(q0), z1(q1), z2(q2);
    {
        std::vector<double> ww(16), dd(256), bb(16 * 16 * 16), db(16 * 16 * 16);
        if (imf_basis_modified_a(2, q0, z0.data(), ww.data(), dd.data(), bb.data(), db.data()) ||
            imf_basis_modified_a(2, q1, z1.data(), ww.data(), dd.data(), bb.data(), db.data()) ||
            imf_basis_modified_b(2, q2, z2.data(), ww.data(), dd.data(), bb.data(), db.data()))
            return fail(IMF_ERR_ARG, k + ": points");
    }
    std::vector<double> tab;
    tab.reserve(imf::prism_table_len(m, q0, q1, q2));
    auto put = [&](const double *p, size_t n) {
        if (p) tab.insert(tab.end(), p, p + n);
        else tab.insert(tab.end(), n, 0.0);
    };
    const int qd[2] = {q0, q1};
    const double *zd[2] = {z0.data(), z1.data()};
    for (int d = 0; d < 2; ++d) {
        put(tabs ? bdata[d] : nullptr, (size_t)m * qd[d]);
        put(helm ? dbdata[d] : nullptr, (size_t)m * qd[d]);
        put(D[d], (size_t)qd[d] * qd[d]);
        put(w[d], qd[d]);
        put(zd[d], qd[d]);
    }
    put(tabs ? bdata[2] : nullptr, (size_t)T * q2);
    put(helm ? dbdata[2] : nullptr, (size_t)T * q2);
    put(D[2], (size_t)q2 * q2);
    for (int j = 0; j < q2; ++j) tab.push_back(0.5 * w[2][j]);
    put(z2.data(), q2);
    imf_op op = new imf_op_s();
    op->kind = kind; op->shape = imf::kPrism; op->deformed = deformed; op->prec = precision;
    op->dim = 3; op->nm = m; op->nq = q0; op->nElmt = nElmt;
    op->collapsed = true;
    op->nmd[0] = op->nmd[1] = op->nmd[2] = m;
    op->nqd[0] = q0;
    op->nqd[1] = q1;
    op->nqd[2] = q2;
    op->M = (size_t)imf::prism_modes(m);
    op->Q = (size_t)q0 * q1 * q2;
    op->w.assign(w[0], w[0] + q0);
    e = hipMalloc(&op->d_atab, tab.size() * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpy(op->d_atab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(op->d_atab);
        delete op;
        return hip_fail(e, "prism basis tables");
    }
    *out = op;
    return IMF_OK;
}

// Tetrahedra: Modified_A x Modified_B x Modified_C in collapsed coordinates (mf_tet.hip),
// FP64.  Direction 0: bdata/dbdata nm*nq[0], D, w (GLL); direction 1: nm(nm+1)/2 * nq[1]
// (Modified_B, Gauss-Radau alpha = 1); direction 2: nm(nm+1)(nm+2)/6 * nq[2] (Modified_C,
// Gauss-Radau alpha = 2).  The collapse factors 1/2 and 1/4 of the weights are applied here
// and the points regenerated; one element's arrays must fit the LDS of a workgroup.
int create_tet(const std::string &k, int kind, int deformed, const int *nm, const int *nq,
               const double *const *bdata, const double *const *dbdata,
               const double *const *D, const double *const *w, int nElmt, int precision,
               imf_op *out)
{
    if (precision != IMF_FP64 ||
        (kind != imf::kBwdTrans && kind != imf::kIProduct && kind != imf::kHelmholtz &&
         kind != imf::kPhysDeriv && kind != imf::kIProdDeriv))
        return fail(IMF_ERR_UNSUPPORTED, k + ": tetrahedra are supported for FP64 BwdTrans, "
                                             "IProduct, Helmholtz, PhysDeriv and IProductWRTDerivBase");
    const int m = nm[0], q0 = nq[0], q1 = nq[1], q2 = nq[2];
    if (m < 2 || m > 8 || q0 < 2 || q1 < 1 || q2 < 1 || q0 > 16 || q1 > 16 || q2 > 16 ||
        (kind != imf::kPhysDeriv &&
         (nm[1] != m || nm[2] != m || q0 < m || q1 < m - 1 || q2 < m - 1)))
        return fail(IMF_ERR_UNSUPPORTED, k + ": # of modes / points combo not implemented.");
    int dev = 0, lds_max = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return hip_fail(e, "LDS size query");
    if (sizeof(double) * (size_t)imf::tet_lds_len(m, q0, q1, q2) > (size_t)lds_max)
        return fail(IMF_ERR_UNSUPPORTED, k + ": an element of this size does not fit the LDS "
                                             "of one workgroup");
    const bool helm = kind == imf::kHelmholtz || kind == imf::kIProdDeriv, tabs = kind != imf::kPhysDeriv;
    if (!D[0] || !D[1] || !D[2] || !w[0] || !w[1] || !w[2] ||
        (tabs && (!bdata[0] || !bdata[1] || !bdata[2])) ||
        (helm && (!dbdata || !dbdata[0] || !dbdata[1] || !dbdata[2])))
        return fail(IMF_ERR_ARG, k + ": missing basis tables");
    std::vector<double> z0(q0), z1(q1), z2(q2);
    {
        std::vector<double> ww(16), dd(256), bb(4 * 16), db(4 * 16);
        if (imf_basis_modified_a(2, q0, z0.data(), ww.data(), dd.data(), bb.data(), db.data()) ||
            imf_basis_modified_b(2, q1, z1.data(), ww.data(), dd.data(), bb.data(), db.data()) ||
            imf_basis_modified_c(2, q2, z2.data(), ww.data(), dd.data(), bb.data(), db.data()))
            return fail(IMF_ERR_ARG, k + ": points");
    }
    const size_t rows[3] = {(size_t)m, (size_t)imf::tri_modes(m), (size_t)imf::tet_modes(m)};
    const int qd[3] = {q0, q1, q2};
    const double fac[3] = {1.0, 0.5, 0.25};
    const double *zd[3] = {z0.data(), z1.data(), z2.data()};
    std::vector<double> tab;
    tab.reserve(imf::tet_table_len(m, q0, q1, q2));
    auto put = [&](const double *p, size_t n) {
        if (p) tab.insert(tab.end(), p, p + n);
        else tab.insert(tab.end(), n, 0.0);
    };
    for (int d = 0; d < 3; ++d) {
        put(tabs ? bdata[d] : nullptr, rows[d] * qd[d]);
        put(helm ? dbdata[d] : nullptr, rows[d] * qd[d]);
        put(D[d], (size_t)qd[d] * qd[d]);
        for (int j = 0; j < qd[d]; ++j) tab.push_back(fac[d] * w[d][j]);
        put(zd[d], qd[d]);
    }
    imf_op op = new imf_op_s();
    op->kind = kind; op->shape = imf::kTet; op->deformed = deformed; op->prec = precision;
    op->dim = 3; op->nm = m; op->nq = q0; op->nElmt = nElmt;
    op->collapsed = true;
    op->nmd[0] = op->nmd[1] = op->nmd[2] = m;
    op->nqd[0] = q0;
    op->nqd[1] = q1;
    op->nqd[2] = q2;
    op->M = (size_t)imf::tet_modes(m);
    op->Q = (size_t)q0 * q1 * q2;
    op->w.assign(w[0], w[0] + q0);
    e = hipMalloc(&op->d_atab, tab.size() * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpy(op->d_atab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(op->d_atab);
        delete op;
        return hip_fail(e, "tetrahedron basis tables");
    }
    *out = op;
    return IMF_OK;
}

// Pyramids: Modified_A x Modified_A x ModifiedPyr_C in collapsed coordinates (mf_pyr.hip),
// FP64.  Directions 0 and 1: bdata/dbdata nm*nq[d], D, w (GLL); direction 2:
// nm(nm+1)(2nm+1)/6 * nq[2] (ModifiedPyr_C, Gauss-Radau alpha = 2; the factor 1/4 of the
// weights applied here).  The points are regenerated; one element's arrays must fit the
// LDS of a workgroup.
int create_pyr(const std::string &k, int kind, int deformed, const int *nm, const int *nq,
               const double *const *bdata, const double *const *dbdata,
               const double *const *D, const double *const *w, int nElmt, int precision,
               imf_op *out)
{
    if (precision != IMF_FP64 ||
        (kind != imf::kBwdTrans && kind != imf::kIProduct && kind != imf::kHelmholtz &&
         kind != imf::kPhysDeriv && kind != imf::kIProdDeriv))
        return fail(IMF_ERR_UNSUPPORTED, k + ": pyramids are supported for FP64 BwdTrans, "
                                             "IProduct, Helmholtz, PhysDeriv and IProductWRTDerivBase");
    const int m = nm[0], q0 = nq[0], q1 = nq[1], q2 = nq[2];
    if (m < 2 || m > 8 || q0 < 2 || q1 < 2 || q2 < 1 || q0 > 16 || q1 > 16 || q2 > 16 ||
        (kind != imf::kPhysDeriv &&
         (nm[1] != m || nm[2] != m || q0 < m || q1 < m || q2 < m - 1)))
        return fail(IMF_ERR_UNSUPPORTED, k + ": # of modes / points combo not implemented.");
    int dev = 0, lds_max = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return hip_fail(e, "LDS size query");
    if (sizeof(double) * (size_t)imf::pyr_lds_len(m, q0, q1, q2) > (size_t)lds_max)
        return fail(IMF_ERR_UNSUPPORTED, k + ": an element of this size does not fit the LDS "
                                             "of one workgroup");
    const bool helm = kind == imf::kHelmholtz || kind == imf::kIProdDeriv, tabs = kind != imf::kPhysDeriv;
    if (!D[0] || !D[1] || !D[2] || !w[0] || !w[1] || !w[2] ||
        (tabs && (!bdata[0] || !bdata[1] || !bdata[2])) ||
        (helm && (!dbdata || !dbdata[0] || !dbdata[1] || !dbdata[2])))
        return fail(IMF_ERR_ARG, k + ": missing basis tables");
    std::vector<double> z0(q0), z1(q1), z2(q2);
    {
        std::vector<double> ww(16), dd(256), bb(5 * 16), db(5 * 16);
        if (imf_basis_modified_a(2, q0, z0.data(), ww.data(), dd.data(), bb.data(), db.data()) ||
            imf_basis_modified_a(2, q1, z1.data(), ww.data(), dd.data(), bb.data(), db.data()) ||
            imf_basis_modified_pyr_c(2, q2, z2.data(), ww.data(), dd.data(), bb.data(), db.data()))
            return fail(IMF_ERR_ARG, k + ": points");
    }
    const size_t rows[3] = {(size_t)m, (size_t)m, (size_t)imf::pyr_modes(m)};
    const int qd[3] = {q0, q1, q2};
    const double fac[3] = {1.0, 1.0, 0.25};
    const double *zd[3] = {z0.data(), z1.data(), z2.data()};
    std::vector<double> tab;
    tab.reserve(imf::pyr_table_len(m, q0, q1, q2));
    auto put = [&](const double *p, size_t n) {
        if (p) tab.insert(tab.end(), p, p + n);
        else tab.insert(tab.end(), n, 0.0);
    };
    for (int d = 0; d < 3; ++d) {
        put(tabs ? bdata[d] : nullptr, rows[d] * qd[d]);
        put(helm ? dbdata[d] : nullptr, rows[d] * qd[d]);
        put(D[d], (size_t)qd[d] * qd[d]);
        for (int j = 0; j < qd[d]; ++j) tab.push_back(fac[d] * w[d][j]);
        put(zd[d], qd[d]);
    }
    imf_op op = new imf_op_s();
    op->kind = kind; op->shape = imf::kPyr; op->deformed = deformed; op->prec = precision;
    op->dim = 3; op->nm = m; op->nq = q0; op->nElmt = nElmt;
    op->collapsed = true;
    op->nmd[0] = op->nmd[1] = op->nmd[2] = m;
    op->nqd[0] = q0;
    op->nqd[1] = q1;
    op->nqd[2] = q2;
    op->M = (size_t)imf::pyr_modes(m);
    op->Q = (size_t)q0 * q1 * q2;
    op->w.assign(w[0], w[0] + q0);
    e = hipMalloc(&op->d_atab, tab.size() * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpy(op->d_atab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(op->d_atab);
        delete op;
        return hip_fail(e, "pyramid basis tables");
    }
    *out = op;
    return IMF_OK;
}

}  // namespace

imf::AnisoArgs aniso_args(const imf_op_s *op, const void *in, void *out, int nvec, long ldin,
                          long ldout)
{
    imf::AnisoArgs a{};
    a.kind = op->kind;
    a.geo = op->kind == imf::kBwdTrans ? imf::kUnit : (op->deformed ? imf::kDeformed : imf::kRegular);
    for (int d = 0; d < 3; ++d) {
        a.nm[d] = op->nmd[d];
        a.nq[d] = op->nqd[d];
    }
    a.nElmt = op->nElmt;
    a.bufLen = op->bufLen;
    a.lambda = op->lambda;
    a.tab = op->d_atab;
    a.in = (const double *)in;
    a.out = (double *)out;
    a.jac = op->d_jac;
    a.df = op->d_df;
    a.nvec = nvec;
    a.ldin = ldin;
    a.ldout = ldout;
    return a;
}

int imf_apply_gather_device(imf_op_s *op, const double *glob, const int *gmap,
                            const signed char *gsign, double *out, hipStream_t stream)
{
    if (int s = check_op(op)) return s;
    if (op->kind != imf::kHelmholtz || op->shape != imf::kHex || op->prec != IMF_FP64 ||
        op->aniso || !gmap || !gsign)
        return fail(IMF_ERR_UNSUPPORTED, "gathered input: FP64 Helmholtz_Hex only");
    return run(op, op->fwd, glob, out, stream, 1, 0, 0, gmap, gsign);
}

extern "C" {

const char *imf_last_error(void) { return g_err.c_str(); }

int imf_create(const char *key, int dim, const int *nm, const int *nq,
               const double *const *bdata, const double *const *dbdata,
               const double *const *D, const double *const *w, int nElmt, int precision,
               imf_op *out)
{
    if (!key || !nm || !nq || !bdata || !D || !w || !out)
        return fail(IMF_ERR_ARG, "imf_create: null argument");
    // dbdata = D bdata is applied in factored form by the kernels; a copy is kept for
    // the assembled diagonal (mf_assembly.hip)
    *out = nullptr;
    const std::string k(key);
    int kind, shape, deformed;
    auto starts = [&](const char *s) { return k.rfind(s, 0) == 0; };
    if (starts("BwdTrans_")) kind = imf::kBwdTrans;
    else if (starts("IProduct_")) kind = imf::kIProduct;
    else if (starts("Helmholtz_")) kind = imf::kHelmholtz;
    else if (starts("Mass_")) kind = imf::kMass;
    else if (starts("PhysDeriv_")) kind = imf::kPhysDeriv;
    else if (starts("IProductWRTDerivBase_")) kind = imf::kIProdDeriv;
    else return fail(IMF_ERR_UNSUPPORTED, "No such operator key: " + k);
    const std::string rest = k.substr(k.find('_') + 1);
    if (rest == "Quad_Regular") { shape = imf::kQuad; deformed = 0; }
    else if (rest == "Quad_Deformed") { shape = imf::kQuad; deformed = 1; }
    else if (rest == "Hex_Regular") { shape = imf::kHex; deformed = 0; }
    else if (rest == "Hex_Deformed") { shape = imf::kHex; deformed = 1; }
    else if (rest == "Tri_Regular") { shape = imf::kTri; deformed = 0; }
    else if (rest == "Tri_Deformed") { shape = imf::kTri; deformed = 1; }
    else if (rest == "Prism_Regular") { shape = imf::kPrism; deformed = 0; }
    else if (rest == "Prism_Deformed") { shape = imf::kPrism; deformed = 1; }
    else if (rest == "Tet_Regular") { shape = imf::kTet; deformed = 0; }
    else if (rest == "Tet_Deformed") { shape = imf::kTet; deformed = 1; }
    else if (rest == "Pyr_Regular") { shape = imf::kPyr; deformed = 0; }
    else if (rest == "Pyr_Deformed") { shape = imf::kPyr; deformed = 1; }
    else if (rest == "Seg_Regular") { shape = imf::kSeg; deformed = 0; }
    else if (rest == "Seg_Deformed") { shape = imf::kSeg; deformed = 1; }
    else return fail(IMF_ERR_UNSUPPORTED, "No such operator key: " + k);
    // The reference registers BwdTrans only as _Regular (MatrixFreeOps/BwdTrans.cpp:7-33).
    if (kind == imf::kBwdTrans && deformed)
        return fail(IMF_ERR_UNSUPPORTED, "No such operator key: " + k);
    if (kind == imf::kMass && shape != imf::kHex)
        return fail(IMF_ERR_UNSUPPORTED, "No such operator key: " + k);
    if (shape == imf::kSeg) {
        // the reference registers no segment Helmholtz (MatrixFreeOps/Helmholtz.cpp)
        if (kind == imf::kHelmholtz) return fail(IMF_ERR_UNSUPPORTED, "No such operator key: " + k);
        if (dim < 1 || dim > 3)
            return fail(IMF_ERR_ARG, "imf_create: dim of " + k + " is the coordinate dimension, 1 to 3");
        if (nElmt <= 0) return fail(IMF_ERR_ARG, "imf_create: nElmt must be positive");
        return create_seg(k, kind, deformed, dim, nm[0], nq[0], bdata, dbdata, D, w, nElmt,
                          precision, out);
    }
    if (shape == imf::kTri || shape == imf::kPrism || shape == imf::kTet || shape == imf::kPyr) {
        if (dim != (shape == imf::kTri ? 2 : 3))
            return fail(IMF_ERR_ARG, "imf_create: dim does not match the shape of " + k);
        if (precision != IMF_FP64 && precision != IMF_FP32)
            return fail(IMF_ERR_ARG, "imf_create: unknown precision");
        if (nElmt <= 0) return fail(IMF_ERR_ARG, "imf_create: nElmt must be positive");
        if (shape == imf::kPyr)
            return create_pyr(k, kind, deformed, nm, nq, bdata, dbdata, D, w, nElmt, precision,
                              out);
        if (shape == imf::kTet)
            return create_tet(k, kind, deformed, nm, nq, bdata, dbdata, D, w, nElmt, precision,
                              out);
        if (shape == imf::kPrism)
            return create_prism(k, kind, deformed, nm, nq, bdata, dbdata, D, w, nElmt, precision,
                                out);
        return create_tri(k, kind, deformed, nm, nq, bdata, dbdata, D, w, nElmt, precision, out);
    }
    if (dim != shape)
        return fail(IMF_ERR_ARG, "imf_create: dim does not match the shape of " + k);
    if (precision != IMF_FP64 && precision != IMF_FP32)
        return fail(IMF_ERR_ARG, "imf_create: unknown precision");
    if (nElmt <= 0) return fail(IMF_ERR_ARG, "imf_create: nElmt must be positive");
    bool same = true;
    for (int d = 1; d < dim; ++d) {
        if (nm[d] != nm[0] || nq[d] != nq[0]) {
            same = false;
            break;
        }
        const size_t nb = (size_t)nm[0] * nq[0];
        if (std::memcmp(bdata[d], bdata[0], nb * sizeof(double)) ||
            std::memcmp(D[d], D[0], (size_t)nq[0] * nq[0] * sizeof(double)) ||
            std::memcmp(w[d], w[0], (size_t)nq[0] * sizeof(double)))
            same = false;
    }
    if (!same) return create_aniso(k, kind, shape, deformed, nm, nq, bdata, dbdata, D, w,
                                   nElmt, precision, out);
    const int m = nm[0], q = nq[0];
    const char *oname = kind == imf::kHelmholtz ? "Helmholtz" : kind == imf::kIProduct ? "IProduct"
                      : kind == imf::kMass ? "Mass" : kind == imf::kPhysDeriv ? "PhysDeriv"
                      : kind == imf::kIProdDeriv ? "IProductWRTDB" : "BwdTrans";
    const bool combo_ok = kind == imf::kPhysDeriv ? (q >= 2 && q <= 16)
                                                 : (m >= 2 && m <= 8 && q >= m && q <= 2 * m);
    if (!combo_ok)
        return fail(IMF_ERR_UNSUPPORTED, std::string(oname) + (shape == imf::kHex ? "Hex" : "Quad") +
                                             ": # of modes / points combo not initialised.");
    imf_op op = new imf_op_s();
    op->kind = kind; op->shape = shape; op->deformed = deformed; op->prec = precision;
    op->dim = dim; op->nm = m; op->nq = q; op->nElmt = nElmt;
    op->M = 1; op->Q = 1;
    for (int d = 0; d < dim; ++d) { op->M *= m; op->Q *= q; }
    op->w.assign(w[0], w[0] + q);
    op->bdata.assign(bdata[0], bdata[0] + (size_t)m * q);
    if (dbdata && dbdata[0]) op->dbdata.assign(dbdata[0], dbdata[0] + (size_t)m * q);
    if (precision == IMF_FP64) fill_tab<double>(op, bdata[0], D[0], w[0]);
    else fill_tab<float>(op, bdata[0], D[0], w[0]);
    const int geo = kind == imf::kBwdTrans ? imf::kUnit : (deformed ? imf::kDeformed : imf::kRegular);
    // BwdTrans^T uses no geometry; IProduct^T uses the IProduct geometry.
    const int geo_t = kind == imf::kBwdTrans ? imf::kUnit : geo;
    if (kind == imf::kPhysDeriv) {
        op->fwd = imf::lookup_physderiv(shape, q, geo, precision);
        op->tr = nullptr;
    } else if (shape == imf::kHex) {
        op->fwd = imf::lookup_hex(m, kind, imf::kForward, geo, q, precision);
        op->tr = imf::lookup_hex(m, kind, imf::kTranspose, geo_t, q, precision);
    } else {
        op->fwd = imf::lookup_quad(m, kind, imf::kForward, geo, q, precision);
        op->tr = imf::lookup_quad(m, kind, imf::kTranspose, geo_t, q, precision);
    }
    if (kind == imf::kMass) op->tr = op->fwd;  // symmetric
    if (!op->fwd) {
        delete op;
        return fail(IMF_ERR_UNSUPPORTED, k + ": # of modes / points combo not initialised.");
    }
    *out = op;
    return IMF_OK;
}

int imf_destroy(imf_op op)
{
    if (!op) return IMF_OK;
    (void)hipFree(op->d_jac);
    (void)hipFree(op->d_df);
    (void)hipFree(op->d_df3);
    (void)hipFree(op->d_tmp3);
    (void)hipFree(op->d_geom);
    (void)hipFree(op->d_in);
    (void)hipFree(op->d_out);
    (void)hipFree(op->d_atab);
    for (int b = 0; b < 2; ++b) {
        (void)hipHostFree(op->pin_in[b]);
        (void)hipHostFree(op->pin_out[b]);
        if (op->ev_h2d[b]) (void)hipEventDestroy(op->ev_h2d[b]);
        if (op->ev_kern[b]) (void)hipEventDestroy(op->ev_kern[b]);
        if (op->ev_d2h[b]) (void)hipEventDestroy(op->ev_d2h[b]);
        if (op->hs[b]) (void)hipStreamDestroy(op->hs[b]);
    }
    delete op;
    return IMF_OK;
}

double imf_ndof(imf_op op) { return op ? double(op->M) * op->nElmt : 0.0; }
int imf_needs_jac(imf_op op) { return op && needs_jac(op) ? 1 : 0; }
int imf_needs_df(imf_op op) { return op && needs_df(op) ? 1 : 0; }

size_t imf_in_size(imf_op op)
{
    if (!op) return 0;
    switch (op->kind) {
    case imf::kIProduct:
    case imf::kPhysDeriv: return (size_t)op->nElmt * op->Q;
    case imf::kIProdDeriv: return (size_t)(op->ncoord ? op->ncoord : op->dim) * op->nElmt * op->Q;
    default: return (size_t)op->nElmt * op->M;
    }
}

size_t imf_out_size(imf_op op)
{
    if (!op) return 0;
    switch (op->kind) {
    case imf::kBwdTrans: return (size_t)op->nElmt * op->Q;
    case imf::kPhysDeriv: return (size_t)op->dim * op->nElmt * op->Q;
    default: return (size_t)op->nElmt * op->M;
    }
}

int imf_set_jac(imf_op op, const double *jac, size_t n)
{
    if (int s = check_op(op)) return s;
    return set_array(op, &op->d_jac, &op->n_jac, jac, n, jac_len(op), "SetJac",
                     hipMemcpyHostToDevice);
}

// Quadrilaterals in 3D space: keep the six raw factors for the contraction and hand the
// packing identity factors (1, 0, 0, 1), so that the packed geometry is w J per point.
int set_df_quad3(imf_op op, const double *df, size_t n, hipMemcpyKind kind)
{
    if (int s = set_array(op, &op->d_df3, &op->n_df3, df, n, df_len(op), "SetDF", kind)) return s;
    const size_t np = jac_len(op);
    if (!op->d_tmp3) {
        hipError_t e = hipMalloc(&op->d_tmp3, 2 * (size_t)op->nElmt * op->Q * sizeof(double));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(contracted fields)");
    }
    std::vector<double> id(4 * np, 0.0);
    std::fill(id.begin(), id.begin() + np, 1.0);
    std::fill(id.begin() + 3 * np, id.end(), 1.0);
    return set_array(op, &op->d_df, &op->n_df, id.data(), 4 * np, 4 * np, "SetDF",
                     hipMemcpyHostToDevice);
}

int imf_set_df(imf_op op, const double *df, size_t n)
{
    if (int s = check_op(op)) return s;
    if (op->ncoord == 3 && op->shape == imf::kQuad)
        return set_df_quad3(op, df, n, hipMemcpyHostToDevice);
    return set_array(op, &op->d_df, &op->n_df, df, n, df_len(op), "SetDF",
                     hipMemcpyHostToDevice);
}

int imf_set_jac_device(imf_op op, const double *d_jac, size_t n)
{
    if (int s = check_op(op)) return s;
    return set_array(op, &op->d_jac, &op->n_jac, d_jac, n, jac_len(op), "SetJac",
                     hipMemcpyDeviceToDevice);
}

int imf_set_df_device(imf_op op, const double *d_df, size_t n)
{
    if (int s = check_op(op)) return s;
    if (op->ncoord == 3 && op->shape == imf::kQuad)
        return set_df_quad3(op, d_df, n, hipMemcpyDeviceToDevice);
    return set_array(op, &op->d_df, &op->n_df, d_df, n, df_len(op), "SetDF",
                     hipMemcpyDeviceToDevice);
}

int imf_set_coord_dim(imf_op op, int ncoord)
{
    if (int s = check_op(op)) return s;
    if ((op->shape != imf::kTri && op->shape != imf::kQuad) || op->kind != imf::kIProdDeriv ||
        op->prec != IMF_FP64 || op->aniso)
        return fail(IMF_ERR_UNSUPPORTED, "SetCoordDim: only the FP64 IProductWRTDerivBase of "
                                         "triangles and quadrilaterals takes 3D coordinates");
    if (ncoord != 2 && ncoord != 3) return fail(IMF_ERR_ARG, "SetCoordDim: ncoord must be 2 or 3");
    if (op->d_df) return fail(IMF_ERR_ARG, "SetCoordDim: call before SetDF");
    op->ncoord = ncoord == 3 ? 3 : 0;
    return IMF_OK;
}

int imf_set_lambda(imf_op op, double lambda)
{
    if (int s = check_op(op)) return s;
    op->lambda = lambda;
    return IMF_OK;
}

int imf_apply_device(imf_op op, const void *d_in, void *d_out, void *stream)
{
    if (int s = check_op(op)) return s;
    return run(op, op->fwd, d_in, d_out, stream);
}

int imf_apply_transpose_device(imf_op op, const void *d_in, void *d_out, void *stream)
{
    if (int s = check_op(op)) return s;
    if (op->aniso && (op->kind == imf::kHelmholtz || op->kind == imf::kMass))
        return run(op, op->fwd, d_in, d_out, stream);
    if (op->collapsed) {
        if (op->kind == imf::kHelmholtz) return run(op, op->fwd, d_in, d_out, stream);
        if (op->kind == imf::kPhysDeriv || op->kind == imf::kIProdDeriv)
            return fail(IMF_ERR_UNSUPPORTED, "transpose not available");
        return run(op, tri_transpose_marker(), d_in, d_out, stream);
    }
    if (!op->tr) return fail(IMF_ERR_UNSUPPORTED, "transpose not available");
    if (op->kind == imf::kHelmholtz || op->kind == imf::kMass)
        return run(op, op->fwd, d_in, d_out, stream);
    return run(op, op->tr, d_in, d_out, stream);
}

int imf_apply_many_device(imf_op op, int nvec, const void *d_in, long ld_in, void *d_out,
                          long ld_out, void *stream)
{
    if (int s = check_op(op)) return s;
    if (nvec < 0) return fail(IMF_ERR_ARG, "imf_apply_many_device: nvec < 0");
    if (nvec == 0) return IMF_OK;
    if (ld_in < (long)imf_in_size(op) || ld_out < (long)imf_out_size(op))
        return fail(IMF_ERR_ARG, "imf_apply_many_device: leading dimension smaller than the "
                                 "vector size");
    if (op->aniso || op->kind == imf::kMass ||
        (op->kind == imf::kHelmholtz && op->shape == imf::kHex))
        return run(op, op->fwd, d_in, d_out, stream, nvec, ld_in, ld_out);
    const size_t es = elt_size(op);
    for (int v = 0; v < nvec; ++v)
        if (int s = run(op, op->fwd, (const char *)d_in + (size_t)v * ld_in * es,
                        (char *)d_out + (size_t)v * ld_out * es, stream))
            return s;
    return IMF_OK;
}

}  // extern "C"

namespace {

// Host memcpy on a small persistent pool of threads (pageable user arrays to and from the
// pinned staging buffers): one thread copies at about a third of what the PCIe link
// moves, so the copies are split over the pool.
class CopyPool {
public:
    static CopyPool &get()
    {
        static CopyPool p;
        return p;
    }
    void copy(void *dst, const void *src, size_t bytes)
    {
        const int n = (int)workers_.size() + 1;
        if (bytes < (1u << 20) || n == 1) {
            std::memcpy(dst, src, bytes);
            return;
        }
        const size_t part = (bytes / n + 4095) / 4096 * 4096;
        std::unique_lock<std::mutex> lk(mu_);
        dst_ = (char *)dst;
        src_ = (const char *)src;
        bytes_ = bytes;
        part_ = part;
        pending_ = n - 1;
        ++gen_;
        cv_.notify_all();
        lk.unlock();
        chunk(0);
        lk.lock();
        done_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    CopyPool()
    {
        const unsigned hc = std::thread::hardware_concurrency();
        const int nw = (int)std::min(15u, hc > 1 ? hc - 1 : 0u);
        for (int i = 0; i < nw; ++i) workers_.emplace_back([this, i] { run(i + 1); });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto &t : workers_) t.join();
    }
    void chunk(int k)
    {
        const size_t b = std::min(bytes_, (size_t)k * part_), e = std::min(bytes_, b + part_);
        if (e > b) std::memcpy(dst_ + b, src_ + b, e - b);
    }
    void run(int k)
    {
        unsigned long seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            if (stop_) return;
            lk.unlock();
            chunk(k);
            lk.lock();
            if (--pending_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0, part_ = 0;
    int pending_ = 0;
    unsigned long gen_ = 0;
    bool stop_ = false;
};

// Element-major operators (inputs, outputs and packed geometry laid out element by
// element) can be applied to a range of elements by offsetting the three pointers.
bool element_major(const imf_op op)
{
    return !op->aniso && !op->collapsed && op->fwd && op->prec == IMF_FP64 &&
           (op->kind == imf::kBwdTrans || op->kind == imf::kIProduct ||
            op->kind == imf::kHelmholtz || op->kind == imf::kMass);
}

// imf_apply for element-major operators: the elements in NC chunks through two pinned
// buffers each way; chunk c's host copy into pinned memory, its H2D copy and kernel (one
// stream) and its D2H copy (a second stream) overlap the neighbouring chunks' steps.
int apply_pipelined(imf_op op, const double *in, double *out)
{
    const size_t nin = imf_in_size(op), nout = imf_out_size(op);
    const size_t ine = nin / op->nElmt, oute = nout / op->nElmt;
    const size_t gstride = (needs_jac(op) || needs_df(op)) ? geom_len(op) / op->nElmt : 0;
    constexpr int NC = 16;
    const long per = (op->nElmt + NC - 1) / NC;
    const size_t need = (size_t)per * std::max(ine, oute) * sizeof(double);
    hipError_t e = hipSuccess;
    if (op->pin_bytes < need) {
        for (int b = 0; b < 2; ++b) {
            (void)hipHostFree(op->pin_in[b]);
            (void)hipHostFree(op->pin_out[b]);
            op->pin_in[b] = op->pin_out[b] = nullptr;
        }
        op->pin_bytes = 0;
        for (int b = 0; b < 2 && e == hipSuccess; ++b) {
            e = hipHostMalloc(&op->pin_in[b], need, hipHostMallocDefault);
            if (e == hipSuccess) e = hipHostMalloc(&op->pin_out[b], need, hipHostMallocDefault);
        }
        if (e != hipSuccess) return hip_fail(e, "hipHostMalloc(staging)");
        op->pin_bytes = need;
    }
    if (!op->hs[0]) {
        for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipStreamCreateWithFlags(&op->hs[i], hipStreamNonBlocking);
        for (int b = 0; b < 2 && e == hipSuccess; ++b) {
            e = hipEventCreateWithFlags(&op->ev_h2d[b], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&op->ev_kern[b], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&op->ev_d2h[b], hipEventDisableTiming);
        }
        if (e != hipSuccess) return hip_fail(e, "streams / events");
    }
    // the streams are non-blocking: finish what earlier calls queued (e.g. the geometry pack)
    if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(e, "synchronize");
    CopyPool &pool = CopyPool::get();
    const int nc = (int)((op->nElmt + per - 1) / per);
    for (int c = 0; c <= nc; ++c) {
        if (c < nc) {
            const int b = c & 1;
            const long e0 = (long)c * per, ne = std::min<long>(per, op->nElmt - e0);
            if (c >= 2 && (e = hipEventSynchronize(op->ev_h2d[b])) != hipSuccess)
                return hip_fail(e, "staging");
            pool.copy(op->pin_in[b], in + e0 * ine, ne * ine * sizeof(double));
            double *din = (double *)op->d_in + e0 * ine, *dout = (double *)op->d_out + e0 * oute;
            if ((e = hipMemcpyAsync(din, op->pin_in[b], ne * ine * sizeof(double),
                                    hipMemcpyHostToDevice, op->hs[0])) != hipSuccess ||
                (e = hipEventRecord(op->ev_h2d[b], op->hs[0])) != hipSuccess)
                return hip_fail(e, "H2D");
            const void *geom = gstride ? (const void *)((const double *)op->d_geom + e0 * gstride)
                                       : (const void *)op->d_geom;
            imf::LaunchArgs a{op->tab.data(), (int)ne, op->lambda, geom, din, dout, op->hs[0]};
            if ((e = op->fwd(a)) != hipSuccess) return hip_fail(e, "kernel launch");
            if ((e = hipEventRecord(op->ev_kern[b], op->hs[0])) != hipSuccess ||
                (e = hipStreamWaitEvent(op->hs[1], op->ev_kern[b], 0)) != hipSuccess ||
                (e = hipMemcpyAsync(op->pin_out[b], dout, ne * oute * sizeof(double),
                                    hipMemcpyDeviceToHost, op->hs[1])) != hipSuccess ||
                (e = hipEventRecord(op->ev_d2h[b], op->hs[1])) != hipSuccess)
                return hip_fail(e, "D2H");
        }
        if (c >= 1) {  // the previous chunk's result to the caller
            const int b = (c - 1) & 1;
            const long e0 = (long)(c - 1) * per, ne = std::min<long>(per, op->nElmt - e0);
            if ((e = hipEventSynchronize(op->ev_d2h[b])) != hipSuccess) return hip_fail(e, "D2H");
            pool.copy(out + e0 * oute, op->pin_out[b], ne * oute * sizeof(double));
        }
    }
    return IMF_OK;
}

}  // namespace

extern "C" {

int imf_apply(imf_op op, const double *in, double *out)
{
    if (int s = check_op(op)) return s;
    if (op->prec != IMF_FP64) return fail(IMF_ERR_STATE, "imf_apply: host applies are FP64 only");
    if (!in || !out) return fail(IMF_ERR_ARG, "null input/output array");
    const size_t nin = imf_in_size(op), nout = imf_out_size(op);
    hipError_t e;
    if (!op->d_in) {
        if ((e = hipMalloc(&op->d_in, nin * sizeof(double))) != hipSuccess)
            return hip_fail(e, "hipMalloc(staging)");
        if ((e = hipMalloc(&op->d_out, nout * sizeof(double))) != hipSuccess)
            return hip_fail(e, "hipMalloc(staging)");
    }
    if (element_major(op)) {
        if ((needs_jac(op) || needs_df(op)) && !op->geom_ready)
            return fail(IMF_ERR_STATE, !needs_jac(op) ? "derivative factors not set"
                                       : needs_df(op) ? "Jacobian and derivative factors not set"
                                                      : "Jacobian not set");
        return apply_pipelined(op, in, out);
    }
    if ((e = hipMemcpy(op->d_in, in, nin * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(in)");
    if (int s = run(op, op->fwd, op->d_in, op->d_out, nullptr)) return s;
    if ((e = hipMemcpy(out, op->d_out, nout * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(out)");
    return IMF_OK;
}

int imf_geom_hex_aniso_device(const double *d_verts, int nElmt, const int *nq,
                              const double *const *z, int deformed, double *d_jac, double *d_df,
                              void *stream)
{
    if (!d_verts || !nq || !z || !d_jac || !d_df || nElmt <= 0)
        return fail(IMF_ERR_ARG, "imf_geom_hex_aniso_device: bad argument");
    for (int d = 0; d < 3; ++d)
        if (nq[d] < 2 || nq[d] > 16 || !z[d])
            return fail(IMF_ERR_ARG, "imf_geom_hex_aniso_device: bad points");
    hipError_t e = imf::geom_hex_aniso(d_verts, nElmt, nq, z, deformed, d_jac, d_df,
                                       (hipStream_t)stream);
    return e == hipSuccess ? IMF_OK : hip_fail(e, "geom_hex_aniso");
}

int imf_geom_hex_device(const double *d_verts, int nElmt, int nq, const double *z, int deformed,
                        double *d_jac, double *d_df, void *stream)
{
    if (!d_verts || !z || !d_jac || !d_df || nElmt <= 0 || nq < 2 || nq > 16)
        return fail(IMF_ERR_ARG, "imf_geom_hex_device: bad argument");
    hipError_t e = imf::geom_from_vertices(imf::kHex, d_verts, nElmt, nq, z, deformed, d_jac,
                                           d_df, (hipStream_t)stream);
    return e == hipSuccess ? IMF_OK : hip_fail(e, "geom_hex");
}

int imf_geom_quad_device(const double *d_verts, int nElmt, int nq, const double *z, int deformed,
                         double *d_jac, double *d_df, void *stream)
{
    if (!d_verts || !z || !d_jac || !d_df || nElmt <= 0 || nq < 2 || nq > 16)
        return fail(IMF_ERR_ARG, "imf_geom_quad_device: bad argument");
    hipError_t e = imf::geom_from_vertices(imf::kQuad, d_verts, nElmt, nq, z, deformed, d_jac,
                                           d_df, (hipStream_t)stream);
    return e == hipSuccess ? IMF_OK : hip_fail(e, "geom_quad");
}

int imf_geom_tri_device(const double *d_verts, int nElmt, int nq0, int nq1, int deformed,
                        double *d_jac, double *d_df, void *stream)
{
    if (!d_verts || !d_jac || !d_df || nElmt <= 0 || nq0 < 1 || nq1 < 1 || nq0 > 64 || nq1 > 64)
        return fail(IMF_ERR_ARG, "imf_geom_tri_device: bad argument");
    hipError_t e = imf::geom_tri(d_verts, nElmt, nq0 * nq1, deformed, d_jac, d_df,
                                 (hipStream_t)stream);
    return e == hipSuccess ? IMF_OK : hip_fail(e, "geom_tri");
}

int imf_geom_seg_device(const double *d_verts, int nElmt, int ncoord, int nq, int deformed,
                        double *d_jac, double *d_df, void *stream)
{
    if (!d_verts || !d_jac || !d_df || nElmt <= 0 || ncoord < 1 || ncoord > 3 || nq < 1 ||
        nq > 64)
        return fail(IMF_ERR_ARG, "imf_geom_seg_device: bad argument");
    hipError_t e = imf::geom_seg(d_verts, nElmt, ncoord, nq, deformed, d_jac, d_df,
                                 (hipStream_t)stream);
    return e == hipSuccess ? IMF_OK : hip_fail(e, "geom_seg");
}

int imf_geom_pyr_device(const double *d_verts, int nElmt, int nq0, int nq1, int nq2,
                        int deformed, double *d_jac, double *d_df, void *stream)
{
    if (!d_verts || !d_jac || !d_df || nElmt <= 0 || nq0 < 2 || nq1 < 2 || nq2 < 1 ||
        nq0 > 16 || nq1 > 16 || nq2 > 16)
        return fail(IMF_ERR_ARG, "imf_geom_pyr_device: bad argument");
    // nm = 2: at most 5 * 16 table values per direction
    std::vector<double> z0(nq0), z1(nq1), z2(nq2), ww(16), dd(256), bb(80), db(80);
    if (imf_basis_modified_a(2, nq0, z0.data(), ww.data(), dd.data(), bb.data(), db.data()) ||
        imf_basis_modified_a(2, nq1, z1.data(), ww.data(), dd.data(), bb.data(), db.data()) ||
        imf_basis_modified_pyr_c(2, nq2, z2.data(), ww.data(), dd.data(), bb.data(), db.data()))
        return fail(IMF_ERR_ARG, "imf_geom_pyr_device: points");
    const double *z[3] = {z0.data(), z1.data(), z2.data()};
    const int nq[3] = {nq0, nq1, nq2};
    hipError_t e = imf::geom_pyr(d_verts, nElmt, nq, z, deformed, d_jac, d_df,
                                 (hipStream_t)stream);
    return e == hipSuccess ? IMF_OK : hip_fail(e, "geom_pyr");
}

int imf_geom_tet_device(const double *d_verts, int nElmt, int nq0, int nq1, int nq2,
                        int deformed, double *d_jac, double *d_df, void *stream)
{
    if (!d_verts || !d_jac || !d_df || nElmt <= 0 || nq0 < 1 || nq1 < 1 || nq2 < 1 ||
        nq0 > 64 || nq1 > 64 || nq2 > 64)
        return fail(IMF_ERR_ARG, "imf_geom_tet_device: bad argument");
    hipError_t e = imf::geom_tet(d_verts, nElmt, nq0 * nq1 * nq2, deformed, d_jac, d_df,
                                 (hipStream_t)stream);
    return e == hipSuccess ? IMF_OK : hip_fail(e, "geom_tet");
}

int imf_geom_prism_device(const double *d_verts, int nElmt, int nq0, int nq1, int nq2,
                          int deformed, double *d_jac, double *d_df, void *stream)
{
    if (!d_verts || !d_jac || !d_df || nElmt <= 0 || nq0 < 2 || nq1 < 2 || nq2 < 1 ||
        nq0 > 16 || nq1 > 16 || nq2 > 16)
        return fail(IMF_ERR_ARG, "imf_geom_prism_device: bad argument");
    // nm = 2: at most 3 * 16 table values per direction
    std::vector<double> z0(nq0), z1(nq1), z2(nq2), ww(16), dd(256), bb(48), db(48);
    if (imf_basis_modified_a(2, nq0, z0.data(), ww.data(), dd.data(), bb.data(), db.data()) ||
        imf_basis_modified_a(2, nq1, z1.data(), ww.data(), dd.data(), bb.data(), db.data()) ||
        imf_basis_modified_b(2, nq2, z2.data(), ww.data(), dd.data(), bb.data(), db.data()))
        return fail(IMF_ERR_ARG, "imf_geom_prism_device: points");
    const double *z[3] = {z0.data(), z1.data(), z2.data()};
    const int nq[3] = {nq0, nq1, nq2};
    hipError_t e = imf::geom_prism(d_verts, nElmt, nq, z, deformed, d_jac, d_df,
                                   (hipStream_t)stream);
    return e == hipSuccess ? IMF_OK : hip_fail(e, "geom_prism");
}

}  // extern "C"
