// Drop-in check of the C++ host mirror: a caller written against the reference's
// Collections API (Collection::Initialise / ApplyOperator with a ConstFactorMap, operator
// keys, NekError on bad requests) runs the gfx950 operators and is compared with the CPU
// oracle element by element.  Prints one line per case; exit status 0 iff all pass.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "ithaca/Collections.hpp"

extern "C" {
#include "mf_oracle.h"
}

using namespace Nektar;

static double rel(const double *a, const double *b, size_t n)
{
    double num = 0, den = 0;
    for (size_t i = 0; i < n; ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return std::sqrt(num / den);
}

static std::vector<double> hex_mesh(int n, double jit, std::mt19937_64 &g)
{
    std::uniform_real_distribution<double> U(-jit, jit);
    const double h = 1.0 / n;
    std::vector<double> X((n + 1) * (n + 1) * (n + 1) * 3);
    auto at = [&](int i, int j, int k) { return &X[((k * (n + 1) + j) * (n + 1) + i) * 3]; };
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) {
                double *x = at(i, j, k);
                x[0] = i * h; x[1] = j * h; x[2] = k * h;
                if (i > 0 && j > 0 && k > 0 && i < n && j < n && k < n)
                    for (int c = 0; c < 3; ++c) x[c] += U(g) * h;
            }
    const int C[8][3] = {{0,0,0},{1,0,0},{1,1,0},{0,1,0},{0,0,1},{1,0,1},{1,1,1},{0,1,1}};
    std::vector<double> v(n * n * n * 24);
    for (int k = 0, e = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i, ++e)
                for (int c = 0; c < 8; ++c)
                    for (int d = 0; d < 3; ++d)
                        v[e * 24 + c * 3 + d] = at(i + C[c][0], j + C[c][1], k + C[c][2])[d];
    return v;
}

// The collapsed shapes through the Collections API: BwdTrans and Helmholtz of deformed
// triangles, prisms, tetrahedra and pyramids against the oracle.
static int collapsed_cases(std::mt19937_64 &g)
{
    using namespace LibUtilities;
    std::uniform_real_distribution<double> U(-1, 1), J(-0.1, 0.1);
    const int nm = 5, E = 37;
    const double lam = 1.5;
    int fails = 0;
    auto scaled = [](const BasisSharedPtr &b, double f) {
        std::vector<double> w(b->GetW());
        for (auto &x : w) x *= f;
        return w;
    };
    auto run = [&](const char *name, ShapeType shape, std::vector<BasisSharedPtr> bases,
                   const double *ref_verts, int nv, int M, auto geom_fn, auto bwd_fn, auto helm_fn,
                   auto ipdb_fn) {
        const int dim = (int)bases.size();
        int Q = 1;
        for (auto &b : bases) Q *= b->GetNumPoints();
        std::vector<double> verts(E * nv * dim);
        for (int e = 0; e < E; ++e)
            for (int i = 0; i < nv * dim; ++i) verts[e * nv * dim + i] = ref_verts[i] + J(g);
        auto jac = std::make_shared<std::vector<double>>(E * Q);
        auto df = std::make_shared<std::vector<double>>(dim * dim * E * Q);
        geom_fn(verts.data(), jac->data(), df->data());
        Collections::CollectionExp exp{shape, bases, E};
        if (exp.GetNcoeffs() != M) {
            printf("%s GetNcoeffs %d != %d (FAIL)\n", name, exp.GetNcoeffs(), M);
            return 1;
        }
        auto geom = std::make_shared<Collections::CoalescedGeomData>(true, jac, df);
        Collections::Collection coll(exp, geom,
                                     Collections::SetFixedImpType(Collections::eMatrixFreeHip));
        Array<OneD, NekDouble> u(E * M), phys(E * Q), out(E * M);
        for (int i = 0; i < E * M; ++i) u[i] = U(g);
        std::vector<double> rb(E * Q), rh(E * M);
        coll.Initialise(Collections::eBwdTrans);
        coll.ApplyOperator(Collections::eBwdTrans, u, phys);
        bwd_fn(u.data(), rb.data());
        double r = rel(phys.data(), rb.data(), E * Q);
        printf("BwdTrans_MatrixFreeHip %s rel=%.3e\n", name, r);
        int f = !(r < 1e-12);
        StdRegions::ConstFactorMap factors;
        factors[StdRegions::eFactorLambda] = lam;
        coll.Initialise(Collections::eHelmholtz);
        coll.ApplyOperator(Collections::eHelmholtz, u, out, factors);
        helm_fn(jac->data(), df->data(), u.data(), rh.data());
        r = rel(out.data(), rh.data(), E * M);
        printf("Helmholtz_MatrixFreeHip %s deformed rel=%.3e\n", name, r);
        f += !(r < 1e-12);
        // IProductWRTDerivBase through the multi-array ApplyOperator forms
        Array<OneD, NekDouble> d0(E * Q), d1(E * Q), d2(E * Q);
        for (int i = 0; i < E * Q; ++i) {
            d0[i] = U(g);
            d1[i] = U(g);
            d2[i] = U(g);
        }
        coll.Initialise(Collections::eIProductWRTDerivBase);
        if (dim == 2) coll.ApplyOperator(Collections::eIProductWRTDerivBase, d0, d1, out);
        else coll.ApplyOperator(Collections::eIProductWRTDerivBase, d0, d1, d2, out);
        ipdb_fn(jac->data(), df->data(), d0.data(), d1.data(), d2.data(), rh.data());
        r = rel(out.data(), rh.data(), E * M);
        printf("IProductWRTDerivBase_MatrixFreeHip %s deformed rel=%.3e\n", name, r);
        return f + !(r < 1e-12);
    };
    {   // triangles: Modified_A x Modified_B
        auto a = MakeModifiedA(nm, nm + 1), b = MakeModifiedB(nm, nm);
        const double v[6] = {0, 0, 1, 0, 0.3, 1};
        auto w1 = scaled(b, 0.5);
        fails += run("tri", eTriangle, {a, b}, v, 3, nm * (nm + 1) / 2,
            [&](const double *vv, double *j, double *d) { or_geom_tri(vv, E, nm + 1, nm, 1, j, d); },
            [&](const double *in, double *o) {
                or_bwdtrans_tri(nm, nm + 1, nm, a->GetBdata().data(), b->GetBdata().data(), E, in, o); },
            [&](const double *j, const double *d, const double *in, double *o) {
                or_helmholtz_tri(nm, nm + 1, nm, a->GetBdata().data(), a->GetDbdata().data(),
                                 b->GetBdata().data(), b->GetDbdata().data(), a->GetD().data(),
                                 b->GetD().data(), a->GetZ().data(), b->GetZ().data(),
                                 a->GetW().data(), w1.data(), E, 1, j, d, lam, in, o); },
            [&](const double *j, const double *d, const double *i0, const double *i1, const double *,
                double *o) {
                or_iprodderiv_tri(nm, nm + 1, nm, a->GetBdata().data(), a->GetDbdata().data(),
                                  b->GetBdata().data(), b->GetDbdata().data(), a->GetZ().data(),
                                  b->GetZ().data(), a->GetW().data(), w1.data(), E, 1, j, d, i0,
                                  i1, o); });
    }
    {   // prisms: Modified_A x Modified_A x Modified_B
        auto a = MakeModifiedA(nm, nm + 1), b = MakeModifiedB(nm, nm);
        const double v[18] = {0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0, 0.5, 0, 1, 0.5, 1, 1};
        auto w2 = scaled(b, 0.5);
        fails += run("prism", ePrism, {a, a, b}, v, 6, nm * nm * (nm + 1) / 2,
            [&](const double *vv, double *j, double *d) {
                or_geom_prism(vv, E, nm + 1, nm + 1, nm, a->GetZ().data(), a->GetZ().data(),
                              b->GetZ().data(), 1, j, d); },
            [&](const double *in, double *o) {
                or_bwdtrans_prism(nm, nm + 1, nm + 1, nm, a->GetBdata().data(), a->GetBdata().data(),
                                  b->GetBdata().data(), E, in, o); },
            [&](const double *j, const double *d, const double *in, double *o) {
                or_helmholtz_prism(nm, nm + 1, nm + 1, nm, a->GetBdata().data(), a->GetDbdata().data(),
                                   a->GetBdata().data(), a->GetDbdata().data(), b->GetBdata().data(),
                                   b->GetDbdata().data(), a->GetD().data(), a->GetD().data(),
                                   b->GetD().data(), a->GetZ().data(), b->GetZ().data(),
                                   a->GetW().data(), a->GetW().data(), w2.data(), E, 1, j, d, lam,
                                   in, o); },
            [&](const double *j, const double *d, const double *i0, const double *i1,
                const double *i2, double *o) {
                or_iprodderiv_prism(nm, nm + 1, nm + 1, nm, a->GetBdata().data(),
                                    a->GetDbdata().data(), a->GetBdata().data(),
                                    a->GetDbdata().data(), b->GetBdata().data(),
                                    b->GetDbdata().data(), a->GetZ().data(), b->GetZ().data(),
                                    a->GetW().data(), a->GetW().data(), w2.data(), E, 1, j, d, i0,
                                    i1, i2, o); });
    }
    {   // tetrahedra: Modified_A x Modified_B x Modified_C
        auto a = MakeModifiedA(nm, nm + 1), b = MakeModifiedB(nm, nm), c = MakeModifiedC(nm, nm);
        const double v[12] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1};
        auto w1 = scaled(b, 0.5), w2 = scaled(c, 0.25);
        fails += run("tet", eTetrahedron, {a, b, c}, v, 4, nm * (nm + 1) * (nm + 2) / 6,
            [&](const double *vv, double *j, double *d) { or_geom_tet(vv, E, nm + 1, nm, nm, 1, j, d); },
            [&](const double *in, double *o) {
                or_bwdtrans_tet(nm, nm + 1, nm, nm, a->GetBdata().data(), b->GetBdata().data(),
                                c->GetBdata().data(), E, in, o); },
            [&](const double *j, const double *d, const double *in, double *o) {
                or_helmholtz_tet(nm, nm + 1, nm, nm, a->GetBdata().data(), a->GetDbdata().data(),
                                 b->GetBdata().data(), b->GetDbdata().data(), c->GetBdata().data(),
                                 c->GetDbdata().data(), a->GetD().data(), b->GetD().data(),
                                 c->GetD().data(), a->GetZ().data(), b->GetZ().data(),
                                 c->GetZ().data(), a->GetW().data(), w1.data(), w2.data(), E, 1, j,
                                 d, lam, in, o); },
            [&](const double *j, const double *d, const double *i0, const double *i1,
                const double *i2, double *o) {
                or_iprodderiv_tet(nm, nm + 1, nm, nm, a->GetBdata().data(), a->GetDbdata().data(),
                                  b->GetBdata().data(), b->GetDbdata().data(), c->GetBdata().data(),
                                  c->GetDbdata().data(), a->GetZ().data(), b->GetZ().data(),
                                  c->GetZ().data(), a->GetW().data(), w1.data(), w2.data(), E, 1, j,
                                  d, i0, i1, i2, o); });
    }
    {   // pyramids: Modified_A x Modified_A x ModifiedPyr_C
        auto a = MakeModifiedA(nm, nm + 1), c = MakeModifiedPyrC(nm, nm);
        const double v[15] = {0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0, 0.5, 0.5, 0.866};
        auto w2 = scaled(c, 0.25);
        fails += run("pyr", ePyramid, {a, a, c}, v, 5, nm * (nm + 1) * (2 * nm + 1) / 6,
            [&](const double *vv, double *j, double *d) {
                or_geom_pyr(vv, E, nm + 1, nm + 1, nm, a->GetZ().data(), a->GetZ().data(),
                            c->GetZ().data(), 1, j, d); },
            [&](const double *in, double *o) {
                or_bwdtrans_pyr(nm, nm + 1, nm + 1, nm, a->GetBdata().data(), a->GetBdata().data(),
                                c->GetBdata().data(), E, in, o); },
            [&](const double *j, const double *d, const double *in, double *o) {
                or_helmholtz_pyr(nm, nm + 1, nm + 1, nm, a->GetBdata().data(), a->GetDbdata().data(),
                                 a->GetBdata().data(), a->GetDbdata().data(), c->GetBdata().data(),
                                 c->GetDbdata().data(), a->GetD().data(), a->GetD().data(),
                                 c->GetD().data(), a->GetZ().data(), a->GetZ().data(),
                                 c->GetZ().data(), a->GetW().data(), a->GetW().data(), w2.data(), E,
                                 1, j, d, lam, in, o); },
            [&](const double *j, const double *d, const double *i0, const double *i1,
                const double *i2, double *o) {
                or_iprodderiv_pyr(nm, nm + 1, nm + 1, nm, a->GetBdata().data(), a->GetDbdata().data(),
                                  a->GetBdata().data(), a->GetDbdata().data(), c->GetBdata().data(),
                                  c->GetDbdata().data(), a->GetZ().data(), a->GetZ().data(),
                                  c->GetZ().data(), a->GetW().data(), a->GetW().data(), w2.data(), E,
                                  1, j, d, i0, i1, i2, o); });
    }
    return fails;
}

// Segments of a one-dimensional mesh through the Collections API: the four operators the
// reference registers for segments against the oracle.
static int seg_cases(std::mt19937_64 &g)
{
    using namespace LibUtilities;
    std::uniform_real_distribution<double> U(-1, 1);
    const int nm = 5, nq = 7, E = 41;
    auto b = MakeModifiedA(nm, nq);
    std::vector<double> verts(2 * E);
    for (int e = 0; e < E; ++e) {
        verts[2 * e] = 0.1 * e;
        verts[2 * e + 1] = 0.1 * e + 0.05 + 0.04 * (U(g) + 1.0);
    }
    auto jac = std::make_shared<std::vector<double>>(E * nq);
    auto df = std::make_shared<std::vector<double>>(E * nq);
    or_geom_seg(verts.data(), E, 1, nq, 1, jac->data(), df->data());
    Collections::CollectionExp exp{eSegment, {b}, E};
    auto geom = std::make_shared<Collections::CoalescedGeomData>(true, jac, df);
    Collections::Collection coll(exp, geom,
                                 Collections::SetFixedImpType(Collections::eMatrixFreeHip));
    Array<OneD, NekDouble> u(E * nm), f(E * nq), phys(E * nq), out(E * nm), dd(E * nq);
    for (int i = 0; i < E * nm; ++i) u[i] = U(g);
    for (int i = 0; i < E * nq; ++i) f[i] = U(g);
    std::vector<double> rp(E * nq), rm(E * nm);
    int fails = 0;
    auto report = [&](const char *what, const double *a, const double *r, size_t n) {
        const double x = rel(a, r, n);
        printf("%s seg deformed rel=%.3e\n", what, x);
        fails += !(x < 1e-12);
    };
    coll.Initialise(Collections::eBwdTrans);
    coll.ApplyOperator(Collections::eBwdTrans, u, phys);
    or_bwdtrans_seg(nm, nq, b->GetBdata().data(), E, u.data(), rp.data());
    report("BwdTrans_MatrixFreeHip", phys.data(), rp.data(), E * nq);
    coll.Initialise(Collections::eIProductWRTBase);
    coll.ApplyOperator(Collections::eIProductWRTBase, f, out);
    or_iproduct_seg(nm, nq, b->GetBdata().data(), b->GetW().data(), E, 1, jac->data(), f.data(),
                    rm.data());
    report("IProductWRTBase_MatrixFreeHip", out.data(), rm.data(), E * nm);
    coll.Initialise(Collections::ePhysDeriv);
    coll.ApplyOperator(Collections::ePhysDeriv, 0, f, dd);
    or_physderiv_seg(nq, b->GetD().data(), E, 1, 1, df->data(), f.data(), rp.data());
    report("PhysDeriv_MatrixFreeHip", dd.data(), rp.data(), E * nq);
    coll.Initialise(Collections::eIProductWRTDerivBase);
    coll.ApplyOperator(Collections::eIProductWRTDerivBase, f, out);
    or_iprodderiv_seg(nm, nq, b->GetDbdata().data(), b->GetW().data(), E, 1, 1, jac->data(),
                      df->data(), f.data(), rm.data());
    report("IProductWRTDerivBase_MatrixFreeHip", out.data(), rm.data(), E * nm);
    return fails;
}

int main()
{
    int fails = 0;
    std::mt19937_64 g(12345);
    const int nm = 6, nq = 7, n = 4, E = n * n * n;
    auto B = LibUtilities::MakeModifiedA(nm, nq);
    std::vector<double> verts = hex_mesh(n, 0.1, g);
    const int Q = nq * nq * nq, M = nm * nm * nm;
    auto jac = std::make_shared<std::vector<double>>(E * Q);
    auto df = std::make_shared<std::vector<double>>(9 * E * Q);
    or_geom_hex(verts.data(), E, nq, B->GetZ().data(), 1, jac->data(), df->data());

    Collections::CollectionExp exp{LibUtilities::eHexahedron, {B, B, B}, E};
    auto geom = std::make_shared<Collections::CoalescedGeomData>(true, jac, df);
    Collections::Collection coll(exp, geom,
                                 Collections::SetFixedImpType(Collections::eMatrixFreeHip));

    std::uniform_real_distribution<double> U(-1, 1);
    Array<OneD, NekDouble> u(E * M), f(E * Q), out(E * M), phys(E * Q);
    for (int i = 0; i < E * M; ++i) u[i] = U(g);
    for (int i = 0; i < E * Q; ++i) f[i] = U(g);
    std::vector<double> ref(E * Q);

    coll.Initialise(Collections::eBwdTrans);
    coll.ApplyOperator(Collections::eBwdTrans, u, phys);
    or_bwdtrans_hex(nm, nq, B->GetBdata().data(), E, u.data(), ref.data());
    double r = rel(phys.data(), ref.data(), E * Q);
    printf("BwdTrans_MatrixFreeHip hex rel=%.3e\n", r);
    fails += !(r < 1e-12);

    coll.Initialise(Collections::eIProductWRTBase);
    coll.ApplyOperator(Collections::eIProductWRTBase, f, out);
    or_iproduct_hex(nm, nq, B->GetBdata().data(), B->GetW().data(), E, 1, jac->data(), f.data(),
                    ref.data());
    r = rel(out.data(), ref.data(), E * M);
    printf("IProductWRTBase_MatrixFreeHip hex rel=%.3e\n", r);
    fails += !(r < 1e-12);

    StdRegions::ConstFactorMap factors;
    factors[StdRegions::eFactorLambda] = 10.0;
    coll.Initialise(Collections::eHelmholtz);
    coll.ApplyOperator(Collections::eHelmholtz, u, out, factors);
    or_helmholtz_hex(nm, nq, B->GetBdata().data(), B->GetDbdata().data(), B->GetD().data(),
                     B->GetW().data(), E, 1, jac->data(), df->data(), 10.0, u.data(), ref.data());
    r = rel(out.data(), ref.data(), E * M);
    printf("Helmholtz_MatrixFreeHip hex rel=%.3e ndof=%.0f\n", r,
           (double)coll.GetOpSharedPtr(Collections::eHelmholtz)->GetNumElmt() * M);
    fails += !(r < 1e-12);

    // lambda is re-read on every call
    factors[StdRegions::eFactorLambda] = 0.5;
    coll.ApplyOperator(Collections::eHelmholtz, u, out, factors);
    or_helmholtz_hex(nm, nq, B->GetBdata().data(), B->GetDbdata().data(), B->GetD().data(),
                     B->GetW().data(), E, 1, jac->data(), df->data(), 0.5, u.data(), ref.data());
    r = rel(out.data(), ref.data(), E * M);
    printf("Helmholtz_MatrixFreeHip hex lambda=0.5 rel=%.3e\n", r);
    fails += !(r < 1e-12);

    // PhysDeriv and IProductWRTDerivBase through the multi-array ApplyOperator forms
    {
        Array<OneD, NekDouble> d0(E * Q), d1(E * Q), d2(E * Q), dd(E * Q), oi(E * M);
        std::vector<double> r0(E * Q), r1(E * Q), r2(E * Q);
        coll.Initialise(Collections::ePhysDeriv);
        coll.ApplyOperator(Collections::ePhysDeriv, f, d0, d1, d2);
        or_physderiv_hex(nq, B->GetD().data(), E, 1, df->data(), f.data(), r0.data(), r1.data(),
                         r2.data());
        r = std::max({rel(d0.data(), r0.data(), E * Q), rel(d1.data(), r1.data(), E * Q),
                      rel(d2.data(), r2.data(), E * Q)});
        printf("PhysDeriv_MatrixFreeHip hex rel=%.3e\n", r);
        fails += !(r < 1e-12);
        coll.ApplyOperator(Collections::ePhysDeriv, 1, f, dd);
        r = rel(dd.data(), r1.data(), E * Q);
        printf("PhysDeriv_MatrixFreeHip hex dir=1 rel=%.3e\n", r);
        fails += !(r < 1e-12);
        coll.Initialise(Collections::eIProductWRTDerivBase);
        coll.ApplyOperator(Collections::eIProductWRTDerivBase, d0, d1, d2, oi);
        or_iprod_deriv_hex(nm, nq, B->GetBdata().data(), B->GetDbdata().data(), B->GetW().data(), E,
                           1, jac->data(), df->data(), d0.data(), d1.data(), d2.data(), ref.data());
        r = rel(oi.data(), ref.data(), E * M);
        printf("IProductWRTDerivBase_MatrixFreeHip hex rel=%.3e\n", r);
        fails += !(r < 1e-12);
    }

    // Regular quads through the same API
    {
        const int nmq = 5, nqq = 6, nn = 6, Eq = nn * nn, Qq = nqq * nqq, Mq = nmq * nmq;
        auto Bq = LibUtilities::MakeModifiedA(nmq, nqq);
        std::vector<double> vq(Eq * 8);
        for (int j = 0, e = 0; j < nn; ++j)
            for (int i = 0; i < nn; ++i, ++e) {
                const double x0 = double(i) / nn, y0 = double(j) / nn, h = 1.0 / nn;
                const double c[8] = {x0, y0, x0 + h, y0, x0 + h, y0 + h, x0, y0 + h};
                for (int k = 0; k < 8; ++k) vq[e * 8 + k] = c[k];
            }
        auto jq = std::make_shared<std::vector<double>>(Eq);
        auto dq = std::make_shared<std::vector<double>>(4 * Eq);
        or_geom_quad(vq.data(), Eq, nqq, Bq->GetZ().data(), 0, jq->data(), dq->data());
        Collections::CollectionExp eq{LibUtilities::eQuadrilateral, {Bq, Bq}, Eq};
        Collections::Collection cq(eq, std::make_shared<Collections::CoalescedGeomData>(false, jq, dq),
                                   Collections::SetFixedImpType(Collections::eMatrixFreeHip));
        Array<OneD, NekDouble> uq(Eq * Mq), oq(Eq * Mq);
        for (int i = 0; i < Eq * Mq; ++i) uq[i] = U(g);
        std::vector<double> rq(Eq * Mq);
        StdRegions::ConstFactorMap fq;
        fq[StdRegions::eFactorLambda] = 1.0;
        cq.Initialise(Collections::eHelmholtz);
        cq.ApplyOperator(Collections::eHelmholtz, uq, oq, fq);
        or_helmholtz_quad(nmq, nqq, Bq->GetBdata().data(), Bq->GetDbdata().data(),
                          Bq->GetD().data(), Bq->GetW().data(), Eq, 0, jq->data(), dq->data(), 1.0,
                          uq.data(), rq.data());
        r = rel(oq.data(), rq.data(), Eq * Mq);
        printf("Helmholtz_MatrixFreeHip quad regular rel=%.3e\n", r);
        fails += !(r < 1e-12);
        (void)Qq;
    }

    // Error behaviour: unknown operator key, dir overload, missing lambda.
    auto expect_throw = [&](const char *what, auto fn) {
        try {
            fn();
            printf("%s: no exception (FAIL)\n", what);
            ++fails;
        } catch (const NekError &e) {
            printf("%s: NekError \"%s\"\n", what, e.what());
        }
    };
    expect_throw("unknown key", [&] {
        Collections::CollectionExp et{LibUtilities::eSegment, {B}, 4};
        Collections::Collection ct(et, geom, Collections::SetFixedImpType(Collections::eMatrixFreeHip));
        ct.Initialise(Collections::eHelmholtz);
    });
    expect_throw("dir overload", [&] {
        Array<OneD, NekDouble> o(E * M);
        (*coll.GetOpSharedPtr(Collections::eHelmholtz))(0, u, o);
    });
    expect_throw("missing lambda", [&] {
        coll.ApplyOperator(Collections::eHelmholtz, u, out, StdRegions::NullConstFactorMap);
    });
    expect_throw("unsupported combo", [&] {
        MatrixFree::GetOperatorFactory().CreateInstance(
            "Helmholtz_Hex_Deformed", {LibUtilities::MakeModifiedA(6, 14), LibUtilities::MakeModifiedA(6, 14),
                                       LibUtilities::MakeModifiedA(6, 14)}, 2);
    });
    expect_throw("bases that do not match the shape", [&] {
        Collections::CollectionExp et{LibUtilities::eTetrahedron, {B, B, B}, 4};
        Collections::Collection ct(et, geom, Collections::SetFixedImpType(Collections::eMatrixFreeHip));
        ct.Initialise(Collections::eHelmholtz);
    });
    fails += collapsed_cases(g);
    fails += seg_cases(g);
    printf("%s\n", fails ? "FAIL" : "PASS");
    return fails ? 1 : 0;
}
