// Implementation of the host mirror (Basis, MatrixFree operators, Collections adapters)
// over the C ABI of include/ithaca_mf.h.
#include <algorithm>
#include <iostream>
#include <sstream>

#include "../../../include/ithaca_mf.h"
#include "ithaca/Collections.hpp"

namespace Nektar {

namespace {

void check(int status)
{
    if (status != IMF_OK) throw NekError(imf_last_error());
}

}  // namespace

// ---------------------------------------------------------------------------
// LibUtilities::Basis
// ---------------------------------------------------------------------------
namespace LibUtilities {

Basis::Basis(const BasisKey &key) : m_key(key)
{
    const BasisType b = key.GetBasisType();
    const PointsType pt = key.GetPointsKey().GetPointsType();
    ASSERTL0((b == eModified_A && pt == eGaussLobattoLegendre) ||
                 (b == eModified_B && pt == eGaussRadauMAlpha1Beta0) ||
                 ((b == eModified_C || b == eModifiedPyr_C) && pt == eGaussRadauMAlpha2Beta0),
             "Basis: provided are Modified_A on Gauss-Lobatto-Legendre points, Modified_B on "
             "Gauss-Radau alpha = 1 points and Modified_C / ModifiedPyr_C on Gauss-Radau "
             "alpha = 2 points");
    const int nm = key.GetNumModes(), nq = key.GetPointsKey().GetNumPoints();
    const int rows = b == eModified_A ? nm
                   : b == eModified_B ? nm * (nm + 1) / 2
                   : b == eModified_C ? nm * (nm + 1) * (nm + 2) / 6
                                      : nm * (nm + 1) * (2 * nm + 1) / 6;
    m_z.resize(nq);
    m_w.resize(nq);
    m_D.resize(nq * nq);
    m_bdata.resize((size_t)rows * nq);
    m_dbdata.resize((size_t)rows * nq);
    auto fn = b == eModified_A ? imf_basis_modified_a
            : b == eModified_B ? imf_basis_modified_b
            : b == eModified_C ? imf_basis_modified_c
                               : imf_basis_modified_pyr_c;
    check(fn(nm, nq, m_z.data(), m_w.data(), m_D.data(), m_bdata.data(), m_dbdata.data()));
}

}  // namespace LibUtilities

// ---------------------------------------------------------------------------
// MatrixFree
// ---------------------------------------------------------------------------
namespace MatrixFree {

namespace {

// The basis types each shape's kernels read (StdTriExp, StdPrismExp, StdTetExp, StdPyrExp
// and the tensor shapes): imf_create takes the tables by pointer and cannot check their
// lengths, so a mismatch is refused here.
void check_bases(const std::string &key, const std::vector<LibUtilities::BasisSharedPtr> &basis)
{
    using namespace LibUtilities;
    const std::string shape = key.substr(key.find('_') + 1, key.rfind('_') - key.find('_') - 1);
    std::vector<BasisType> want;
    if (shape == "Seg") want = {eModified_A};
    else if (shape == "Quad") want = {eModified_A, eModified_A};
    else if (shape == "Hex") want = {eModified_A, eModified_A, eModified_A};
    else if (shape == "Tri") want = {eModified_A, eModified_B};
    else if (shape == "Prism") want = {eModified_A, eModified_A, eModified_B};
    else if (shape == "Tet") want = {eModified_A, eModified_B, eModified_C};
    else if (shape == "Pyr") want = {eModified_A, eModified_A, eModifiedPyr_C};
    bool ok = want.size() == basis.size();
    for (size_t d = 0; ok && d < want.size(); ++d) ok = basis[d]->GetBasisType() == want[d];
    ASSERTL0(ok, key + ": the bases do not match the shape's expansion");
}

}  // namespace

Operator::Operator(const std::string &key, std::vector<LibUtilities::BasisSharedPtr> basis,
                   int nElmt)
    : m_basis(basis), m_nElmt(nElmt)
{
    check_bases(key, basis);
    const int dim = (int)basis.size();
    std::vector<int> nm(dim), nq(dim);
    std::vector<const double *> b(dim), db(dim), D(dim), w(dim);
    for (int d = 0; d < dim; ++d) {
        nm[d] = basis[d]->GetNumModes();
        nq[d] = basis[d]->GetNumPoints();
        b[d] = basis[d]->GetBdata().data();
        db[d] = basis[d]->GetDbdata().data();
        D[d] = basis[d]->GetD().data();
        w[d] = basis[d]->GetW().data();
    }
    check(imf_create(key.c_str(), dim, nm.data(), nq.data(), b.data(), db.data(), D.data(),
                     w.data(), nElmt, IMF_FP64, &m_op));
}

Operator::~Operator() { imf_destroy(m_op); }

NekDouble Operator::Ndof() { return imf_ndof(m_op); }

void Operator::SetJac(const std::shared_ptr<std::vector<NekDouble>> &jac)
{
    check(imf_set_jac(m_op, jac->data(), jac->size()));
}

void Operator::SetDF(const std::shared_ptr<std::vector<NekDouble>> &df)
{
    check(imf_set_df(m_op, df->data(), df->size()));
}

void Operator::SetJacDevice(const NekDouble *d_jac, size_t n) { check(imf_set_jac_device(m_op, d_jac, n)); }
void Operator::SetDFDevice(const NekDouble *d_df, size_t n) { check(imf_set_df_device(m_op, d_df, n)); }

void Operator::ApplyDevice(const NekDouble *d_in, NekDouble *d_out, void *stream)
{
    check(imf_apply_device(m_op, d_in, d_out, stream));
}

void Operator::ApplyHost(const NekDouble *in, NekDouble *out) { check(imf_apply(m_op, in, out)); }

void BwdTrans::operator()(const Array<OneD, const NekDouble> &in, Array<OneD, NekDouble> &out)
{
    ASSERTL0(in.size() >= imf_in_size(m_op) && out.size() >= imf_out_size(m_op),
             "BwdTrans: array too small");
    ApplyHost(in.data(), out.data());
}

void IProduct::operator()(const Array<OneD, const NekDouble> &in, Array<OneD, NekDouble> &out)
{
    ASSERTL0(in.size() >= imf_in_size(m_op) && out.size() >= imf_out_size(m_op),
             "IProduct: array too small");
    ApplyHost(in.data(), out.data());
}

void Helmholtz::operator()(const Array<OneD, const NekDouble> &in, Array<OneD, NekDouble> &out)
{
    ASSERTL0(in.size() >= imf_in_size(m_op) && out.size() >= imf_out_size(m_op),
             "Helmholtz: array too small");
    ApplyHost(in.data(), out.data());
}

void Helmholtz::SetLambda(NekDouble lambda)
{
    m_lambda = lambda;
    check(imf_set_lambda(m_op, lambda));
}

void PhysDeriv::operator()(const Array<OneD, const NekDouble> &in,
                           Array<OneD, Array<OneD, NekDouble>> &out)
{
    const size_t nq = imf_in_size(m_op), dim = m_basis.size();
    ASSERTL0(out.size() == dim, "PhysDeriv: need one output per dimension");
    ASSERTL0(in.size() >= nq, "PhysDeriv: array too small");
    for (size_t d = 0; d < dim; ++d) ASSERTL0(out[d].size() >= nq, "PhysDeriv: array too small");
    std::vector<NekDouble> tmp(dim * nq);
    ApplyHost(in.data(), tmp.data());
    for (size_t d = 0; d < dim; ++d)
        std::copy(tmp.begin() + d * nq, tmp.begin() + (d + 1) * nq, out[d].data());
}

void IProductWRTDerivBase::operator()(const Array<OneD, Array<OneD, NekDouble>> &in,
                                      Array<OneD, NekDouble> &out)
{
    const size_t dim = m_basis.size(), nq = imf_in_size(m_op) / dim;
    ASSERTL0(in.size() == dim, "IProductWRTDerivBase: need one input per dimension");
    ASSERTL0(out.size() >= imf_out_size(m_op), "IProductWRTDerivBase: array too small");
    std::vector<NekDouble> tmp(dim * nq);
    for (size_t d = 0; d < dim; ++d) {
        ASSERTL0(in[d].size() >= nq, "IProductWRTDerivBase: array too small");
        std::copy(in[d].data(), in[d].data() + nq, tmp.begin() + d * nq);
    }
    ApplyHost(tmp.data(), out.data());
}

OperatorFactory &GetOperatorFactory()
{
    static OperatorFactory tmp;
    return tmp;
}

std::string GetOpstring(LibUtilities::ShapeType shape, bool deformed)
{
    std::string s = "_";
    switch (shape) {
    case LibUtilities::eSegment: s += "Seg"; break;
    case LibUtilities::eTriangle: s += "Tri"; break;
    case LibUtilities::eQuadrilateral: s += "Quad"; break;
    case LibUtilities::eTetrahedron: s += "Tet"; break;
    case LibUtilities::ePyramid: s += "Pyr"; break;
    case LibUtilities::ePrism: s += "Prism"; break;
    case LibUtilities::eHexahedron: s += "Hex"; break;
    default: break;
    }
    s += deformed ? "_Deformed" : "_Regular";
    return s;
}

namespace {

template <class T>
OperatorSharedPtr make(const std::string &key, std::vector<LibUtilities::BasisSharedPtr> b, int n)
{
    return std::make_shared<T>(key, b, n);
}

// Keys registered by the reference: MatrixFreeOps/BwdTrans.cpp:7-35, IProduct{Seg,Quad,Hex,
// Tri,Prism,Tet,Pyr}.cpp, Helmholtz.cpp:7-42 (no segments), PhysDeriv.cpp:14-49,
// IProductWRTDerivBase{Seg,Quad,Hex,Tri,Prism,Tet,Pyr}.cpp.  Segments made here lie in a
// one-dimensional mesh (one derivative factor per point: the basis count is the dimension);
// segments of 2D and 3D meshes are created through imf_create with dim = the coordinate
// dimension.
const bool registered = [] {
    auto &f = GetOperatorFactory();
    for (const char *shape : {"Seg", "Quad", "Hex", "Tri", "Prism", "Tet", "Pyr"}) {
        const std::string s(shape);
        const std::string kb = "BwdTrans_" + s + "_Regular";
        f.RegisterCreatorFunction(kb, [kb](std::vector<LibUtilities::BasisSharedPtr> b, int n) {
            return make<BwdTrans>(kb, b, n);
        });
        for (const char *d : {"_Regular", "_Deformed"}) {
            const std::string kp = "PhysDeriv_" + s + d, kd = "IProductWRTDerivBase_" + s + d;
            f.RegisterCreatorFunction(kp, [kp](std::vector<LibUtilities::BasisSharedPtr> b, int n) {
                return make<PhysDeriv>(kp, b, n);
            });
            f.RegisterCreatorFunction(kd, [kd](std::vector<LibUtilities::BasisSharedPtr> b, int n) {
                return make<IProductWRTDerivBase>(kd, b, n);
            });
            const std::string ki = "IProduct_" + s + d, kh = "Helmholtz_" + s + d;
            f.RegisterCreatorFunction(ki, [ki](std::vector<LibUtilities::BasisSharedPtr> b, int n) {
                return make<IProduct>(ki, b, n);
            });
            if (s == "Seg") continue;
            f.RegisterCreatorFunction(kh, [kh](std::vector<LibUtilities::BasisSharedPtr> b, int n) {
                return make<Helmholtz>(kh, b, n);
            });
        }
    }
    return true;
}();

}  // namespace
}  // namespace MatrixFree

// ---------------------------------------------------------------------------
// Collections
// ---------------------------------------------------------------------------
namespace Collections {

// StdExpansion::GetNcoeffs of the shape (StdTriData / StdTetData / StdPyrData /
// StdPrismData::getNumberOfCoefficients for the collapsed shapes).
int CollectionExp::GetNcoeffs() const
{
    const int m = basis[0]->GetNumModes();
    switch (shape) {
    case LibUtilities::eTriangle: return m * (m + 1) / 2;
    case LibUtilities::ePrism: return m * m * (m + 1) / 2;
    case LibUtilities::eTetrahedron: return m * (m + 1) * (m + 2) / 6;
    case LibUtilities::ePyramid: return m * (m + 1) * (2 * m + 1) / 6;
    default: {
        int n = 1;
        for (auto &b : basis) n *= b->GetNumModes();
        return n;
    }
    }
}

int CollectionExp::GetTotPoints() const
{
    int n = 1;
    for (auto &b : basis) n *= b->GetNumPoints();
    return n;
}

OperatorImpMap SetFixedImpType(ImplementationType defaultType)
{
    OperatorImpMap m;
    for (int i = 0; i < SIZE_OperatorType; ++i) m[(OperatorType)i] = defaultType;
    return m;
}

std::ostream &operator<<(std::ostream &os, const OperatorKey &p)
{
    os << LibUtilities::ShapeTypeMap[std::get<0>(p)] << ", "
       << OperatorTypeMap[std::get<1>(p)] << ", " << ImplementationTypeMap[std::get<2>(p)];
    if (std::get<3>(p)) os << ", Nodal";
    return os;
}

OperatorFactory &GetOperatorFactory()
{
    static OperatorFactory tmp;
    return tmp;
}

Operator::Operator(const CollectionExp &exp, CoalescedGeomDataSharedPtr geom)
    : m_isDeformed(geom ? geom->IsDeformed() : false), m_numElmt(exp.nElmt)
{
}

namespace {

std::shared_ptr<MatrixFree::Operator> create_mf(const std::string &op, const CollectionExp &exp,
                                                bool deformed)
{
    const std::string key = op + MatrixFree::GetOpstring(exp.shape, deformed);
    return MatrixFree::GetOperatorFactory().CreateInstance(key, exp.basis, exp.nElmt);
}

// Collections/BwdTrans.cpp:140-240 (BwdTrans_MatrixFree), on the GPU.
class BwdTrans_MatrixFreeHip final : public Operator {
public:
    BwdTrans_MatrixFreeHip(const CollectionExp &exp, CoalescedGeomDataSharedPtr geom)
        : Operator(exp, geom)
    {
        m_oper = std::dynamic_pointer_cast<MatrixFree::BwdTrans>(create_mf("BwdTrans", exp, false));
        ASSERTL0(m_oper, "Failed to cast pointer.");
    }
    void operator()(const Array<OneD, const NekDouble> &input, Array<OneD, NekDouble> &output0,
                    Array<OneD, NekDouble> &, Array<OneD, NekDouble> &, Array<OneD, NekDouble> &,
                    const StdRegions::ConstFactorMap &) final
    {
        (*m_oper)(input, output0);
    }
    void operator()(int, const Array<OneD, const NekDouble> &, Array<OneD, NekDouble> &,
                    Array<OneD, NekDouble> &) final
    {
        NEKERROR(ErrorUtil::efatal, "BwdTrans_MatrixFree: Not valid for this operator.");
    }
    std::shared_ptr<MatrixFree::BwdTrans> m_oper;
};

// Collections/IProductWRTBase.cpp:168-277 (IProductWRTBase_MatrixFree), on the GPU.
class IProductWRTBase_MatrixFreeHip final : public Operator {
public:
    IProductWRTBase_MatrixFreeHip(const CollectionExp &exp, CoalescedGeomDataSharedPtr geom)
        : Operator(exp, geom)
    {
        auto oper = create_mf("IProduct", exp, m_isDeformed);
        oper->SetJac(geom->GetJac());
        m_oper = std::dynamic_pointer_cast<MatrixFree::IProduct>(oper);
        ASSERTL0(m_oper, "Failed to cast pointer.");
    }
    void operator()(const Array<OneD, const NekDouble> &input, Array<OneD, NekDouble> &output,
                    Array<OneD, NekDouble> &, Array<OneD, NekDouble> &, Array<OneD, NekDouble> &,
                    const StdRegions::ConstFactorMap &) final
    {
        (*m_oper)(input, output);
    }
    void operator()(int, const Array<OneD, const NekDouble> &, Array<OneD, NekDouble> &,
                    Array<OneD, NekDouble> &) final
    {
        NEKERROR(ErrorUtil::efatal, "Not valid for this operator.");
    }
    std::shared_ptr<MatrixFree::IProduct> m_oper;
};

// Collections/Helmholtz.cpp:391-503 (Helmholtz_MatrixFree), on the GPU.  lambda is read
// from factors[eFactorLambda] on every call (Helmholtz.cpp:408-414).
class Helmholtz_MatrixFreeHip final : public Operator {
public:
    Helmholtz_MatrixFreeHip(const CollectionExp &exp, CoalescedGeomDataSharedPtr geom)
        : Operator(exp, geom)
    {
        auto oper = create_mf("Helmholtz", exp, m_isDeformed);
        oper->SetJac(geom->GetJac());
        oper->SetDF(geom->GetDerivFactors());
        m_oper = std::dynamic_pointer_cast<MatrixFree::Helmholtz>(oper);
        ASSERTL0(m_oper, "Failed to cast pointer.");
    }
    void operator()(const Array<OneD, const NekDouble> &input, Array<OneD, NekDouble> &output0,
                    Array<OneD, NekDouble> &, Array<OneD, NekDouble> &, Array<OneD, NekDouble> &,
                    const StdRegions::ConstFactorMap &factors) final
    {
        auto x = factors.find(StdRegions::eFactorLambda);
        ASSERTL0(x != factors.end(), "Constant factor not defined: FactorLambda");
        m_oper->SetLambda(x->second);
        (*m_oper)(input, output0);
    }
    void operator()(int, const Array<OneD, const NekDouble> &, Array<OneD, NekDouble> &,
                    Array<OneD, NekDouble> &) final
    {
        NEKERROR(ErrorUtil::efatal, "Not valid for this operator.");
    }
    std::shared_ptr<MatrixFree::Helmholtz> m_oper;
};

// Collections/PhysDeriv.cpp:270-400 (PhysDeriv_MatrixFree), on the GPU: outputs
// output0..output{dim-1}; the dir overload returns one direction.
class PhysDeriv_MatrixFreeHip final : public Operator {
public:
    PhysDeriv_MatrixFreeHip(const CollectionExp &exp, CoalescedGeomDataSharedPtr geom)
        : Operator(exp, geom), m_dim((int)exp.basis.size()), m_nq(exp.GetTotPoints() * exp.nElmt)
    {
        auto oper = create_mf("PhysDeriv", exp, m_isDeformed);
        oper->SetDF(geom->GetDerivFactors());
        m_oper = std::dynamic_pointer_cast<MatrixFree::PhysDeriv>(oper);
        ASSERTL0(m_oper, "Failed to cast pointer.");
        m_output = Array<OneD, Array<OneD, NekDouble>>(m_dim);
        for (int d = 0; d < m_dim; ++d) m_output[d] = Array<OneD, NekDouble>(m_nq);
    }
    void operator()(const Array<OneD, const NekDouble> &input, Array<OneD, NekDouble> &output0,
                    Array<OneD, NekDouble> &output1, Array<OneD, NekDouble> &output2,
                    Array<OneD, NekDouble> &, const StdRegions::ConstFactorMap &) final
    {
        (*m_oper)(input, m_output);
        Array<OneD, NekDouble> *outs[3] = {&output0, &output1, &output2};
        for (int d = 0; d < m_dim; ++d)
            std::copy(m_output[d].data(), m_output[d].data() + m_nq, outs[d]->data());
    }
    void operator()(int dir, const Array<OneD, const NekDouble> &input,
                    Array<OneD, NekDouble> &output, Array<OneD, NekDouble> &) final
    {
        ASSERTL0(dir >= 0 && dir < m_dim, "PhysDeriv_MatrixFree: direction out of range");
        (*m_oper)(input, m_output);
        std::copy(m_output[dir].data(), m_output[dir].data() + m_nq, output.data());
    }
    int m_dim;
    size_t m_nq;
    Array<OneD, Array<OneD, NekDouble>> m_output;
    std::shared_ptr<MatrixFree::PhysDeriv> m_oper;
};

// Collections/IProductWRTDerivBase.cpp:274-400 (IProductWRTDerivBase_MatrixFree), on the
// GPU: inputs entry0..entry{dim-1}, output entry{dim}.
class IProductWRTDerivBase_MatrixFreeHip final : public Operator {
public:
    IProductWRTDerivBase_MatrixFreeHip(const CollectionExp &exp, CoalescedGeomDataSharedPtr geom)
        : Operator(exp, geom), m_dim((int)exp.basis.size())
    {
        auto oper = create_mf("IProductWRTDerivBase", exp, m_isDeformed);
        oper->SetJac(geom->GetJac());
        oper->SetDF(geom->GetDerivFactors());
        m_oper = std::dynamic_pointer_cast<MatrixFree::IProductWRTDerivBase>(oper);
        ASSERTL0(m_oper, "Failed to cast pointer.");
    }
    void operator()(const Array<OneD, const NekDouble> &entry0, Array<OneD, NekDouble> &entry1,
                    Array<OneD, NekDouble> &entry2, Array<OneD, NekDouble> &entry3,
                    Array<OneD, NekDouble> &, const StdRegions::ConstFactorMap &) final
    {
        Array<OneD, Array<OneD, NekDouble>> in(m_dim);
        const size_t n = entry0.size();
        in[0] = Array<OneD, NekDouble>(n, entry0.data());
        Array<OneD, NekDouble> *ents[3] = {&entry1, &entry2, &entry3};
        for (int d = 1; d < m_dim; ++d) in[d] = *ents[d - 1];
        (*m_oper)(in, *ents[m_dim - 1]);
    }
    void operator()(int, const Array<OneD, const NekDouble> &, Array<OneD, NekDouble> &,
                    Array<OneD, NekDouble> &) final
    {
        NEKERROR(ErrorUtil::efatal, "Not valid for this operator.");
    }
    int m_dim;
    std::shared_ptr<MatrixFree::IProductWRTDerivBase> m_oper;
};

template <class T>
OperatorSharedPtr create(const CollectionExp &e, CoalescedGeomDataSharedPtr g)
{
    return std::make_shared<T>(e, g);
}

const bool registered = [] {
    auto &f = GetOperatorFactory();
    for (auto shape : {LibUtilities::eSegment, LibUtilities::eQuadrilateral,
                       LibUtilities::eHexahedron, LibUtilities::eTriangle, LibUtilities::ePrism,
                       LibUtilities::eTetrahedron, LibUtilities::ePyramid}) {
        f.RegisterCreatorFunction(OperatorKey(shape, eBwdTrans, eMatrixFreeHip, false),
                                  create<BwdTrans_MatrixFreeHip>, "BwdTrans_MatrixFreeHip");
        f.RegisterCreatorFunction(OperatorKey(shape, eIProductWRTBase, eMatrixFreeHip, false),
                                  create<IProductWRTBase_MatrixFreeHip>,
                                  "IProductWRTBase_MatrixFreeHip");
        if (shape != LibUtilities::eSegment)  // no segment Helmholtz in the reference
            f.RegisterCreatorFunction(OperatorKey(shape, eHelmholtz, eMatrixFreeHip, false),
                                      create<Helmholtz_MatrixFreeHip>, "Helmholtz_MatrixFreeHip");
        f.RegisterCreatorFunction(OperatorKey(shape, ePhysDeriv, eMatrixFreeHip, false),
                                  create<PhysDeriv_MatrixFreeHip>, "PhysDeriv_MatrixFreeHip");
        f.RegisterCreatorFunction(OperatorKey(shape, eIProductWRTDerivBase, eMatrixFreeHip, false),
                                  create<IProductWRTDerivBase_MatrixFreeHip>,
                                  "IProductWRTDerivBase_MatrixFreeHip");
    }
    return true;
}();

}  // namespace

Collection::Collection(const CollectionExp &exp, CoalescedGeomDataSharedPtr geom,
                       const OperatorImpMap &impTypes)
    : m_exp(exp), m_geomData(geom), m_impTypes(impTypes)
{
}

void Collection::Initialise(const OperatorType opType)
{
    if (HasOperator(opType)) return;
    auto it = m_impTypes.find(opType);
    if (it == m_impTypes.end()) {
        NEKERROR(ErrorUtil::ewarning, "Failed to determine implmentation to initialise "
                                      "collection operator");
        return;
    }
    OperatorKey key(m_exp.shape, opType, it->second, m_exp.nodal);
    std::stringstream ss;
    ss << key;
    ASSERTL0(GetOperatorFactory().ModuleExists(key), "Requested unknown operator " + ss.str());
    m_ops[opType] = GetOperatorFactory().CreateInstance(key, m_exp, m_geomData);
}

void Collection::ApplyOperator(const OperatorType &op, const Array<OneD, const NekDouble> &inarray,
                               Array<OneD, NekDouble> &output,
                               const StdRegions::ConstFactorMap &factors)
{
    Array<OneD, NekDouble> wsp(m_ops[op]->GetWspSize());
    (*m_ops[op])(inarray, output, NullNekDouble1DArray, NullNekDouble1DArray, wsp, factors);
}

void Collection::ApplyOperator(const OperatorType &op, const Array<OneD, const NekDouble> &inarray,
                               Array<OneD, NekDouble> &output0, Array<OneD, NekDouble> &output1)
{
    Array<OneD, NekDouble> wsp(m_ops[op]->GetWspSize());
    (*m_ops[op])(inarray, output0, output1, NullNekDouble1DArray, wsp);
}

void Collection::ApplyOperator(const OperatorType &op, const Array<OneD, const NekDouble> &inarray,
                               Array<OneD, NekDouble> &output0, Array<OneD, NekDouble> &output1,
                               Array<OneD, NekDouble> &output2,
                               const StdRegions::ConstFactorMap &factors)
{
    Array<OneD, NekDouble> wsp(m_ops[op]->GetWspSize());
    (*m_ops[op])(inarray, output0, output1, output2, wsp, factors);
}

void Collection::ApplyOperator(const OperatorType &op, int dir,
                               const Array<OneD, const NekDouble> &inarray,
                               Array<OneD, NekDouble> &output)
{
    Array<OneD, NekDouble> wsp(m_ops[op]->GetWspSize());
    (*m_ops[op])(dir, inarray, output, wsp);
}

}  // namespace Collections
}  // namespace Nektar
