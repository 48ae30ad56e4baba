// Nektar::MatrixFree operator layer, implemented on the gfx950 C ABI.
//
// Mirrors MatrixFreeOps/Operator.hpp:21-310 and Operator.cpp:6-56: the Operator base
// (Ndof / NeedsDF / NeedsJac / SetJac / SetDF), the BwdTrans / IProduct / Helmholtz /
// PhysDeriv / IProductWRTDerivBase operator bases with operator()(in, out),
// Helmholtz::SetLambda, the string-keyed
// OperatorFactory ("Helmholtz_Hex_Deformed", ...) and GetOpstring.  The registered
// implementations run on the GPU through include/ithaca_mf.h.
//
// Differences from the reference, by design:
//  * SetJac / SetDF take the un-interleaved CoalescedGeomData::GetJac / GetDerivFactors
//    arrays (the vec_t interleave is a CPU SIMD layout; the GPU packs its own).
//  * nElmt needs no padding to a SIMD width.
//  * Ndof() is defined for every operator, including Helmholtz (the reference returns
//    0 there, Operator.hpp:29-32).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "Basis.hpp"
#include "Nektar.hpp"

struct imf_op_s;

namespace Nektar {
namespace MatrixFree {

class Operator {
public:
    virtual ~Operator();
    virtual NekDouble Ndof();
    virtual bool NeedsDF() { return false; }
    virtual bool NeedsJac() { return false; }
    // jac: nElmt*nq^dim (deformed) or nElmt values; df: dim*dim blocks of the same size.
    virtual void SetJac(const std::shared_ptr<std::vector<NekDouble>> &jac);
    virtual void SetDF(const std::shared_ptr<std::vector<NekDouble>> &df);
    // GPU-resident variants of SetJac / SetDF (device pointers, n doubles).
    void SetJacDevice(const NekDouble *d_jac, size_t n);
    void SetDFDevice(const NekDouble *d_df, size_t n);
    // operator()(in, out) on device arrays, asynchronous on a hipStream_t (void *).
    void ApplyDevice(const NekDouble *d_in, NekDouble *d_out, void *stream = nullptr);
    imf_op_s *Handle() const { return m_op; }

protected:
    Operator(const std::string &key, std::vector<LibUtilities::BasisSharedPtr> basis, int nElmt);
    void ApplyHost(const NekDouble *in, NekDouble *out);

    std::vector<LibUtilities::BasisSharedPtr> m_basis;
    int m_nElmt;
    imf_op_s *m_op = nullptr;
};

using OperatorSharedPtr = std::shared_ptr<Operator>;

class BwdTrans : public Operator {
public:
    BwdTrans(const std::string &key, std::vector<LibUtilities::BasisSharedPtr> basis, int nElmt)
        : Operator(key, basis, nElmt)
    {
    }
    virtual void operator()(const Array<OneD, const NekDouble> &input, Array<OneD, NekDouble> &output);
};

class IProduct : public Operator {
public:
    IProduct(const std::string &key, std::vector<LibUtilities::BasisSharedPtr> basis, int nElmt)
        : Operator(key, basis, nElmt)
    {
    }
    bool NeedsJac() final { return true; }
    virtual void operator()(const Array<OneD, const NekDouble> &input, Array<OneD, NekDouble> &output);
};

class Helmholtz : public Operator {
public:
    Helmholtz(const std::string &key, std::vector<LibUtilities::BasisSharedPtr> basis, int nElmt)
        : Operator(key, basis, nElmt)
    {
    }
    bool NeedsJac() final { return true; }
    bool NeedsDF() final { return true; }
    virtual void operator()(const Array<OneD, const NekDouble> &input, Array<OneD, NekDouble> &output);
    void SetLambda(NekDouble lambda);

protected:
    NekDouble m_lambda = 1.0;
};

// MatrixFreeOps/PhysDeriv.h:18-48: one input of quadrature values, dim outputs
// out[d] = d/dx_d of the input (nElmt*Q values each).
class PhysDeriv : public Operator {
public:
    PhysDeriv(const std::string &key, std::vector<LibUtilities::BasisSharedPtr> basis, int nElmt)
        : Operator(key, basis, nElmt)
    {
    }
    bool NeedsDF() final { return true; }
    virtual void operator()(const Array<OneD, const NekDouble> &input,
                            Array<OneD, Array<OneD, NekDouble>> &output);
};

// MatrixFreeOps/IProductWRTDerivBase.h:18-48: dim inputs of quadrature values, one output
// of coefficients.
class IProductWRTDerivBase : public Operator {
public:
    IProductWRTDerivBase(const std::string &key, std::vector<LibUtilities::BasisSharedPtr> basis,
                         int nElmt)
        : Operator(key, basis, nElmt)
    {
    }
    bool NeedsJac() final { return true; }
    bool NeedsDF() final { return true; }
    virtual void operator()(const Array<OneD, Array<OneD, NekDouble>> &input,
                            Array<OneD, NekDouble> &output);
};

using OperatorFactory =
    LibUtilities::NekFactory<std::string, Operator, std::vector<LibUtilities::BasisSharedPtr>, int>;

OperatorFactory &GetOperatorFactory();

std::string GetOpstring(LibUtilities::ShapeType shape, bool deformed = false);

}  // namespace MatrixFree
}  // namespace Nektar
