// Nektar::Collections operator layer with the gfx950 implementation type.
//
// Mirrors Collections/Operator.h:40-188 (OperatorType, ImplementationType, OperatorKey,
// Operator::operator()(in, out0, out1, out2, wsp, factors), the OperatorKey factory),
// Collections/Collection.{h,cpp} (Initialise / ApplyOperator) and the *_MatrixFree
// adapters (Collections/BwdTrans.cpp:140-240, IProductWRTBase.cpp:168-277,
// Helmholtz.cpp:391-503, PhysDeriv.cpp:270-400, IProductWRTDerivBase.cpp:274-400).  One value, eMatrixFreeHip, is appended to
// ImplementationType; every existing value keeps its number.
//
// The reference operators receive vector<StdExpansionSharedPtr>; here a collection is
// described by CollectionExp (shape, bases, element count) and its geometry by
// CoalescedGeomData holding the GetJac / GetDerivFactors arrays.
#pragma once

#include <memory>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "MatrixFreeOps.hpp"

namespace Nektar {
namespace Collections {

enum OperatorType {
    eBwdTrans,
    eHelmholtz,
    eIProductWRTBase,
    eIProductWRTDerivBase,
    ePhysDeriv,
    SIZE_OperatorType
};

const char *const OperatorTypeMap[] = {"BwdTrans", "Helmholtz", "IProductWRTBase",
                                       "IProductWRTDerivBase", "PhysDeriv"};

enum ImplementationType {
    eNoImpType,
    eNoCollection,
    eIterPerExp,
    eStdMat,
    eSumFac,
    eMatrixFree,
    eMatrixFreeHip,  // gfx950 matrix-free operators (this library)
    SIZE_ImplementationType
};

const char *const ImplementationTypeMap[] = {"NoImplementationType", "NoCollection",
                                             "IterPerExp", "StdMat", "SumFac", "MatrixFree",
                                             "MatrixFreeHip"};

using ExpansionIsNodal = bool;
using OperatorImpMap = std::map<OperatorType, ImplementationType>;
OperatorImpMap SetFixedImpType(ImplementationType defaultType);

// Description of a collection of elements of one shape, order and deformation type
// (what the reference adapters query from pCollExp[0]->GetStdExp()).
struct CollectionExp {
    LibUtilities::ShapeType shape;
    std::vector<LibUtilities::BasisSharedPtr> basis;  // one per direction
    int nElmt;
    bool nodal = false;
    int GetShapeDimension() const { return (int)basis.size(); }
    int GetNcoeffs() const;
    int GetTotPoints() const;
};

// Geometric factors of a collection, in the CoalescedGeomData::GetJac /
// GetDerivFactors layouts (Collections/CoalescedGeomData.cpp:53-112, :251-313).
class CoalescedGeomData {
public:
    CoalescedGeomData(bool deformed, std::shared_ptr<std::vector<NekDouble>> jac,
                      std::shared_ptr<std::vector<NekDouble>> df)
        : m_deformed(deformed), m_jac(jac), m_df(df)
    {
    }
    bool IsDeformed() const { return m_deformed; }
    const std::shared_ptr<std::vector<NekDouble>> &GetJac() const { return m_jac; }
    const std::shared_ptr<std::vector<NekDouble>> &GetDerivFactors() const { return m_df; }

private:
    bool m_deformed;
    std::shared_ptr<std::vector<NekDouble>> m_jac, m_df;
};

using CoalescedGeomDataSharedPtr = std::shared_ptr<CoalescedGeomData>;

class Operator {
public:
    Operator(const CollectionExp &exp, CoalescedGeomDataSharedPtr geom);
    virtual ~Operator() = default;

    virtual void operator()(const Array<OneD, const NekDouble> &input,
                            Array<OneD, NekDouble> &output0, Array<OneD, NekDouble> &output1,
                            Array<OneD, NekDouble> &output2,
                            Array<OneD, NekDouble> &wsp = NullNekDouble1DArray,
                            const StdRegions::ConstFactorMap &factors =
                                StdRegions::NullConstFactorMap) = 0;

    virtual void operator()(int dir, const Array<OneD, const NekDouble> &input,
                            Array<OneD, NekDouble> &output,
                            Array<OneD, NekDouble> &wsp = NullNekDouble1DArray) = 0;

    unsigned int GetWspSize() { return m_wspSize; }
    unsigned int GetNumElmt() { return m_numElmt; }

protected:
    bool m_isDeformed;
    unsigned int m_numElmt;
    unsigned int m_wspSize = 0;
};

using OperatorSharedPtr = std::shared_ptr<Operator>;

using OperatorKey = std::tuple<LibUtilities::ShapeType, OperatorType, ImplementationType,
                               ExpansionIsNodal>;

std::ostream &operator<<(std::ostream &os, const OperatorKey &p);

using OperatorFactory = LibUtilities::NekFactory<OperatorKey, Operator, const CollectionExp &,
                                                 CoalescedGeomDataSharedPtr>;

OperatorFactory &GetOperatorFactory();

// Collection::Initialise / ApplyOperator (Collections/Collection.cpp:46-84,
// Collection.h:119-175).
class Collection {
public:
    Collection(const CollectionExp &exp, CoalescedGeomDataSharedPtr geom,
               const OperatorImpMap &impTypes);
    void Initialise(const OperatorType opType);
    bool HasOperator(const OperatorType &op) { return m_ops.count(op) > 0; }
    OperatorSharedPtr GetOpSharedPtr(const OperatorType &op) { return m_ops[op]; }
    void ApplyOperator(const OperatorType &op, const Array<OneD, const NekDouble> &inarray,
                       Array<OneD, NekDouble> &output,
                       const StdRegions::ConstFactorMap &factors = StdRegions::NullConstFactorMap);
    // Two-output form (2D PhysDeriv; 2D IProductWRTDerivBase as in0, in1, out).
    void ApplyOperator(const OperatorType &op, const Array<OneD, const NekDouble> &inarray,
                       Array<OneD, NekDouble> &output0, Array<OneD, NekDouble> &output1);
    // Three-output form (3D PhysDeriv; 3D IProductWRTDerivBase as in0, in1, in2, out).
    void ApplyOperator(const OperatorType &op, const Array<OneD, const NekDouble> &inarray,
                       Array<OneD, NekDouble> &output0, Array<OneD, NekDouble> &output1,
                       Array<OneD, NekDouble> &output2,
                       const StdRegions::ConstFactorMap &factors = StdRegions::NullConstFactorMap);
    // One direction (PhysDeriv).
    void ApplyOperator(const OperatorType &op, int dir, const Array<OneD, const NekDouble> &inarray,
                       Array<OneD, NekDouble> &output);

private:
    std::unordered_map<int, OperatorSharedPtr> m_ops;
    CollectionExp m_exp;
    CoalescedGeomDataSharedPtr m_geomData;
    OperatorImpMap m_impTypes;
};

}  // namespace Collections
}  // namespace Nektar
