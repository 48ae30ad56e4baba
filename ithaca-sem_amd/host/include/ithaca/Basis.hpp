// LibUtilities::Basis for the modified bases of the MatrixFree shapes
// (LibUtilities/Foundations/Basis.cpp:217-668; Points/GaussPoints.cpp): Modified_A on
// Gauss-Lobatto-Legendre points, Modified_B on Gauss-Radau alpha = 1 points, Modified_C and
// ModifiedPyr_C on Gauss-Radau alpha = 2 points, with the accessors the MatrixFree Helper
// reads (MatrixFreeOps/Operator.hpp:223-273).
#pragma once

#include <memory>
#include <vector>

#include "Nektar.hpp"

namespace Nektar {
namespace LibUtilities {

enum BasisType { eNoBasisType, eModified_A, eModified_B, eModified_C, eModifiedPyr_C };
enum PointsType {
    eNoPointsType,
    eGaussLobattoLegendre,
    eGaussRadauMAlpha1Beta0,
    eGaussRadauMAlpha2Beta0
};

class PointsKey {
public:
    PointsKey(int nq, PointsType t) : m_nq(nq), m_type(t) {}
    int GetNumPoints() const { return m_nq; }
    PointsType GetPointsType() const { return m_type; }

private:
    int m_nq;
    PointsType m_type;
};

class BasisKey {
public:
    BasisKey(BasisType b, int nm, const PointsKey &pk) : m_type(b), m_nm(nm), m_pk(pk) {}
    BasisType GetBasisType() const { return m_type; }
    int GetNumModes() const { return m_nm; }
    const PointsKey &GetPointsKey() const { return m_pk; }

private:
    BasisType m_type;
    int m_nm;
    PointsKey m_pk;
};

class Basis {
public:
    explicit Basis(const BasisKey &key);
    BasisType GetBasisType() const { return m_key.GetBasisType(); }
    int GetNumModes() const { return m_key.GetNumModes(); }
    int GetNumPoints() const { return m_key.GetPointsKey().GetNumPoints(); }
    PointsType GetPointsType() const { return m_key.GetPointsKey().GetPointsType(); }
    const BasisKey &GetBasisKey() const { return m_key; }
    // bdata[mode*nq + i] = phi_mode(z_i); dbdata[mode*nq + i] = phi_mode'(z_i) (Modified_A:
    // nm modes; Modified_B: nm(nm+1)/2; Modified_C: nm(nm+1)(nm+2)/6; ModifiedPyr_C:
    // nm(nm+1)(2nm+1)/6)
    const std::vector<NekDouble> &GetBdata() const { return m_bdata; }
    const std::vector<NekDouble> &GetDbdata() const { return m_dbdata; }
    const std::vector<NekDouble> &GetZ() const { return m_z; }
    const std::vector<NekDouble> &GetW() const { return m_w; }
    // Flat derivative matrix D[a*nq + b] = l_a'(z_b) (column-major D(b, a)).
    const std::vector<NekDouble> &GetD() const { return m_D; }

private:
    BasisKey m_key;
    std::vector<NekDouble> m_z, m_w, m_D, m_bdata, m_dbdata;
};

using BasisSharedPtr = std::shared_ptr<Basis>;

inline BasisSharedPtr MakeModifiedA(int nm, int nq)
{
    return std::make_shared<Basis>(BasisKey(eModified_A, nm, PointsKey(nq, eGaussLobattoLegendre)));
}
inline BasisSharedPtr MakeModifiedB(int nm, int nq)
{
    return std::make_shared<Basis>(BasisKey(eModified_B, nm, PointsKey(nq, eGaussRadauMAlpha1Beta0)));
}
inline BasisSharedPtr MakeModifiedC(int nm, int nq)
{
    return std::make_shared<Basis>(BasisKey(eModified_C, nm, PointsKey(nq, eGaussRadauMAlpha2Beta0)));
}
inline BasisSharedPtr MakeModifiedPyrC(int nm, int nq)
{
    return std::make_shared<Basis>(
        BasisKey(eModifiedPyr_C, nm, PointsKey(nq, eGaussRadauMAlpha2Beta0)));
}

}  // namespace LibUtilities
}  // namespace Nektar
