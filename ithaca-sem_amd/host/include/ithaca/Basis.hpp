// LibUtilities::Basis for the Modified_A basis on Gauss-Lobatto-Legendre points
// (LibUtilities/Foundations/Basis.cpp:217, :392-421; Points/GaussPoints.cpp), with the
// accessors the MatrixFree Helper reads (MatrixFreeOps/Operator.hpp:223-273).
#pragma once

#include <memory>
#include <vector>

#include "Nektar.hpp"

namespace Nektar {
namespace LibUtilities {

enum BasisType { eNoBasisType, eModified_A };
enum PointsType { eNoPointsType, eGaussLobattoLegendre };

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
    // bdata[p*nq + i] = phi_p(z_i); dbdata[p*nq + i] = phi_p'(z_i)
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

}  // namespace LibUtilities
}  // namespace Nektar
