// Reference linear algebra of the ROM path, compiled from the Eigen 3.3.7 headers that the
// reference vendors (solvers/IncNavierStokesSolver/EquationSystems/ThirdParty/Eigen) --
// TEST INFRASTRUCTURE (oracle/_ref), never part of the product.
//   colpivqr_solve : Eigen::ColPivHouseholderQR::solve, the reference's online reduced
//                    solve (CoupledLinearNS_ROM.cpp:4053)
//   bdcsvd         : Eigen::BDCSVD with ComputeThinU, the reference's POD
//                    (CoupledLinearNS_ROM.cpp:3925-3954)
//   selfadjoint_eig: Eigen::SelfAdjointEigenSolver (eigen-decomposition of a Gram matrix)
#include <Eigen/Dense>
#include <Eigen/SVD>

extern "C" {

// A row-major n x n; b, x length n.
void ref_colpivqr_solve(int n, const double *A, const double *b, double *x)
{
    Eigen::Map<const Eigen::Matrix<double, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor>> M(A, n, n);
    Eigen::Map<const Eigen::VectorXd> rhs(b, n);
    Eigen::Map<Eigen::VectorXd> out(x, n);
    out = M.colPivHouseholderQr().solve(rhs);
}

// S column-major m x k; sv (k), U column-major m x k (thin).
void ref_bdcsvd(int m, int k, const double *S, double *sv, double *U)
{
    Eigen::Map<const Eigen::MatrixXd> M(S, m, k);
    Eigen::BDCSVD<Eigen::MatrixXd> svd(M, Eigen::ComputeThinU);
    Eigen::Map<Eigen::VectorXd>(sv, svd.singularValues().size()) = svd.singularValues();
    Eigen::Map<Eigen::MatrixXd>(U, m, svd.matrixU().cols()) = svd.matrixU();
}

// G column-major symmetric n x n; ascending eigenvalues, eigenvectors column-major.
void ref_selfadjoint_eig(int n, const double *G, double *evals, double *evecs)
{
    Eigen::Map<const Eigen::MatrixXd> M(G, n, n);
    Eigen::SelfAdjointEigenSolver<Eigen::MatrixXd> es(M);
    Eigen::Map<Eigen::VectorXd>(evals, n) = es.eigenvalues();
    Eigen::Map<Eigen::MatrixXd>(evecs, n, n) = es.eigenvectors();
}
}
