"""CPU pins of the reduced-solve oracle: LAPACK LU against the reference's own solver
(Eigen 3.3.7 ColPivHouseholderQR, CoupledLinearNS_ROM.cpp:4053, built in oracle/_ref),
generator properties, and the adjoint parameter gradient against finite differences."""
import numpy as np
import pytest

import oracle.rom as R
from ithaca_hip.rom import affine_batch


@pytest.fixture(scope="module")
def batch():
    return affine_batch(n_inst=40, seed=3, nmin=1, nmax=64)


def test_generator_properties(batch):
    for i in range(len(batch["N"])):
        Aq, th, sl = R.instance(batch, i)
        ev = np.linalg.eigvalsh(Aq[0])
        assert np.allclose(Aq[0], Aq[0].T, atol=1e-9 * ev.max())
        assert ev.min() > 1.0 - 1e-9 and ev.max() < 1e4 * (1 + 1e-9)
        for q in range(1, len(th)):
            assert np.array_equal(Aq[q], Aq[q].T)
        assert np.all((th >= 0.1) & (th <= 1.0))


def test_lapack_oracle_matches_reference_eigen_solver(batch):
    if R.eigen_lib() is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    idx = range(len(batch["N"]))
    x_lapack = R.solve(batch)
    x_eigen = R.solve_eigen(batch, idx)
    for i in idx:
        _, _, sl = R.instance(batch, i)
        err = np.linalg.norm(x_lapack[sl] - x_eigen[sl]) / np.linalg.norm(x_eigen[sl])
        assert err < 1e-10, (i, err)


def test_residual_and_transpose(batch):
    x = R.solve(batch)
    xt = R.solve(batch, transpose=True)
    for i in range(len(batch["N"])):
        Aq, th, sl = R.instance(batch, i)
        A = R.assemble(Aq, th)
        bb = batch["b"][sl]
        assert np.linalg.norm(A @ x[sl] - bb) <= 1e-12 * np.linalg.norm(A, 2) * np.linalg.norm(x[sl]) + 1e-14
        assert np.linalg.norm(A.T @ xt[sl] - bb) <= 1e-12 * np.linalg.norm(A, 2) * np.linalg.norm(xt[sl]) + 1e-14


def test_param_grad_finite_differences(batch):
    rng = np.random.default_rng(4)
    l = rng.standard_normal(len(batch["b"]))
    x = R.solve(batch)
    lam = R.solve(batch, rhs=l, transpose=True)
    g = R.param_grad(batch, lam, x)
    h = 1e-6
    for i in [0, 7, 23]:
        Aq, th, sl = R.instance(batch, i)
        t0 = int(batch["t_off"][i])
        for q in range(len(th)):
            def J(d):
                t = th.copy(); t[q] += d
                return l[sl] @ np.linalg.solve(R.assemble(Aq, t), batch["b"][sl])
            fd = (J(h) - J(-h)) / (2 * h)
            assert abs(fd - g[t0 + q]) <= 1e-6 * max(1.0, abs(fd)), (i, q, fd, g[t0 + q])
