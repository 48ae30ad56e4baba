"""Adjoint gradients through torch.autograd (FP64) against central finite differences
(step 1e-6) and the CPU oracle gradient.  Bar: 1e-6 relative agreement."""
import numpy as np
import pytest
import torch

import ithaca_hip as H
import oracle as O
import oracle.rom as R
from ithaca_hip import mesh, rom
from ithaca_hip.autograd import BwdTransFn, L2Misfit, reduced_solve

pytestmark = pytest.mark.gpu
H_FD = 1e-6


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", torch.float64)


@pytest.mark.parametrize("kind", ["normalised", "nonsymmetric"])
@pytest.mark.parametrize("method", [rom.LU_FP64, rom.LU_FP32_IR])
def test_reduced_solve_parameter_gradient(method, kind):
    """With non-symmetric A_q the adjoint must solve with A^T and form -lambda^T A_q x
    (not x^T A_q lambda); the oracle does both with numpy on the transposed matrices."""
    bt = rom.affine_batch(n_inst=512, seed=3, kind=kind)
    print(f"[case] reduced-solve adjoint method={method} kind={kind}", flush=True)
    B = rom.Batch(bt["N"], bt["Q"])
    A = _dev(bt["A"])
    theta = _dev(bt["theta"]).requires_grad_(True)
    b = _dev(bt["b"])
    l = _dev(np.random.default_rng(4).standard_normal(bt["b"].size))
    x = reduced_solve(B, A, theta, b, method)
    J = torch.dot(l, x)
    J.backward()
    g = theta.grad.detach().clone()
    # CPU oracle gradient
    xr = R.solve(bt)
    lr = R.solve(bt, rhs=l.cpu().numpy(), transpose=True)
    gr = R.param_grad(bt, lr, xr)
    assert np.linalg.norm(g.cpu().numpy() - gr) / np.linalg.norm(gr) < 1e-8
    # central finite differences along random directions in parameter space
    rng = np.random.default_rng(5)
    for _ in range(3):
        d = _dev(rng.standard_normal(bt["theta"].size))
        with torch.no_grad():
            Jp = torch.dot(l, B.solve(A, (theta + H_FD * d).contiguous(), b))
            Jm = torch.dot(l, B.solve(A, (theta - H_FD * d).contiguous(), b))
        fd = ((Jp - Jm) / (2 * H_FD)).item()
        an = torch.dot(g, d).item()
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_bwdtrans_misfit_gradient_full_hex_mesh():
    """J(u_hat) = 0.5 ||BwdTrans(u_hat) - f||^2_{L2} on the 48^3 hex mesh; gradient by the
    IProductWRTBase kernel, by autograd through BwdTrans^T, by FD and by the oracle."""
    w = mesh.workload_hex_helmholtz()
    nm, nq, E = w["nm"], w["nq"], w["nElmt"]
    print("[case] misfit adjoint on the full hex mesh", flush=True)
    b = H.basis_modified_a(nm, nq)
    gj, gd = H.geom_device("hex", _dev(w["verts"]), nq, b["z"], True)
    bwd = H.Operator("BwdTrans_Hex_Regular", nm, nq, E)
    ipr = H.Operator("IProduct_Hex_Deformed", nm, nq, E)
    ipr.set_jac(gj)
    W = torch.from_numpy(np.kron(np.kron(b["w"], b["w"]), b["w"])).cuda()
    jw = (gj.view(E, -1) * W).reshape(-1)
    # random smooth field f (seed 4): a few low Fourier modes at the quadrature points
    xyz = torch.from_numpy(O.coords_hex(w["verts"], nq, O.basis(nm, nq)["z"])).cuda()
    rng = np.random.default_rng(4)
    f = torch.zeros(E * nq ** 3, dtype=torch.float64, device="cuda")
    for _ in range(6):
        k = rng.integers(1, 4, size=3)
        a = rng.standard_normal()
        f += a * torch.sin(np.pi * k[0] * xyz[0]) * torch.cos(np.pi * k[1] * xyz[1]) * \
            torch.sin(np.pi * k[2] * xyz[2] + 0.3)
    u = _dev(w["u"]).requires_grad_(True)
    f.requires_grad_(True)
    J = L2Misfit.apply(u, f, bwd, ipr, jw)
    J.backward()
    g1 = u.grad.detach().clone()
    gf = f.grad.detach().clone()
    f = f.detach()
    # gradient in the target f: -diag(wJ) r, checked by central differences
    with torch.no_grad():
        d = torch.rand(f.numel(), dtype=torch.float64, device="cuda",
                       generator=torch.Generator("cuda").manual_seed(11)) - 0.5
        Jp = L2Misfit.apply(u.detach(), f + H_FD * d, bwd, ipr, jw)
        Jm = L2Misfit.apply(u.detach(), f - H_FD * d, bwd, ipr, jw)
        fd = ((Jp - Jm) / (2 * H_FD)).item()
        an = torch.dot(gf, d).item()
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    # autograd through the Euclidean transpose of BwdTrans
    u2 = _dev(w["u"]).requires_grad_(True)
    r = BwdTransFn.apply(u2, bwd) - f
    (0.5 * torch.dot(jw * r, r)).backward()
    g2 = u2.grad.detach()
    assert (torch.linalg.norm(g1 - g2) / torch.linalg.norm(g1)).item() < 1e-12
    # oracle gradient on a slice of elements (the operator is element-local)
    Es = 512
    bo = O.basis(nm, nq)
    jac_s, _ = O.geom_hex(w["verts"][:Es], nq, bo["z"], True)
    fs = f[:Es * nq ** 3].cpu().numpy()
    rs = O.bwdtrans("hex", bo, Es, w["u"][:Es * nm ** 3]) - fs
    go = O.iproduct("hex", bo, Es, True, jac_s, rs)
    gs = g1[:Es * nm ** 3].cpu().numpy()
    assert np.linalg.norm(gs - go) / np.linalg.norm(go) < 1e-12
    # central finite differences along random directions
    with torch.no_grad():
        for s in range(3):
            d = torch.rand(u.numel(), dtype=torch.float64, device="cuda",
                           generator=torch.Generator("cuda").manual_seed(s)) - 0.5
            Jp = L2Misfit.apply(u.detach() + H_FD * d, f, bwd, ipr, jw)
            Jm = L2Misfit.apply(u.detach() - H_FD * d, f, bwd, ipr, jw)
            fd = ((Jp - Jm) / (2 * H_FD)).item()
            an = torch.dot(g1, d).item()
            assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_helmsolve_adjoint():
    """J(rhs, lambda) = l . u with u = A(lambda)^{-1} rhs, the assembled deformed-hex
    Helmholtz solve (Dirichlet modes fixed at zero): the rhs gradient equals the sparse
    direct solve A^{-1} l on the free modes (assembled on the CPU from independent dense
    element matrices), and the lambda and rhs gradients match central finite differences."""
    import scipy.sparse.linalg as spla
    from test_assembly_cpu import _mesh, helmholtz_system
    from ithaca_hip import assembly
    from ithaca_hip.autograd import helm_solve
    verts, ids = _mesh(2, True, seed=3, jitter=0.05)
    nm, nq, E = 5, 6, len(ids)
    print("[case] helmsolve adjoint", flush=True)
    b = H.basis_modified_a(nm, nq)
    jac, df = H.geom_device("hex", _dev(verts), nq, b["z"], True)
    op = H.Operator("Helmholtz_Hex_Deformed", nm, nq, E, basis=b)
    op.set_jac(jac)
    op.set_df(df)
    mass = H.Operator("Mass_Hex_Deformed", nm, nq, E, basis=b)
    mass.set_jac(jac)
    asm = assembly.Assembly(ids, nm)
    ng = asm.nglobal
    rng = np.random.default_rng(7)
    rhs = _dev(rng.standard_normal(ng)).requires_grad_(True)
    lam = torch.tensor(3.0, dtype=torch.float64, requires_grad=True)
    l = _dev(rng.standard_normal(ng))
    J = torch.dot(l, helm_solve(asm, op, rhs, lam, mass, tol=1e-13))
    J.backward()
    A, _, dm, _, _ = helmholtz_system(verts, ids, nm, nq, 3.0, lambda X: np.zeros(len(X)))
    free = ~dm
    wref = np.zeros(ng)
    wref[free] = spla.spsolve(A[free][:, free].tocsc(), l.cpu().numpy()[free])
    g = rhs.grad.cpu().numpy()
    assert np.linalg.norm(g - wref) <= 1e-9 * np.linalg.norm(wref)
    assert np.all(g[dm] == 0.0)

    # dJ/dlambda = -w^T M u with u, w from sparse direct solves and M = A(4) - A(3)
    A4 = helmholtz_system(verts, ids, nm, nq, 4.0, lambda X: np.zeros(len(X)))[0]
    Mf = (A4 - A)[free][:, free]
    uref = spla.spsolve(A[free][:, free].tocsc(), rhs.detach().cpu().numpy()[free])
    gref = -wref[free] @ (Mf @ uref)
    an = lam.grad.item()
    assert abs(an - gref) <= 1e-8 * abs(gref)

    def Jof(r, lv):
        with torch.no_grad():
            return torch.dot(l, helm_solve(asm, op, r, torch.tensor(lv, dtype=torch.float64),
                                           None, 1e-13)).item()
    # finite differences: a step large enough that the solver's 1e-13 residual does not
    # dominate, small enough that the O(h^2) term stays below the bar
    h = 1e-3
    fd = (Jof(rhs.detach(), 3.0 + h) - Jof(rhs.detach(), 3.0 - h)) / (2 * h)
    print(f"dJ/dlambda adjoint {an:.10e} sparse {gref:.10e} fd {fd:.10e}")
    assert abs(fd - an) <= 1e-6 * abs(an)
    d = _dev(rng.standard_normal(ng))
    fd = (Jof(rhs.detach() + H_FD * d, 3.0) - Jof(rhs.detach() - H_FD * d, 3.0)) / (2 * H_FD)
    an = torch.dot(rhs.grad, d).item()
    assert abs(fd - an) <= 1e-6 * abs(an)


def test_helmsolve_adjoint_variable_order():
    """The same adjoint over a variable-order assembly (one Helmholtz and one mass operator
    per run of equal-order elements): gradients equal sparse direct solves of the
    variable-order system."""
    import scipy.sparse.linalg as spla
    from test_assembly_cpu import _mesh
    from test_assembly_varp import varp_system
    from ithaca_hip import assembly
    from ithaca_hip.autograd import helm_solve
    verts, ids = _mesh(2, True, seed=3, jitter=0.05)
    nmv = np.repeat([4, 6], 4)  # elements are z-major: lower and upper layers
    print("[case] helmsolve adjoint, variable order", flush=True)
    ops, masses = [], []
    for e0, e1, nm in ((0, 4, 4), (4, 8, 6)):
        nq = nm + 1
        b = H.basis_modified_a(nm, nq)
        jac, df = H.geom_device("hex", _dev(verts[e0:e1]), nq, b["z"], True)
        op = H.Operator("Helmholtz_Hex_Deformed", nm, nq, e1 - e0, basis=b)
        op.set_jac(jac)
        op.set_df(df)
        ms = H.Operator("Mass_Hex_Deformed", nm, nq, e1 - e0, basis=b)
        ms.set_jac(jac)
        ops.append(op)
        masses.append(ms)
    asm = assembly.Assembly(ids, nmv)
    ng = asm.nglobal
    rng = np.random.default_rng(8)
    rhs = _dev(rng.standard_normal(ng)).requires_grad_(True)
    lam = torch.tensor(2.0, dtype=torch.float64, requires_grad=True)
    l = _dev(rng.standard_normal(ng))
    torch.dot(l, helm_solve(asm, ops, rhs, lam, masses, tol=1e-13)).backward()
    zero = lambda X: np.zeros(len(X))  # noqa: E731
    A, _, dm, _, _ = varp_system(verts, ids, nmv, 2.0, zero)
    A3 = varp_system(verts, ids, nmv, 3.0, zero)[0]
    free = ~dm
    Af = A[free][:, free].tocsc()
    w = spla.spsolve(Af, l.cpu().numpy()[free])
    u = spla.spsolve(Af, rhs.detach().cpu().numpy()[free])
    g = rhs.grad.cpu().numpy()
    assert np.linalg.norm(g[free] - w) <= 1e-9 * np.linalg.norm(w)
    gl = -w @ ((A3 - A)[free][:, free] @ u)
    assert abs(lam.grad.item() - gl) <= 1e-8 * abs(gl)
