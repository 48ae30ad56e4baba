"""GPU parity of the batched online reduced solves against the FP64 CPU oracle.

Bar: every solution within 1e-10 relative error of the FP64 oracle, for the
FP64 LU path and for FP32 LU + FP64 iterative refinement."""
import numpy as np
import pytest
import torch

import oracle.rom as R
from ithaca_hip import rom

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _dev(a, dt=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt)


def _max_rel(batch, got, ref):
    worst = 0.0
    for i in range(len(batch["N"])):
        _, _, sl = R.instance(batch, i)
        worst = max(worst, np.linalg.norm(got[sl] - ref[sl]) / np.linalg.norm(ref[sl]))
    return worst


@pytest.fixture(scope="module")
def small():
    bt = affine_batch_edges()
    B = rom.Batch(bt["N"], bt["Q"])
    assert B.a_size == bt["A"].size and B.v_size == bt["b"].size
    assert np.array_equal(B.a_off, bt["a_off"]) and np.array_equal(B.v_off, bt["v_off"])
    return bt, B


def affine_batch_edges():
    bt = rom.affine_batch(n_inst=300, seed=3, nmin=1, nmax=256)
    return bt


@pytest.mark.parametrize("method", [rom.LU_FP64, rom.LU_FP32_IR])
def test_batched_solve_parity(small, method):
    bt, B = small
    print(f"[case] batched solve method={method} n={len(bt['N'])}", flush=True)
    A, th, b = _dev(bt["A"]), _dev(bt["theta"]), _dev(bt["b"])
    iters = torch.zeros(B.n, dtype=torch.int32, device="cuda")
    status = torch.full((B.n,), -1, dtype=torch.int32, device="cuda")
    x = B.solve(A, th, b, method=method, iters=iters, status=status)
    ref = R.solve(bt)
    err = _max_rel(bt, x.cpu().numpy(), ref)
    print(f"max rel err {err:.3e}, iters max {int(iters.max())}")
    assert err < TOL
    assert int(status.max()) == 0 and int(status.min()) == 0
    if method == rom.LU_FP32_IR:
        assert int(iters.max()) <= 6


@pytest.mark.parametrize("method", [rom.LU_FP64, rom.LU_FP32_IR])
def test_transpose_solve(small, method):
    bt, B = small
    A, th, b = _dev(bt["A"]), _dev(bt["theta"]), _dev(bt["b"])
    x = B.solve(A, th, b, method=method, transpose=True)
    assert _max_rel(bt, x.cpu().numpy(), R.solve(bt, transpose=True)) < TOL


def test_every_tile_variant(small):
    bt, B = small
    A, th, b = _dev(bt["A"]), _dev(bt["theta"]), _dev(bt["b"])
    ref = R.solve(bt)
    for v, name in enumerate(B.variants()):
        for method in (rom.LU_FP64, rom.LU_FP32_IR):
            print(f"[case] variant {name} method {method}", flush=True)
            B.set_variant(method, v)
            x = B.solve(A, th, b, method=method)
            assert _max_rel(bt, x.cpu().numpy(), ref) < TOL, name
    chosen = B.autotune(A, th, b, torch.empty_like(b), rom.LU_FP64, reps=1)
    assert 0 <= chosen < len(B.variants())


def test_graph_replay_matches_eager(small):
    bt, B = small
    A, th, b = _dev(bt["A"]), _dev(bt["theta"]), _dev(bt["b"])
    eager = B.solve(A, th, b, method=rom.LU_FP32_IR).clone()
    x = torch.zeros_like(b)
    B.capture(A, th, b, x, method=rom.LU_FP32_IR)
    for r in range(4):
        x.zero_()
        B.replay()
        torch.cuda.synchronize()
        assert torch.equal(x, eager), r
        _ = (A[:1000] * 2.0).sum().item()  # eager work between replays


def test_param_grad(small):
    bt, B = small
    rng = np.random.default_rng(4)
    l = rng.standard_normal(bt["b"].size)
    A, th, b = _dev(bt["A"]), _dev(bt["theta"]), _dev(bt["b"])
    x = B.solve(A, th, b)
    lam = B.solve(A, th, _dev(l), transpose=True)
    g = B.param_grad(A, lam, x).cpu().numpy()
    xr = R.solve(bt)
    lr = R.solve(bt, rhs=l, transpose=True)
    gr = R.param_grad(bt, lr, xr)
    assert np.linalg.norm(g - gr) / np.linalg.norm(gr) < 1e-10


def test_full_batch_4096():
    bt = rom.affine_batch()  # 4096 instances, seed 3
    print("[case] full 4096-instance batch", flush=True)
    B = rom.Batch(bt["N"], bt["Q"])
    A, th, b = _dev(bt["A"]), _dev(bt["theta"]), _dev(bt["b"])
    ref = R.solve(bt)
    for method in (rom.LU_FP64, rom.LU_FP32_IR):
        x = B.solve(A, th, b, method=method)
        err = _max_rel(bt, x.cpu().numpy(), ref)
        print(f"method {method}: max rel err vs oracle {err:.3e}")
        assert err < TOL


@pytest.mark.parametrize("method", [rom.LU_FP64, rom.LU_FP32_IR])
def test_status_codes_and_degenerate_batches(method):
    """A singular instance (every A_q zero) reports a zero pivot (status 1) without
    disturbing the other instances; one-term systems (Q = 1) of every size solve; a
    refinement budget of zero steps leaves the FP32 solutions unconverged (status 2, no
    refinement) and the default budget converges them; a batch of one instance solves."""
    bt = rom.affine_batch(n_inst=40, seed=5, nmin=1, nmax=256, qmin=1, qmax=3)
    k = 7
    Aq, _, _ = R.instance(bt, k)
    a0 = int(bt["a_off"][k])
    bt["A"][a0:a0 + Aq.size] = 0.0
    ok = [i for i in range(40) if i != k]
    print(f"[case] status codes method={method}", flush=True)
    B = rom.Batch(bt["N"], bt["Q"])
    A, th, b = _dev(bt["A"]), _dev(bt["theta"]), _dev(bt["b"])
    iters = torch.zeros(B.n, dtype=torch.int32, device="cuda")
    status = torch.full((B.n,), -1, dtype=torch.int32, device="cuda")
    x = B.solve(A, th, b, method=method, iters=iters, status=status).cpu().numpy()
    st = status.cpu().numpy()
    assert st[k] == 1 and np.all(st[ok] == 0)
    ref = R.solve(bt, idx=ok)
    for i in ok:
        _, _, sl = R.instance(bt, i)
        assert np.linalg.norm(x[sl] - ref[sl]) <= TOL * np.linalg.norm(ref[sl])
    if method == rom.LU_FP32_IR:
        status.fill_(-1)
        B.solve(A, th, b, method=method, max_iter=0, iters=iters, status=status)
        st0 = status.cpu().numpy()
        assert np.all(st0[ok] == 2) and int(iters.max()) == 0 and st0[k] == 1
    one = rom.affine_batch(n_inst=1, seed=6, nmin=256, nmax=256, qmin=1, qmax=1)
    B1 = rom.Batch(one["N"], one["Q"])
    x1 = B1.solve(_dev(one["A"]), _dev(one["theta"]), _dev(one["b"]), method=method)
    r1 = R.solve(one)
    assert np.linalg.norm(x1.cpu().numpy() - r1) <= TOL * np.linalg.norm(r1)
