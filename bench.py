"""Benchmark of the matrix-free hex Helmholtz apply on one MI355X.

Workload (default): 48^3 = 110592 deformed trilinear hexes on [0,1]^3 (interior vertices
jittered by U[-0.05,0.05]*h, seed 1), modified basis P=5 (6 modes), 7 GLL points per
direction, lambda = 10, FP64, coefficients U[-1,1].  One step = one Helmholtz apply
over all elements with inputs resident in HBM.

Prints ONE JSON line.  value = GDOF/s (DOF = nElmt * 216 per apply).
roofline.achieved = algorithmic bytes per apply / mean kernel time, with the algorithmic
bytes 8*(2M + 7Q) = 22664 B per element (read u_hat, write out, 7 geometry doubles per
quadrature point), measured with HIP events on the launch stream.
cpu_baseline = the oracle's restatement of the reference CPU path (scalar element loop,
reference summation order), one OpenMP thread per core, on a bounded element sample.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "ithaca-sem_amd", "python"))

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    p.add_argument("--op", choices=["helmholtz", "bwdtrans", "iproduct", "physderiv",
                                    "iprodderiv", "copy"],
                   default="helmholtz",
                   help="copy: plain 1 GiB device copy (HBM bandwidth / counter calibration)")
    p.add_argument("--n", type=int, default=48, help="elements per direction")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-rom", action="store_true", help="skip the batched reduced solves")
    p.add_argument("--rom-instances", type=int, default=4096)
    p.add_argument("--no-pod", action="store_true", help="skip the full-size POD offline stage")
    p.add_argument("--headline-only", action="store_true",
                   help="time the Helmholtz apply only (no further sections; for counter runs)")
    p.add_argument("--pod-snapshots", type=int, default=512)
    p.add_argument("--rom-only", action="store_true",
                   help="the batched reduced solves only, one JSON object (for counter runs)")
    p.add_argument("--rom-host-gen", action="store_true",
                   help="generate the batched systems on the host (numpy) instead of the GPU")
    return p.parse_args()


def algorithmic_bytes(op, nm, nq, elt):
    M, Q = nm ** 3, nq ** 3
    if op == "helmholtz":
        return elt * (2 * M + 7 * Q)      # u_hat + out + 7 geometry values per point
    if op == "bwdtrans":
        return elt * (M + Q)              # u_hat + u
    if op == "physderiv":
        return elt * 13 * Q               # u + 9 derivative factors + 3 derivatives per point
    if op == "iprodderiv":
        return elt * (12 * Q + M)         # 3 fields + 9 w J df per point + out
    return elt * (2 * Q + M)              # f + J w + out


def cpu_baseline(w, seconds):
    """Oracle Helmholtz on a contiguous sample of elements, timed on host cores."""
    import numpy as np

    import oracle as O

    threads = min(16, os.cpu_count() or 1)
    O.set_threads(threads)
    nm, nq = w["nm"], w["nq"]
    E = min(8192, w["nElmt"])
    b = O.basis(nm, nq)
    jac, df = O.geom_hex(np.ascontiguousarray(w["verts"][:E]), nq, b["z"], True)
    u = np.ascontiguousarray(w["u"][:E * nm ** 3])
    O.helmholtz("hex", b, E, True, jac, df, w["lam"], u)  # warm-up
    times = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end or len(times) < 3:
        t0 = time.perf_counter()
        O.helmholtz("hex", b, E, True, jac, df, w["lam"], u)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    # the AVX2 SIMD ABI (four elements per vector), same sample and budget
    O.helmholtz("hex", b, E, True, jac, df, w["lam"], u, simd="avx2")
    tv = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end or len(tv) < 3:
        t0 = time.perf_counter()
        O.helmholtz("hex", b, E, True, jac, df, w["lam"], u, simd="avx2")
        tv.append(time.perf_counter() - t0)
    # one core on a smaller contiguous sample (the reference runs one MPI rank per core)
    O.set_threads(1)
    E1 = 1024
    u1 = np.ascontiguousarray(u[:E1 * nm ** 3])
    jac1 = np.ascontiguousarray(jac[:E1 * nq ** 3])
    df1 = np.ascontiguousarray(df.reshape(9, -1)[:, :E1 * nq ** 3])
    one = {}
    for simd in ("scalar", "avx2"):
        O.helmholtz("hex", b, E1, True, jac1, df1, w["lam"], u1, simd=simd)
        t1s = []
        for _ in range(5):
            t0 = time.perf_counter()
            O.helmholtz("hex", b, E1, True, jac1, df1, w["lam"], u1, simd=simd)
            t1s.append(time.perf_counter() - t0)
        one[simd] = E1 * nm ** 3 / float(np.median(t1s)) / 1e9
    O.set_threads(threads)
    return {"value": E * nm ** 3 / t / 1e9, "unit": "GDOF/s", "cores": threads,
            "kind": "port",
            "sample": f"{E} of {w['nElmt']} elements, {len(times)} applies, median; "
                      "oracle restatement of HelmholtzHexImpl (scalar element loop, "
                      f"width 1), OpenMP over contiguous element chunks, {threads} threads",
            "avx2_value": E * nm ** 3 / float(np.median(tv)) / 1e9,
            "avx2_sample": f"same elements, {len(tv)} applies, median; the same restatement "
                           "with four elements interleaved per 256-bit vector (the "
                           "reference's AVX2 SIMD ABI)",
            "single_core_value": one["scalar"],
            "single_core_avx2_value": one["avx2"],
            "single_core_sample": f"{E1} elements, 5 applies, median",
            "cpu_model": _cpu_model(), "host_cpus": os.cpu_count()}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _event_time(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def rom_bench(args):
    """Batched online reduced solves: 4096 affine systems A(mu) = sum_q theta_q A_q,
    N ~ U{16..256}, Q ~ U{3..8}, theta ~ U[0.1, 1], b ~ N(0,1) (seed 3), A_0 with
    eigenvalues log-uniform in [1, 1e4], in two distributions: "normalised" (A_q random
    symmetric of spectral norm ~0.1: SPD, condition ~1e4) and "literal" (A_q = 0.1 sym(randn)
    unnormalised, as the workload states: ~38 % indefinite, condition up to ~1e7).
    One step = assemble + factor + solve every instance in one persistent launch (two for
    the FP32 path with the FP64 fallback)."""
    import numpy as np
    import torch

    from ithaca_hip import rom

    res = {}
    for kind in ("normalised", "literal"):
        if args.rom_host_gen:  # no device QR in the process (e.g. under counter collection)
            bt = rom.affine_batch(args.rom_instances, seed=3, kind=kind)
            bt["A"] = torch.from_numpy(bt["A"]).cuda()
        else:
            bt = rom.affine_batch_device(args.rom_instances, seed=3, kind=kind)
        print(f"[bench] batched solves: {kind} batch generated", file=sys.stderr, flush=True)
        B = rom.Batch(bt["N"], bt["Q"])
        A = bt["A"]
        th = torch.from_numpy(bt["theta"]).cuda()
        b = torch.from_numpy(bt["b"]).cuda()
        n = B.n
        a_bytes = float(bt["A"].numel() * 8)
        r = {"workload": f"{n} affine reduced systems ({kind}), N~U{{16..256}}, Q~U{{3..8}}, "
                         "seed 3",
             "flops_lu_total": float(np.sum((2.0 / 3.0) * bt["N"].astype(float) ** 3)),
             "bytes_A_q_total": a_bytes,
             "hbm_floor_ms": a_bytes / (HBM_PEAK_GBS * 1e9) * 1e3}
        mixed = rom.LU_FP32_IR if kind == "normalised" else rom.LU_FP32_IR_FALLBACK
        xs, times = {}, {}
        for method, name in ((rom.LU_FP64, "fp64_lu"), (mixed, "fp32_lu_fp64_ir")):
            x = torch.empty_like(b)
            v = B.autotune(A, th, b, x, method, reps=2)
            ms = _event_time(lambda: B.solve(A, th, b, x, method=method), 10, 2)
            iters = torch.zeros(n, dtype=torch.int32, device="cuda")
            status = torch.full((n,), -1, dtype=torch.int32, device="cuda")
            B.solve(A, th, b, x, method=method, iters=iters, status=status)
            torch.cuda.synchronize()
            xs[name] = x.cpu().numpy()
            times[name] = ms
            print(f"[bench] batched solves: {kind} {name} {ms:.3f} ms", file=sys.stderr, flush=True)
            st = status.cpu().numpy()
            r[name] = {"solves_per_s": n / (ms * 1e-3), "ms_per_batch": ms,
                       "frac_of_hbm_floor": r["hbm_floor_ms"] / ms,
                       "variant": B.variants()[v],
                       "refinement_iters_max": int(iters.max()),
                       "refinement_iters_mean": float(iters.float().mean()),
                       "status_counts": np.bincount(st, minlength=4).tolist()}
            if method == rom.LU_FP32_IR_FALLBACK:
                r[name]["method"] = "FP32 LU + FP64 refinement, FP64 LU fallback (status 3)"
        if kind == "normalised":
            xg = torch.empty_like(b)
            B.capture(A, th, b, xg, method=rom.LU_FP32_IR)
            r["fp32_lu_fp64_ir"]["graph_replay_ms"] = _event_time(lambda: B.replay(), 10, 2)
        r["speedup_fp32ir_over_fp64"] = times["fp64_lu"] / times["fp32_lu_fp64_ir"]
        if not args.no_cpu:
            r["cpu_baseline"] = rom_cpu_baseline(bt, xs)
        res[kind] = r
        del bt, B, A, th, b
        torch.cuda.empty_cache()
    return res


def rom_cpu_baseline(bt, xs):
    """The reference's online solve (Eigen 3.3.7 ColPivHouseholderQR after the affine
    assembly, CoupledLinearNS_ROM.cpp:4053, :4694-4704) on every 16th instance, one core;
    its solutions are the FP64 reference for the GPU errors."""
    import numpy as np

    import oracle.rom as R

    idx = list(range(0, len(bt["N"]), 16))
    host = dict(bt)
    A_host = {}
    for i in idx:
        n, q = int(bt["N"][i]), int(bt["Q"][i])
        a0 = int(bt["a_off"][i])
        A_host[i] = bt["A"][a0:a0 + q * n * n].cpu().numpy().reshape(q, n, n)
    use_eigen = R.eigen_lib() is not None
    errs = {k: 0.0 for k in xs}
    bwd = {k: 0.0 for k in xs}
    t_tot = 0.0
    for i in idx:
        n = int(bt["N"][i])
        th = bt["theta"][int(bt["t_off"][i]):int(bt["t_off"][i]) + int(bt["Q"][i])]
        v0 = int(bt["v_off"][i])
        rhs = np.ascontiguousarray(bt["b"][v0:v0 + n])
        t0 = time.perf_counter()
        Am = np.ascontiguousarray(R.assemble(A_host[i], th))
        if use_eigen:
            xr = R.eigen_colpivqr(Am, rhs)
        else:
            xr = np.linalg.solve(Am, rhs)
        t_tot += time.perf_counter() - t0
        for k, x in xs.items():
            xi = x[v0:v0 + n]
            errs[k] = max(errs[k], float(np.linalg.norm(xi - xr) / np.linalg.norm(xr)))
            rr = rhs - Am @ xi
            bwd[k] = max(bwd[k], float(np.abs(rr).max() / (np.abs(Am).sum(1).max()
                                                         * np.abs(xi).max() + np.abs(rhs).max())))
    del host
    return {"value": len(idx) / t_tot, "unit": "solves/s", "cores": 1,
            "kind": "reference" if use_eigen else "port",
            "sample": f"every 16th instance ({len(idx)} of {len(bt['N'])}): affine assembly + "
                      + ("Eigen 3.3.7 ColPivHouseholderQR (reference solver, oracle/_ref)"
                         if use_eigen else "LAPACK LU (numpy)"),
            "max_rel_err_vs_reference": errs, "max_backward_err": bwd}


def pod_bench(args):
    """POD offline stage on the 48^3 deformed hex space (N = 23 887 872 coefficients):
    512 snapshots, each sum_k a_k prod_i sin(m_{k,i} pi x_i) over 20 fields (integer
    wavenumbers U{1..6}, a_k ~ N(0,1) 2^-k, seed 2), the fields projected to coefficients
    by the elemental L2 projection (preconditioned CG, converged to 1e-13), plus N(0,1)*1e-6
    noise added in coefficient space (torch Philox, seed 2).  Then M S (matrix-free mass),
    G = S^T M S in FP64 / FP32x3 / BF16x3 on MFMA, the 512^2 FP64 eigendecomposition,
    truncation at retained energy 1 - 1e-8, the basis V, and the Galerkin projection
    V^T H V of the Helmholtz operator (lambda = 10)."""
    import numpy as np
    import torch

    import ithaca_hip as H
    from ithaca_hip import mesh, pod

    dev = torch.device("cuda:0")
    torch.cuda.empty_cache()
    w = mesh.workload_hex_helmholtz(args.n)
    nm, nq, E = w["nm"], w["nq"], w["nElmt"]
    M, Q = nm ** 3, nq ** 3
    N = E * M
    k_snap = args.pod_snapshots
    b = H.basis_modified_a(nm, nq)
    verts = torch.from_numpy(w["verts"]).to(dev)
    jac, df = H.geom_device("hex", verts, nq, b["z"], True)
    proj = pod.ElementalProjector(E, nm, nq, b, jac)
    res = {"workload": f"{E} deformed hexes (P=5, 7 GLL pts), N={N}, {k_snap} snapshots of "
                       "20 sine-product fields (seed 2) + 1e-6 noise",
           "projection": "elemental L2 (HexExp::v_FwdTrans) by CG preconditioned with the "
                         "constant-Jacobian inverse mass, tol 1e-13; noise added in "
                         "coefficient space"}

    # 1. the 20 fields, projected
    m, a = mesh.pod_fields(20, k_snap, seed=2)
    X = mesh.hex_quad_points(verts, b["z"])
    Phi = torch.empty((20, N), dtype=torch.float64, device=dev)
    its, ress = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(20):
        f = (torch.sin(float(m[k, 0]) * np.pi * X[..., 0])
             * torch.sin(float(m[k, 1]) * np.pi * X[..., 1])
             * torch.sin(float(m[k, 2]) * np.pi * X[..., 2])).reshape(-1)
        Phi[k], it, rr = proj.project(f)
        its.append(it)
        ress.append(rr)
    torch.cuda.synchronize()
    res["projection_s"] = time.perf_counter() - t0
    res["projection_cg_iters_max"] = int(max(its))
    res["projection_rel_residual_max"] = float(max(ress))
    del X, f

    # 2. snapshots S (k_snap, N) = a Phi + noise
    S = torch.empty((k_snap, N), dtype=torch.float64, device=dev)
    torch.matmul(torch.from_numpy(a).to(dev), Phi, out=S)
    gen = torch.Generator(device=dev).manual_seed(2)
    for c0 in range(0, k_snap, 16):
        c1 = min(k_snap, c0 + 16)
        S[c0:c1].add_(torch.randn((c1 - c0, N), generator=gen, dtype=torch.float64,
                                  device=dev), alpha=1e-6)
    del Phi

    # 3. M S: composed (BwdTrans then IProduct per snapshot, the reference's
    # MassMatrixOp_MatFree) and fused (Mass_Hex_Deformed, 64 snapshots per launch,
    # geometry read once per launch); the fused result feeds the Gram
    MS = torch.empty_like(S)
    ms_comp = _event_time(lambda: pod.mass_apply(S, proj.bwd, proj.ipr, out=MS,
                                                 phys=proj.phys), 1, 0)
    MS_comp_row = MS[k_snap // 2].clone()
    ms_fused = _event_time(lambda: pod.mass_apply_fused(S, proj.mass_op, out=MS), 2, 1)
    comp_bytes = k_snap * E * 8 * (2 * M + 3 * Q)
    fused_bytes = E * 8 * (2 * M * k_snap + Q * ((k_snap + 63) // 64))
    res["mass_apply"] = {
        "composed": {"ms_total": ms_comp, "algorithmic_GBs": comp_bytes / (ms_comp * 1e-3) / 1e9,
                     "bytes_per_snapshot_per_element": 8 * (2 * M + 3 * Q)},
        "fused": {"ms_total": ms_fused, "algorithmic_GBs": fused_bytes / (ms_fused * 1e-3) / 1e9,
                  "bytes_per_snapshot_per_element": 8 * 2 * M,
                  "geometry_bytes_per_launch_per_element": 8 * Q, "snapshots_per_launch": 64},
        "fused_vs_composed_rel_diff": (torch.linalg.vector_norm(MS[k_snap // 2] - MS_comp_row)
                                       / torch.linalg.vector_norm(MS_comp_row)).item()}
    del MS_comp_row

    # 4. Gram in three precisions
    work = torch.empty((pod._lib().irom_gram_workspace_size(k_snap, k_snap, N) + 7) // 8,
                       dtype=torch.float64, device=dev)
    G = {}
    flops = 2.0 * k_snap * k_snap * N          # every block of C = S (M S)^T
    flops_sym = 1.0 * k_snap * (k_snap + 1) * N  # upper triangle only (symmetric result)
    for mode, name in ((pod.GRAM_FP64, "fp64"), (pod.GRAM_FP32X3, "fp32x3"),
                       (pod.GRAM_BF16X3, "bf16x3")):
        out = torch.empty((k_snap, k_snap), dtype=torch.float64, device=dev)
        ms_full = _event_time(lambda: pod.gram(S, MS, mode=mode, out=out, work=work), 2, 1)
        ms = _event_time(lambda: pod.gram(S, MS, mode=mode, out=out, work=work,
                                          symmetric_result=True), 2, 1)
        G[name] = out
        res[f"gram_{name}"] = {"ms_symmetric": ms,
                               "TFLOPs_symmetric": flops_sym / (ms * 1e-3) / 1e12,
                               "flops_symmetric": flops_sym,
                               "ms_all_blocks": ms_full,
                               "TFLOPs_all_blocks": flops / (ms_full * 1e-3) / 1e12,
                               "flops_all_blocks": flops,
                               "HBM_floor_ms_reading_S_and_MS_once": 2 * k_snap * N * 8
                               / (HBM_PEAK_GBS * 1e9) * 1e3}
        res[f"gram_{name}"]["frac_of_hbm_floor_symmetric"] = (
            res[f"gram_{name}"]["HBM_floor_ms_reading_S_and_MS_once"] / ms)
    gn = torch.linalg.matrix_norm(G["fp64"]).item()
    for name in ("fp32x3", "bf16x3"):
        res[f"gram_{name}"]["rel_err_vs_gpu_fp64"] = (
            torch.linalg.matrix_norm(G[name] - G["fp64"]).item() / gn)
    res["gram_fp64"]["mfma_frac_of_78.6TF_all_blocks"] = (res["gram_fp64"]["TFLOPs_all_blocks"]
                                                          / 78.6)
    res["gram_fp64"]["mfma_frac_of_78.6TF_symmetric"] = (res["gram_fp64"]["TFLOPs_symmetric"]
                                                         / 78.6)
    if not args.no_cpu:
        # FP64 CPU oracle over the full K = N: the 32 x 32 principal submatrix of the Gram
        # on every 16th snapshot, numpy dot products of the full-length rows
        sub = torch.arange(0, k_snap, 16, device=dev)
        Sh = S[sub].cpu().numpy()
        MSh = MS[sub].cpu().numpy()
        from threadpoolctl import threadpool_limits
        with threadpool_limits(limits=min(16, os.cpu_count() or 1)):
            t0 = time.perf_counter()
            Gsub = Sh @ MSh.T
            t_sub = time.perf_counter() - t0
        gsn = np.linalg.norm(Gsub)
        for name in ("fp64", "fp32x3", "bf16x3"):
            Gg = G[name][sub][:, sub].cpu().numpy()
            res[f"gram_{name}"]["rel_err_vs_cpu_fp64_submatrix"] = float(
                np.linalg.norm(Gg - Gsub) / gsn)
        res["gram_cpu_fp64_oracle"] = {
            "sample": f"{len(sub)} x {len(sub)} principal submatrix (every 16th snapshot), "
                      f"full K = {N}, numpy FP64", "s": t_sub}
        del Sh, MSh
    res["gram_fp64"]["rocblas_dgemm_ms"] = _event_time(lambda: torch.matmul(S, MS.T), 1, 1)

    # 5. eigendecomposition (FP64 Jacobi on the GPU), truncation, basis
    Gs = 0.5 * (G["fp64"] + G["fp64"].T)
    ev, Psi, sweeps = pod.eig_sym(Gs)  # first call: allocates the kept workspace
    te = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev, Psi, sweeps = pod.eig_sym(Gs)
        te.append((time.perf_counter() - t0) * 1e3)
    res["eig_ms"] = float(np.median(te))
    res["eig_sweeps"] = sweeps
    res["eig_method"] = "blocked one-sided Jacobi, one persistent launch (median of 3)"
    r = pod.energy_truncation(ev, 1.0 - 1e-8)
    res["rank_energy_1e-8"] = r
    C = Psi[:r] / torch.sqrt(torch.as_tensor(ev[:r], device=dev))[:, None]
    for name in ("fp32x3", "bf16x3"):
        ev_m, Psi_m, _ = pod.eig_sym(0.5 * (G[name] + G[name].T))
        C_m = Psi_m[:r] / torch.sqrt(torch.as_tensor(ev_m[:r], device=dev))[:, None]
        sgn = torch.sign((C_m * C).sum(1, keepdim=True))
        Dm = C_m * sgn - C
        # ||V_m,i - V_i||_M = sqrt(d_i G d_i^T), ||V_i||_M = 1
        verr = torch.sqrt(torch.clamp(((Dm @ Gs) * Dm).sum(1), min=0.0)).max().item()
        res[f"gram_{name}"]["eig_rel_err_top_r"] = float(
            np.max(np.abs(ev_m[:r] - ev[:r]) / ev[:r]))
        res[f"gram_{name}"]["basis_max_M_norm_err"] = verr
    V = C @ S
    cpu_slice = None
    if not args.no_cpu:
        Ks = 1024 * M
        cpu_slice = (S[:, :Ks].cpu().numpy(), MS[:, :Ks].cpu().numpy(),
                     pod.gram(S[:, :Ks], MS[:, :Ks], work=work).cpu().numpy())
    del MS
    MV = torch.empty_like(V)
    pod.mass_apply_fused(V, proj.mass_op, out=MV)
    O = pod.gram(V, MV)
    res["basis_M_orthonormality_err"] = (O - torch.eye(r, dtype=torch.float64,
                                                       device=dev)).abs().max().item()
    # 6. Galerkin projection of the affine terms of the Helmholtz operator, one at a
    # time (stiffness K = Helmholtz with lambda = 0, mass M): C_q = V A_q V^T, A_q applied
    # to all r basis vectors in one launch, then the product on MFMA; A(10) = K + 10 M
    terms = pod.helmholtz_affine_terms(nm, nq, E, b, jac, df)
    AV = MV
    Cq, msq = {}, {}
    for name, opq in terms.items():
        msq[name] = _event_time(lambda: pod.galerkin_project(opq, V, W=AV), 2, 1)
        Cq[name] = {m: pod.galerkin_project(opq, V, mode=mode, W=AV)[0]
                    for m, mode in (("fp64", pod.GRAM_FP64), ("fp32x3", pod.GRAM_FP32X3),
                                    ("bf16x3", pod.GRAM_BF16X3))}
        if not args.no_cpu:
            # FP64 CPU oracle of the product: V and A_q V (FP64 GPU apply, itself pinned to
            # the oracle to 1e-12) to the host, numpy V (A_q V)^T over the full N
            Vh, AVh = V.cpu().numpy(), AV.cpu().numpy()
            Cq[name]["cpu"] = Vh @ AVh.T
            del Vh, AVh
    Hop = H.Operator("Helmholtz_Hex_Deformed", nm, nq, E, basis=b)
    Hop.set_jac(jac)
    Hop.set_df(df)
    Hop.set_lambda(w["lam"])
    ms_proj = _event_time(lambda: pod.galerkin_project(Hop, V, W=AV), 2, 1)
    Ar, _ = pod.galerkin_project(Hop, V, W=AV)
    comb = w["lam"] * Cq["mass"]["fp64"] + Cq["stiffness"]["fp64"]
    gp = {"r": r, "ms_helmholtz_lambda10": ms_proj,
          "ms_per_term": msq,
          "term_recombination_rel_err": (torch.linalg.matrix_norm(comb - Ar)
                                         / torch.linalg.matrix_norm(Ar)).item(),
          "symmetry_rel_err": (torch.linalg.matrix_norm(Ar - Ar.T)
                               / torch.linalg.matrix_norm(Ar)).item(),
          "bytes_per_launch_apply": E * 8 * (2 * M * r + 7 * Q),
          "entry": "irom_galerkin_project (multi-vector apply + MFMA product)"}
    for name in terms:
        ref = Cq[name].get("cpu")
        if ref is None:
            continue
        for m in ("fp64", "fp32x3", "bf16x3"):
            gp[f"{name}_{m}_rel_err_vs_cpu_fp64"] = float(
                np.linalg.norm(Cq[name][m].cpu().numpy() - ref) / np.linalg.norm(ref))
    res["galerkin_projection"] = gp
    HV = AV

    if not args.no_cpu:
        res["cpu_baseline"] = pod_cpu_baseline(cpu_slice, G["fp64"], ev, r)
    del S, V, MV, HV, Hop, proj, work, G, terms, Cq
    torch.cuda.empty_cache()
    return res


def pod_cpu_baseline(cpu_slice, G64, ev, r):
    """FP64 CPU oracle of the reduced quantities within a bounded CPU budget: the numpy
    (OpenBLAS) Gram of a K-slice of S and M S (the first 1024 elements' coefficients; M is
    block diagonal, so the slice of M S is exact), compared with the GPU FP64 Gram kernel on
    the same slice, and numpy eigh of the full GPU Gram compared with the GPU eigenvalues."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    S_h, MS_h, G_gpu = cpu_slice
    threads = min(16, os.cpu_count() or 1)
    with threadpool_limits(limits=threads):
        S_h @ MS_h[:8].T  # warm-up
        t0 = time.perf_counter()
        G_cpu = S_h @ MS_h.T
        t_gram = time.perf_counter() - t0
        Gh = G64.cpu().numpy()
        t0 = time.perf_counter()
        ev_cpu = np.linalg.eigh(0.5 * (Gh + Gh.T))[0][::-1]
        t_eig = time.perf_counter() - t0
    flops = 2.0 * S_h.shape[0] ** 2 * S_h.shape[1]
    return {"value": flops / t_gram / 1e9, "unit": "GFLOP/s (FP64 Gram)", "cores": threads,
            "kind": "port",
            "sample": f"Gram of a {S_h.shape[0]} x {S_h.shape[1]} K-slice (numpy/OpenBLAS), "
                      "eigh of the full 512^2 Gram (LAPACK)",
            "gram_slice_s": t_gram, "eigh_ms": t_eig * 1e3,
            "gram_slice_rel_err_gpu_vs_cpu": float(np.linalg.norm(G_gpu - G_cpu)
                                                   / np.linalg.norm(G_cpu)),
            "eig_rel_err_top_r_gpu_vs_cpu": float(np.max(np.abs(ev[:r] - ev_cpu[:r])
                                                         / ev_cpu[:r]))}

def regular_bench(args):
    """Helmholtz_Hex_Regular on 48^3 undistorted hexes of [0,1]^3 (P=5, 7 GLL points,
    lambda = 10, FP64, 50 applies): the per-element constant geometry (J and the 6 metric
    terms, 7 doubles per element) replaces the 7 doubles per quadrature point of the
    deformed path, so the apply streams 8*(2M + 7) bytes per element; CPU oracle on an
    8192-element sample."""
    import numpy as np
    import torch

    import ithaca_hip as H
    from ithaca_hip import mesh

    w = mesh.workload_hex_helmholtz(args.n)
    nm, nq, E, lam = w["nm"], w["nq"], w["nElmt"], w["lam"]
    b = H.basis_modified_a(nm, nq)
    verts = mesh.hex_box(args.n, 0.0)
    gj, gd = H.geom_device("hex", torch.from_numpy(verts).cuda(), nq, b["z"], False)
    op = H.Operator("Helmholtz_Hex_Regular", nm, nq, E, basis=b)
    op.set_jac(gj)
    op.set_df(gd)
    op.set_lambda(lam)
    x = torch.from_numpy(w["u"]).cuda()
    y = torch.empty(op.out_size, dtype=torch.float64, device="cuda")
    ms = _event_time(lambda: op.apply(x, y), 50, 5)
    ab = 8 * (2 * nm ** 3 + 7) * E
    res = {"key": "Helmholtz_Hex_Regular", "workload": f"{E} regular hexes, P=5, 7 GLL pts, "
                                                       "lambda=10",
           "us_per_apply": ms * 1e3, "gdof_per_s": E * nm ** 3 / (ms * 1e-3) / 1e9,
           "algorithmic_bytes": ab, "achieved_GBs": ab / (ms * 1e-3) / 1e9,
           "speedup_over_deformed_path": None}
    del op, x, y, gj, gd
    if not args.no_cpu:
        import oracle as O
        threads = min(16, os.cpu_count() or 1)
        O.set_threads(threads)
        Es = 8192
        bo = O.basis(nm, nq)
        jac, df = O.geom_hex(np.ascontiguousarray(verts[:Es]), nq, bo["z"], False)
        u = np.ascontiguousarray(w["u"][:Es * nm ** 3])
        O.helmholtz("hex", bo, Es, False, jac, df, lam, u)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            O.helmholtz("hex", bo, Es, False, jac, df, lam, u)
            ts.append(time.perf_counter() - t0)
        res["cpu_baseline"] = {"gdof_per_s": Es * nm ** 3 / float(np.median(ts)) / 1e9,
                               "cores": threads, "kind": "port",
                               "sample": f"{Es} elements, 3 applies, median; oracle "
                                         "restatement of HelmholtzHexImpl (regular)"}
    torch.cuda.empty_cache()
    return res


def deriv_bench(args):
    """PhysDeriv_Hex_Deformed and IProductWRTDerivBase_Hex_Deformed on the 48^3 mesh
    (FP64, 50 applies each), with their algorithmic bytes per apply and the CPU oracle
    on a 4096-element sample."""
    import numpy as np
    import torch

    import ithaca_hip as H
    from ithaca_hip import mesh

    w = mesh.workload_hex_helmholtz(args.n)
    nm, nq, E = w["nm"], w["nq"], w["nElmt"]
    Q = nq ** 3
    b = H.basis_modified_a(nm, nq)
    verts = torch.from_numpy(w["verts"]).cuda()
    gj, gd = H.geom_device("hex", verts, nq, b["z"], True)
    res = {}
    for name, key in (("physderiv", "PhysDeriv_Hex_Deformed"),
                      ("iprodderiv", "IProductWRTDerivBase_Hex_Deformed")):
        op = H.Operator(key, nm, nq, E, basis=b)
        if op.needs_jac():
            op.set_jac(gj)
        op.set_df(gd)
        x = torch.rand(op.in_size, dtype=torch.float64, device="cuda").mul_(2).sub_(1)
        y = torch.empty(op.out_size, dtype=torch.float64, device="cuda")
        ms = _event_time(lambda: op.apply(x, y), 50, 5)
        ab = algorithmic_bytes(name, nm, nq, 8) * E
        res[name] = {"key": key, "us_per_apply": ms * 1e3, "algorithmic_bytes": ab,
                     "achieved_GBs": ab / (ms * 1e-3) / 1e9,
                     "frac_of_8TBs": ab / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        del op, x, y
    if not args.no_cpu:
        import oracle as O
        threads = min(16, os.cpu_count() or 1)
        O.set_threads(threads)
        Es = 4096
        bo = O.basis(nm, nq)
        jac, df = O.geom_hex(np.ascontiguousarray(w["verts"][:Es]), nq, bo["z"], True)
        f = np.random.default_rng(0).uniform(-1, 1, 3 * Es * Q)
        for name, fn in (("physderiv", lambda: O.physderiv("hex", bo, Es, True, df, f[:Es * Q])),
                         ("iprodderiv", lambda: O.iprod_deriv("hex", bo, Es, True, jac, df,
                                                              np.split(f, 3)))):
            fn()
            t0 = time.perf_counter()
            for _ in range(5):
                fn()
            t = (time.perf_counter() - t0) / 5
            res[name]["cpu_baseline"] = {"us_per_apply_scaled_to_full_mesh": t * E / Es * 1e6,
                                         "cores": threads, "kind": "port",
                                         "sample": f"{Es} elements, 5 applies, oracle"}
    del gj, gd
    torch.cuda.empty_cache()
    return res


def helmsolve_bench(args):
    """C0-assembled Helmholtz solve on the 48^3 jittered mesh (P=5, 7 GLL points, lambda
    = 10): -lap u + lambda u = f on [0,1]^3, u = 0 on the boundary, f from
    u = sin(pi x) sin(pi y) sin(pi z); matrix-free conjugate gradients with the assembled
    diagonal, stopping at the reference's default tolerance 1e-9."""
    import numpy as np
    import torch

    import ithaca_hip as H
    from ithaca_hip import assembly, mesh

    w = mesh.workload_hex_helmholtz(args.n)
    nm, nq, E = w["nm"], w["nq"], w["nElmt"]
    b = H.basis_modified_a(nm, nq)
    vd = torch.from_numpy(w["verts"]).cuda()
    jac, df = H.geom_device("hex", vd, nq, b["z"], True)
    Hop = H.Operator("Helmholtz_Hex_Deformed", nm, nq, E, basis=b)
    Hop.set_jac(jac)
    Hop.set_df(df)
    Hop.set_lambda(w["lam"])
    ipr = H.Operator("IProduct_Hex_Deformed", nm, nq, E, basis=b)
    ipr.set_jac(jac)
    bwd = H.Operator("BwdTrans_Hex_Regular", nm, nq, E, basis=b)
    X = mesh.hex_quad_points(vd, b["z"]).reshape(-1, 3)
    uex = torch.sin(np.pi * X[:, 0]) * torch.sin(np.pi * X[:, 1]) * torch.sin(np.pi * X[:, 2])
    del X
    A = assembly.Assembly(mesh.hex_box_ids(args.n), nm, dirichlet=True)
    rhs = A.assemble(ipr.apply(((3 * np.pi ** 2 + w["lam"]) * uex).contiguous()))
    x = torch.empty_like(rhs)
    A.helmsolve(Hop, rhs, x=x, tol=1e-9, max_iter=20000)  # warm-up (builds the diagonal)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, it, res = A.helmsolve(Hop, rhs, x=x, tol=1e-9, max_iter=20000)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    uh = bwd.apply(A.global_to_local(x))
    wq = np.einsum("k,j,i->kji", b["w"], b["w"], b["w"]).ravel()
    wj = jac * torch.from_numpy(np.tile(wq, E)).cuda()
    err = torch.sqrt(torch.sum(wj * (uh - uex) ** 2) / torch.sum(wj * uex ** 2)).item()
    res_d = {"workload": f"Dirichlet Helmholtz solve, {E} deformed hexes, P=5, lambda=10, "
                         "manufactured u = sin(pi x) sin(pi y) sin(pi z)",
             "global_modes": A.nglobal, "cg_iterations": it, "relative_residual": res,
             "solve_s": t, "ms_per_iteration": t / max(it, 1) * 1e3,
             "local_gdof_per_s_per_iteration": E * nm ** 3 / (t / max(it, 1)) / 1e9,
             "l2_error_vs_exact": err, "tolerance": 1e-9}
    del A, Hop, ipr, bwd, x, rhs, uh, uex, jac, df, wj
    torch.cuda.empty_cache()
    return res_d


def varp_helmsolve_bench(args):
    """The solve of helmsolve_bench on the same 48^3 jittered mesh with two orders: P=4
    (5 modes, 6 GLL points) in the lower half of the box, P=6 (7 modes, 8 points) in the
    upper half; one Helmholtz operator per half, the variable-order C0 numbering between
    them."""
    import numpy as np
    import torch

    import ithaca_hip as H
    from ithaca_hip import assembly, mesh

    w = mesh.workload_hex_helmholtz(args.n)
    E, lam = w["nElmt"], w["lam"]
    half = E // 2  # elements are z-major: the first half is the lower half of the box
    groups = [(0, half, 5), (half, E, 7)]
    nmv = np.concatenate([np.full(e1 - e0, nm, dtype=np.int32) for e0, e1, nm in groups])
    A = assembly.Assembly(mesh.hex_box_ids(args.n), nmv, dirichlet=True)
    ops, loc, parts = [], [], []
    for e0, e1, nm in groups:
        nq = nm + 1
        b = H.basis_modified_a(nm, nq)
        vd = torch.from_numpy(w["verts"][e0:e1]).cuda()
        jac, df = H.geom_device("hex", vd, nq, b["z"], True)
        Hop = H.Operator("Helmholtz_Hex_Deformed", nm, nq, e1 - e0, basis=b)
        Hop.set_jac(jac)
        Hop.set_df(df)
        Hop.set_lambda(lam)
        ipr = H.Operator("IProduct_Hex_Deformed", nm, nq, e1 - e0, basis=b)
        ipr.set_jac(jac)
        X = mesh.hex_quad_points(vd, b["z"]).reshape(-1, 3)
        uex = torch.sin(np.pi * X[:, 0]) * torch.sin(np.pi * X[:, 1]) * torch.sin(np.pi * X[:, 2])
        del X, df
        loc.append(ipr.apply(((3 * np.pi ** 2 + lam) * uex).contiguous()))
        wq = np.einsum("k,j,i->kji", b["w"], b["w"], b["w"]).ravel()
        parts.append((H.Operator("BwdTrans_Hex_Regular", nm, nq, e1 - e0, basis=b), uex,
                      jac * torch.from_numpy(np.tile(wq, e1 - e0)).cuda()))
        ops.append(Hop)
        del ipr
    rhs = A.assemble(torch.cat(loc))
    del loc
    x = torch.empty_like(rhs)
    A.helmsolve(ops, rhs, x=x, tol=1e-9, max_iter=20000)  # warm-up (builds the diagonal)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, it, res = A.helmsolve(ops, rhs, x=x, tol=1e-9, max_iter=20000)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    xl = A.global_to_local(x)
    num = den = 0.0
    o = 0
    for bwd, uex, wj in parts:
        n = int(bwd.in_size)
        uh = bwd.apply(xl[o:o + n].contiguous())
        num += torch.sum(wj * (uh - uex) ** 2).item()
        den += torch.sum(wj * uex ** 2).item()
        o += n
    res_d = {"workload": f"Dirichlet Helmholtz solve, {E} deformed hexes, P=4 in the lower "
                         "half and P=6 in the upper half (variable order), lambda=10, "
                         "manufactured u = sin(pi x) sin(pi y) sin(pi z)",
             "operators": len(ops), "local_modes": A.nlocal, "global_modes": A.nglobal,
             "cg_iterations": it, "relative_residual": res, "solve_s": t,
             "ms_per_iteration": t / max(it, 1) * 1e3,
             "l2_error_vs_exact": float(np.sqrt(num / den)), "tolerance": 1e-9}
    del A, ops, parts, x, xl, rhs
    torch.cuda.empty_cache()
    return res_d


def quad_helmsolve_bench(args):
    """C0-assembled Helmholtz solve on the 32x32 regular quads of the quad workload (P=4,
    6 GLL points, lambda = 1): -lap u + u = f on [0,1]^2, u = 0 on the boundary, f from
    u = sin(pi x) sin(pi y); matrix-free conjugate gradients with the assembled diagonal,
    stopping at the reference's default tolerance 1e-9."""
    import numpy as np
    import torch

    import ithaca_hip as H
    from ithaca_hip import assembly, mesh

    w = mesh.workload_quad_helmholtz()
    nm, nq, E, lam = w["nm"], w["nq"], w["nElmt"], w["lam"]
    n = int(round(E ** 0.5))
    b = H.basis_modified_a(nm, nq)
    vd = torch.from_numpy(w["verts"]).cuda()
    jac, df = H.geom_device("quad", vd, nq, b["z"], False)
    Hop = H.Operator("Helmholtz_Quad_Regular", nm, nq, E, basis=b)
    Hop.set_jac(jac)
    Hop.set_df(df)
    Hop.set_lambda(lam)
    ipr = H.Operator("IProduct_Quad_Regular", nm, nq, E, basis=b)
    ipr.set_jac(jac)
    bwd = H.Operator("BwdTrans_Quad_Regular", nm, nq, E, basis=b)
    X = mesh.quad_points(vd, b["z"]).reshape(-1, 2)
    uex = torch.sin(np.pi * X[:, 0]) * torch.sin(np.pi * X[:, 1])
    A = assembly.Assembly(mesh.quad_box_ids(n), nm, dirichlet=True, shape="quad")
    rhs = A.assemble(ipr.apply(((2 * np.pi ** 2 + lam) * uex).contiguous()))
    x = torch.empty_like(rhs)
    A.helmsolve(Hop, rhs, x=x, tol=1e-9, max_iter=20000)  # warm-up (builds the diagonal)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, it, res = A.helmsolve(Hop, rhs, x=x, tol=1e-9, max_iter=20000)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    uh = bwd.apply(A.global_to_local(x))
    wq = np.einsum("j,i->ji", b["w"], b["w"]).ravel()
    wj = jac.repeat_interleave(nq * nq) * torch.from_numpy(np.tile(wq, E)).cuda()
    err = torch.sqrt(torch.sum(wj * (uh - uex) ** 2) / torch.sum(wj * uex ** 2)).item()
    return {"workload": f"Dirichlet Helmholtz solve, {E} regular quads, P=4, lambda=1, "
                        "manufactured u = sin(pi x) sin(pi y)",
            "global_modes": A.nglobal, "cg_iterations": it, "relative_residual": res,
            "solve_s": t, "ms_per_iteration": t / max(it, 1) * 1e3,
            "l2_error_vs_exact": err, "tolerance": 1e-9}


def quad_bench(args):
    """32x32 regular quads, P=4 (5 modes), 6 GLL points, lambda=1, seed 0, 100 applies."""
    import numpy as np
    import torch

    import ithaca_hip as H
    from ithaca_hip import mesh

    w = mesh.workload_quad_helmholtz()
    nm, nq, E = w["nm"], w["nq"], w["nElmt"]
    b = H.basis_modified_a(nm, nq)
    verts = torch.from_numpy(w["verts"]).cuda()
    gj, gd = H.geom_device("quad", verts, nq, b["z"], False)
    op = H.Operator("Helmholtz_Quad_Regular", nm, nq, E, basis=b)
    op.set_jac(gj)
    op.set_df(gd)
    op.set_lambda(w["lam"])
    x = torch.from_numpy(w["u"]).cuda()
    y = torch.empty(op.out_size, dtype=torch.float64, device="cuda")
    ms = _event_time(lambda: op.apply(x, y), 100, 10)
    res = {"workload": "Helmholtz_Quad_Regular 32x32 quads, P=4, 6 GLL pts, lambda=1",
           "gpu_us_per_apply": ms * 1e3, "ndof": E * nm * nm}
    if not args.no_cpu:
        import oracle as O
        O.set_threads(1)
        bo = O.basis(nm, nq)
        jac, df = O.geom_quad(w["verts"], nq, bo["z"], False)
        O.helmholtz("quad", bo, E, False, jac, df, w["lam"], w["u"])
        t0 = time.perf_counter()
        for _ in range(100):
            O.helmholtz("quad", bo, E, False, jac, df, w["lam"], w["u"])
        res["cpu_baseline"] = {"us_per_apply": (time.perf_counter() - t0) / 100 * 1e6,
                               "cores": 1, "kind": "port",
                               "sample": "100 applies, oracle restatement of HelmholtzQuadImpl"}
        O.set_threads(min(16, os.cpu_count() or 1))
    return res


def copy_calibration(args):
    """Device-to-device copy of 1 GiB: reads 1 GiB, writes 1 GiB per launch."""
    import torch

    n = (1 << 30) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(args.steps):
        y.copy_(x)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    gbs = 2 * n * 8 / (ms * 1e-3) / 1e9
    print(json.dumps({"metric": "device copy bandwidth", "value": gbs, "unit": "GB/s",
                      "bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8,
                      "ms_per_step": ms, "steps": args.steps, "frac_of_8TBs": gbs / HBM_PEAK_GBS}))


OP_KEYS = {"helmholtz": "Helmholtz_Hex_Deformed", "bwdtrans": "BwdTrans_Hex_Regular",
           "iproduct": "IProduct_Hex_Deformed", "physderiv": "PhysDeriv_Hex_Deformed",
           "iprodderiv": "IProductWRTDerivBase_Hex_Deformed"}


def time_operator(w, opname, precision, steps, warmup, keep=False):
    """Mean time of one apply of the hex operator `opname` over the workload w, timed with
    HIP events recorded on the stream the kernel is launched on; algorithmic bytes and the
    fraction of the 8 TB/s HBM peak."""
    import torch

    import ithaca_hip as H

    dev = torch.device("cuda:0")
    nm, nq, E = w["nm"], w["nq"], w["nElmt"]
    prec = H.FP64 if precision == "fp64" else H.FP32
    dtype = torch.float64 if prec == H.FP64 else torch.float32
    b = H.basis_modified_a(nm, nq)
    verts = torch.from_numpy(w["verts"]).to(dev)
    key = OP_KEYS[opname]
    op = H.Operator(key, nm, nq, E, precision=prec, basis=b)
    if op.needs_jac() or op.needs_df():
        gj, gd = H.geom_device("hex", verts, nq, b["z"], True)
        if op.needs_jac():
            op.set_jac(gj)
        if op.needs_df():
            op.set_df(gd)
        del gj, gd
    op.set_lambda(w["lam"])
    if opname in ("iproduct", "physderiv", "iprodderiv"):
        x = torch.rand(op.in_size, dtype=torch.float64, device=dev).mul_(2).sub_(1).to(dtype)
    else:
        x = torch.from_numpy(w["u"]).to(dev, dtype)
    y = torch.empty(op.out_size, dtype=dtype, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        op.apply(x, y, stream)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        op.apply(x, y, stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / steps
    elt = 8 if prec == H.FP64 else 4
    abytes = algorithmic_bytes(opname, nm, nq, elt) * E
    achieved = abytes / (ms * 1e-3) / 1e9
    res = {"key": key, "precision": precision, "ms": ms, "wall": wall,
           "gdof_per_s": E * nm ** 3 / (ms * 1e-3) / 1e9, "achieved": achieved,
           "abytes": abytes, "frac_of_8TBs": achieved / HBM_PEAK_GBS,
           "algorithmic_bytes_per_element": algorithmic_bytes(opname, nm, nq, elt)}
    if keep:
        res.update(_op=op, _x=x, _y=y)
    return res


def hex_operators_bench(args):
    """The other timings of the 48^3 hex workload: Helmholtz in FP32, and the standalone
    BwdTrans and IProductWRTBase (f = BwdTrans(u_hat)-sized random input) in FP64 and FP32,
    20 applies each."""
    from ithaca_hip import mesh

    w = mesh.workload_hex_helmholtz(args.n)
    out = {}
    for opname, prec in (("helmholtz", "fp32"), ("bwdtrans", "fp64"), ("iproduct", "fp64"),
                         ("bwdtrans", "fp32"), ("iproduct", "fp32")):
        r = time_operator(w, opname, prec, 20, 3)
        out[f"{opname}_{prec}"] = {k: r[k] for k in ("key", "ms", "gdof_per_s", "frac_of_8TBs",
                                                      "algorithmic_bytes_per_element")}
    return out


def main():
    args = parse()
    if args.op == "copy":
        return copy_calibration(args)
    if args.rom_only:
        print(json.dumps({"batched_reduced_solves": rom_bench(args)}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: --gpus must be 1 (the path's sharding "
                                   "over several GPUs is not implemented)"}))
        sys.exit(2)
    import numpy as np
    import torch

    import ithaca_hip as H
    from ithaca_hip import mesh

    w = mesh.workload_hex_helmholtz(args.n)
    nm, nq, E = w["nm"], w["nq"], w["nElmt"]
    t = time_operator(w, args.op, args.precision, args.steps, args.warmup, keep=True)
    key, ms, wall, achieved, abytes = t["key"], t["ms"], t["wall"], t["achieved"], t["abytes"]
    op, x, y = t.pop("_op"), t.pop("_x"), t.pop("_y")
    prec = H.FP64 if args.precision == "fp64" else H.FP32
    ndof = E * nm ** 3

    traffic = None
    tpath = os.path.join(REPO, "profiles", f"traffic_{args.op}_hex_{args.precision}.json")
    if os.path.exists(tpath) and args.n == 48:
        with open(tpath) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    out = {
        "metric": "Helmholtz apply GDOF/s (%HBM peak); batched reduced solves/s at 1e-10 vs FP64",
        "value": ndof / (ms * 1e-3) / 1e9,
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == H.FP64 else "f32",
        "data": "synthetic (jittered 48^3 hex mesh, seed 1; coefficients U[-1,1])",
        "config": {"workload": f"{key} {args.n}^3={E} hexes, P=5 (6 modes), 7 GLL pts, "
                               f"lambda=10", "elements": E, "modes": nm, "points": nq,
                   "ndof_per_apply": ndof, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": abytes},
        "wall_s_timed_region": wall,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        out["gpu_over_cpu_avx2"] = out["value"] / out["cpu_baseline"]["avx2_value"]
    if args.op == "helmholtz" and args.precision == "fp64" and args.n == 48 and not args.headline_only:
        # the host-buffer entry (imf_apply: pageable host arrays copied in and out around
        # the same kernel, geometry resident); reported beside value, never as it
        uh = x.cpu().numpy()
        op.apply_host(uh)
        th = []
        for _ in range(3):
            t0 = time.perf_counter()
            op.apply_host(uh)
            th.append(time.perf_counter() - t0)
        th = float(np.median(th))
        out["host_buffer_path"] = {
            "entry": "imf_apply (host in/out, FP64)", "ms_per_apply": th * 1e3,
            "gdof_per_s": ndof / th / 1e9, "bytes_over_pcie": 2 * ndof * 8,
            "sample": "median of 3 applies on the same workload"}
        del uh
        del op, x, y
        torch.cuda.empty_cache()
        def section(name, fn):
            t0 = time.perf_counter()
            out[name] = fn(args)
            print(f"[bench] {name}: {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

        print(f"[bench] headline: {out['value']:.2f} GDOF/s", file=sys.stderr, flush=True)
        if not args.no_rom:
            section("batched_reduced_solves", rom_bench)
        section("hex_operators", hex_operators_bench)
        section("quad_helmholtz", quad_bench)
        section("quad_helmsolve", quad_helmsolve_bench)
        section("deriv_operators", deriv_bench)
        section("regular_hex", regular_bench)
        out["regular_hex"]["speedup_over_deformed_path"] = (
            out["regular_hex"]["gdof_per_s"] / out["value"])
        section("helmsolve", helmsolve_bench)
        section("helmsolve_varp", varp_helmsolve_bench)
        if not args.no_pod:
            try:
                section("pod_offline", pod_bench)
            except Exception as exc:  # noqa: BLE001 -- reported, the main line stands
                out["pod_offline"] = {"error": f"{type(exc).__name__}: {exc}"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
