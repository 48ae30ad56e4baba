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
    p.add_argument("--op", choices=["helmholtz", "bwdtrans", "iproduct", "copy"],
                   default="helmholtz",
                   help="copy: plain 1 GiB device copy (HBM bandwidth / counter calibration)")
    p.add_argument("--n", type=int, default=48, help="elements per direction")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    return p.parse_args()


def algorithmic_bytes(op, nm, nq, elt):
    M, Q = nm ** 3, nq ** 3
    if op == "helmholtz":
        return elt * (2 * M + 7 * Q)      # u_hat + out + 7 geometry values per point
    if op == "bwdtrans":
        return elt * (M + Q)              # u_hat + u
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
    return {"value": E * nm ** 3 / t / 1e9, "unit": "GDOF/s", "cores": threads,
            "kind": "port",
            "sample": f"{E} of {w['nElmt']} elements, {len(times)} applies, median; "
                      "oracle restatement of HelmholtzHexImpl (scalar element loop, "
                      f"width 1), OpenMP over contiguous element chunks, {threads} threads"}


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


def main():
    args = parse()
    if args.op == "copy":
        return copy_calibration(args)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not run)"}))
        sys.exit(2)
    import numpy as np
    import torch

    import ithaca_hip as H
    from ithaca_hip import mesh

    dev = torch.device("cuda:0")
    w = mesh.workload_hex_helmholtz(args.n)
    nm, nq, E = w["nm"], w["nq"], w["nElmt"]
    prec = H.FP64 if args.precision == "fp64" else H.FP32
    dtype = torch.float64 if prec == H.FP64 else torch.float32
    b = H.basis_modified_a(nm, nq)
    verts = torch.from_numpy(w["verts"]).to(dev)
    key = {"helmholtz": "Helmholtz_Hex_Deformed", "bwdtrans": "BwdTrans_Hex_Regular",
           "iproduct": "IProduct_Hex_Deformed"}[args.op]
    op = H.Operator(key, nm, nq, E, precision=prec, basis=b)
    if op.needs_jac():
        gj, gd = H.geom_device("hex", verts, nq, b["z"], True)
        op.set_jac(gj)
        if op.needs_df():
            op.set_df(gd)
        del gj, gd
    op.set_lambda(w["lam"])
    if args.op == "iproduct":
        x = torch.rand(op.in_size, dtype=torch.float64, device=dev).mul_(2).sub_(1).to(dtype)
    else:
        x = torch.from_numpy(w["u"]).to(dev, dtype)
    y = torch.empty(op.out_size, dtype=dtype, device=dev)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        op.apply(x, y, stream)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        op.apply(x, y, stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    ndof = E * nm ** 3
    elt = 8 if prec == H.FP64 else 4
    abytes = algorithmic_bytes(args.op, nm, nq, elt) * E
    achieved = abytes / (ms * 1e-3) / 1e9

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
    print(json.dumps(out))


if __name__ == "__main__":
    main()
