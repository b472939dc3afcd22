"""Benchmark of the ND block-tridiagonal Cholesky on one MI355X.

Workload: n=128, N=512 block-tridiagonal SPD systems from the README
generator (diagonal shift alpha), m right-hand sides, one system (B=1).
A step is one factor_and_solve of the whole system with inputs resident in
HBM; the in-place inputs are restored from pristine device copies before each
step, outside the HIP-event-timed region.

Prints one JSON line (the driver contract) with a live roofline figure for the
dominant kernel (HIP events around its launches) and a CPU baseline (the
oracle's C restatement on the host cores).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode fp64]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--N", type=int, default=512)
    ap.add_argument("--m", type=int, default=1)
    ap.add_argument("--alpha", type=float, default=16.0)
    ap.add_argument("--mode", default="fp64", choices=["fp64"])
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def cpu_baseline(L0, E0, b0, seconds):
    """Time the oracle's C restatement (OpenMP over the nodes of a level) on
    the same problem, repeated until `seconds` of CPU time have elapsed."""
    from oracle import nd_c

    threads = nd_c.set_threads(os.cpu_count() or 1)
    reps, t_total = 0, 0.0
    while t_total < seconds or reps < 2:
        L = L0[None].copy()
        E = E0[None].copy()
        x = b0[None].copy()
        t0 = time.perf_counter()
        nd_c.factor_and_solve(L, E, x)
        t_total += time.perf_counter() - t0
        reps += 1
        if reps >= 200:
            break
    return reps / t_total, threads, reps


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(independent systems shard over GPUs with no communication; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import torch

    from oracle import nd
    from socu_amd import _lib
    from socu_amd.block_tridiag_solver import create_cholesky_factor_and_solve_launch

    n, N, m = args.n, args.N, args.m
    dev = torch.device("cuda", 0)
    log(f"bench: generating n={n} N={N} m={m} alpha={args.alpha}")
    L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=args.alpha, seed=0)
    Lp = torch.tensor(L0, device=dev)
    Ep = torch.tensor(E0, device=dev)
    bp = torch.tensor(b0, device=dev)
    L, E, x = Lp.clone(), Ep.clone(), bp.clone()
    st = torch.cuda.current_stream(dev)

    def restore():
        L.copy_(Lp)
        E.copy_(Ep)
        x.copy_(bp)

    fs = create_cholesky_factor_and_solve_launch(L, E, x, use_cuda_graph=True)
    log("bench: warmup")
    for _ in range(args.warmup):
        restore()
        fs()
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        restore()
        starts[k].record(st)
        fs()
        ends[k].record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])
    ms_per_step = float(per.mean())
    log(f"bench: {args.steps} steps, mean {ms_per_step:.3f} ms p50 {np.median(per):.3f} "
        f"p99 {np.percentile(per, 99):.3f} (wall incl. restores {wall * 1e3 / args.steps:.3f} ms/step)")

    xg = x.double().cpu().numpy()
    relres = float(nd.rel_residual(L0, E0, xg, b0))
    log(f"bench: rel-resid {relres:.3e}")

    # live per-kernel timing (HIP events around each launch, same stream)
    lib = _lib.load()
    reps = 5
    tot_ms = np.zeros(5)
    tot_cnt = np.zeros(5, dtype=np.int64)
    for _ in range(reps):
        restore()
        kms = (ctypes.c_double * 5)()
        kcnt = (ctypes.c_int * 5)()
        rc = lib.socu_profile_op(_lib.OP_FACTOR_AND_SOLVE, L.data_ptr(), E.data_ptr(), x.data_ptr(), None, 1, N, n, m,
                                 _lib.DTYPE_F64, ctypes.c_void_p(st.cuda_stream), kms, kcnt)
        _lib.check(rc, "profile")
        tot_ms += np.array(kms[:])
        tot_cnt += np.array(kcnt[:])
    kernel_ms = {k: float(tot_ms[i] / reps) for i, k in enumerate(_lib.KERNEL_CLASSES)}
    kernel_launch = {k: int(tot_cnt[i] // reps) for i, k in enumerate(_lib.KERNEL_CLASSES)}
    log(f"bench: per-class kernel ms per step {kernel_ms} launches {kernel_launch}")

    ff, fsv = nd.nd_flops(N, n, m)
    bf, bsv, bfs = nd.nd_bytes(N, n, m, 8)
    # dominant kernel class by time
    dom = max(kernel_ms, key=kernel_ms.get)
    # algorithmic flops per class (fp64)
    cls_flops = _class_flops(N, n, m)
    achieved = cls_flops[dom] / (kernel_ms[dom] * 1e-3) / 1e12
    roofline = {
        "kernel": dom,
        "bound": "mfma",
        "achieved": round(achieved, 4),
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
        "traffic": None,
        "launches_per_step": kernel_launch[dom],
        "ms_per_step": round(kernel_ms[dom], 4),
    }

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline")
        rate, threads, creps = cpu_baseline(L0, E0, b0, args.cpu_seconds)
        cpu = {"value": round(rate, 4), "unit": "factor+solve/s", "cores": threads, "kind": "port",
               "sample": f"oracle/nd_oracle.c factor_and_solve fp64, same n={n} N={N} m={m} problem, {creps} runs"}

    value = 1e3 / ms_per_step
    out = {
        "metric": "factor+solve/s, mixed+refine vs FP64, n=128 N=512 @ rel-resid 1e-10",
        "value": round(value, 3),
        "unit": "factor+solve/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (README generator, seed 0)",
        "config": {"workload": f"ND block-tridiagonal Cholesky factor_and_solve, n={n} N={N} m={m} B=1, "
                               f"alpha={args.alpha}, mode={args.mode}",
                   "n": n, "N": N, "m": m, "alpha": args.alpha, "mode": args.mode,
                   "rel_resid": relres, "gflops": round((ff + fsv) / (ms_per_step * 1e-3) / 1e9, 2),
                   "kernel_ms": kernel_ms},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def _class_flops(N, n, m):
    """Algorithmic flops per kernel class for one factor_and_solve."""
    from oracle import nd

    f = {"potrf": 0.0, "trsm": 0.0, "schur": 0.0, "forward": 0.0, "backward": 0.0}
    for lev in nd.level_schedule(N):
        s = lev["s"]
        for i in lev["nodes"]:
            r = i + s < N
            c = i >= s
            f["potrf"] += n ** 3 / 3
            f["trsm"] += (r + c) * n ** 3
            f["schur"] += (r + c) * n ** 3 + (r and c) * 2 * n ** 3
            f["forward"] += n * n * m + 2 * (r + c) * n * n * m
            f["backward"] += n * n * m + 2 * (r + c) * n * n * m
    return f


if __name__ == "__main__":
    main()
