"""Benchmark of the ND block-tridiagonal Cholesky on one MI355X.

Workload: n=128, N=512 block-tridiagonal SPD systems from the README
generator (diagonal shift alpha), m right-hand sides, one system (B=1).
A step is one factor_and_solve of the whole system with inputs resident in
HBM; the in-place inputs are restored from pristine device copies before each
step, outside the HIP-event-timed region.

Prints one JSON line (the driver contract) with a live roofline figure for the
dominant kernel (HIP events around its launches) and a CPU baseline (the
oracle's C restatement on the host cores).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
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
# HBM bytes moved per factor by the potrf kernel of the mixed path at
# n=128, N=512 (FP32 storage), per pivot precision: rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE, fetch scaled by the write/fetch ratio of the
# bench's own device copies (write exact); per-launch numbers and the commands
# in profiles/potrf_traffic_n128_N512_pmc.md.
MEASURED_TRAFFIC = {("potrf", 128, 512, 1, "float32_f32_pivots"): 56884 * 1024,
                    ("potrf", 128, 512, 1, "float32"): 57811 * 1024}
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
# mixed-solve factor precisions (socu_amd.mixed) and their pivot / Schur types
PRECISIONS = {"float32": "FP32 Schur, FP64 pivots", "float32_f32_pivots": "FP32 Schur, FP32 pivots",
              "bfloat16": "BF16 Schur (FP32 accumulation), FP64 pivots"}
# refinement caps of the timed graphs (the graph replays every captured
# iteration; converged ones are skipped kernel by kernel): FP32 factors need
# two correction solves on the timed problem, BF16 Schur updates about eight
DEFAULT_MAX_ITER = {"float32": 3, "float32_f32_pivots": 3, "bfloat16": 12}


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
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--max-iter", type=int, default=0,
                    help="refinement cap of the timed graphs (0: per precision, DEFAULT_MAX_ITER)")
    ap.add_argument("--precision", default="float32", choices=list(PRECISIONS),
                    help="factor precision of the timed mixed solve: FP32 Schur updates with FP64 diagonal-block "
                         "pivots (float32, the default), FP32 pivots, or BF16 Schur updates; the others are timed "
                         "too and reported beside it")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="skip the batched and gradient workloads")
    ap.add_argument("--mpc-count", type=int, default=4096)
    ap.add_argument("--secondary-steps", type=int, default=20)
    return ap.parse_args()


def capi_factor_and_solve(lib, prefix, L, E, x, B, N, n, m, dtype, stream):
    """One factor_and_solve through a library with the product's C-ABI shape
    (include/socu_hip.h): libsocu_hip.so (prefix "socu_", device pointers and
    a stream) or the oracle's C restatement (prefix "ndo_", host pointers)."""
    rc = getattr(lib, prefix + "cholesky_factor_and_solve")(L, E, x, None, B, N, n, m, dtype, stream)
    if rc != 0:
        raise RuntimeError(f"{prefix}cholesky_factor_and_solve returned {rc}")


def cpu_baseline(L0, E0, b0, seconds):
    """Time the oracle's C restatement (OpenMP over the nodes of a level) on
    the same problem through the C-ABI harness, repeated until `seconds` of
    CPU time have elapsed."""
    from oracle import nd_c
    from socu_amd import _lib

    lib = nd_c.capi()
    threads = nd_c.set_threads(os.cpu_count() or 1)
    N, n, m = b0.shape
    reps, t_total = 0, 0.0
    while t_total < seconds or reps < 2:
        L = L0[None].copy()
        E = E0[None].copy()
        x = b0[None].copy()
        t0 = time.perf_counter()
        capi_factor_and_solve(lib, "ndo_", L.ctypes.data, E.ctypes.data, x.ctypes.data, 1, N, n, m, _lib.DTYPE_F64,
                              None)
        t_total += time.perf_counter() - t0
        reps += 1
        if reps >= 200:
            break
    return reps / t_total, threads, reps


def gpu_direct_ms(L0, E0, b0, dev, st, reps=10):
    """The same harness on the GPU library's direct (eager, no plan or graph)
    socu_cholesky_factor_and_solve, HIP events around each call."""
    import torch

    from socu_amd import _lib

    lib = _lib.load()
    N, n, m = b0.shape
    Lp, Ep, bp = (torch.tensor(a, device=dev) for a in (L0, E0, b0))
    L, E, x = Lp.clone(), Ep.clone(), bp.clone()

    def step():
        capi_factor_and_solve(lib, "socu_", L.data_ptr(), E.data_ptr(), x.data_ptr(), 1, N, n, m, _lib.DTYPE_F64,
                              ctypes.c_void_p(st.cuda_stream))

    def restore():
        L.copy_(Lp)
        E.copy_(Ep)
        x.copy_(bp)

    per, _ = _time_steps(step, restore, st, reps, 2)
    return float(per.mean())


def _profile(lib, fn, nclass):
    kms = (ctypes.c_double * nclass)()
    kcnt = (ctypes.c_int * nclass)()
    fn(kms, kcnt)
    return np.array(kms[:]), np.array(kcnt[:])


def _time_steps(step, restore, st, steps, warmup):
    """Run `warmup` untimed steps, then time `steps` steps with HIP events on
    `st` (each step preceded by an untimed restore of its inputs)."""
    import torch

    for _ in range(warmup):
        restore()
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        restore()
        starts[k].record(st)
        step()
        ends[k].record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = np.array([s_.elapsed_time(e) for s_, e in zip(starts, ends)])
    return per, wall


def bench_mpc_batch(args, dev, st):
    """The uneven MPC batch: 4096 independent systems, N ~ U[10, 200],
    n in {8, 12, 16, 24, 32}, m = 1, FP64 factor_and_solve, one tuned,
    graph-captured batched launch per step.  HBM-bound: the roofline figure
    is the compulsory bytes of every system's factor_and_solve over the step
    time."""
    import torch

    from socu_amd import problems
    from socu_amd.batched import create_batched_cholesky_launch

    log(f"bench: generating the {args.mpc_count}-system uneven MPC batch")
    batch = problems.mpc_batch(args.mpc_count, seed=1)
    total = sum(L.size + E.size + b.size for L, E, b in batch)
    host = np.empty(total)
    spans = []
    o = 0
    for L, E, b in batch:
        sp = []
        for a in (L, E, b):
            host[o:o + a.size] = a.ravel()
            sp.append((o, a.shape))
            o += a.size
        spans.append(sp)
    shapes = [(L.shape[0], L.shape[1]) for L, _, _ in batch]
    del batch
    pristine = torch.from_numpy(host).to(dev)
    work = pristine.clone()

    def views(t):
        return [[t[o:o + int(np.prod(sh))].view(*sh) for o, sh in sp] for sp in spans]

    wv = views(work)
    Ls, Es, xs = [v[0] for v in wv], [v[1] for v in wv], [v[2] for v in wv]
    t0 = time.perf_counter()
    run = create_batched_cholesky_launch(Ls, Es, xs, tune=True, use_cuda_graph=True)
    t_create = time.perf_counter() - t0
    per, _ = _time_steps(run, lambda: work.copy_(pristine), st, args.secondary_steps, 3)
    out_host = work.cpu().numpy()
    worst = 0.0
    for sp in spans:
        (oL, shL), (oE, shE), (ob, shb) = sp
        Lh = host[oL:oL + int(np.prod(shL))].reshape(shL)
        Eh = host[oE:oE + int(np.prod(shE))].reshape(shE)
        bh = host[ob:ob + int(np.prod(shb))].reshape(shb)
        xh = out_host[ob:ob + int(np.prod(shb))].reshape(shb)
        worst = max(worst, problems.rel_residual(Lh, Eh, xh, bh))
    byt = sum(problems.nd_bytes(N, n, 1)[2] for N, n in shapes)
    fl = sum(sum(problems.nd_flops(N, n, 1)) for N, n in shapes)
    ms = float(per.mean())
    gbs = byt / (ms * 1e-3) / 1e9
    res = {
        "workload": f"{args.mpc_count} uneven systems, N~U[10,200], n in {{8,12,16,24,32}}, m=1, FP64 "
                    "factor_and_solve, one graph launch",
        "ms_per_batch": round(ms, 4), "p50_ms": round(float(np.median(per)), 4),
        "p99_ms": round(float(np.percentile(per, 99)), 4),
        "solves_per_s": round(args.mpc_count / (ms * 1e-3), 1),
        "gflops": round(fl / (ms * 1e-3) / 1e9, 2),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 5), "algorithmic_bytes": byt},
        "max_rel_resid": worst, "config": run.config(), "create_s_incl_tune": round(t_create, 3),
    }
    log(f"bench: mpc batch {res}")
    return res


def bench_adjoint(args, dev, st):
    """Differentiable solve at B=256, n=32, N=128, m=4 (FP64): one step is
    the forward factor_and_solve plus the backward transpose solve and
    adjoint scatter."""
    import torch

    from socu_amd import problems
    from socu_amd.autograd import block_tridiag_solve

    B, N, n, m = 256, 128, 32, 4
    g = torch.Generator(device=dev).manual_seed(2)
    D = torch.tril(torch.randn(B, N, n, n, generator=g, device=dev, dtype=torch.float64))
    D += 10 * torch.eye(n, device=dev, dtype=torch.float64)
    Ec = torch.randn(B, N - 1, n, n, generator=g, device=dev, dtype=torch.float64)
    L = torch.matmul(D, D.transpose(-1, -2).contiguous())
    L[:, 1:] += torch.matmul(Ec, Ec.transpose(-1, -2).contiguous())
    E = torch.zeros(B, problems.calculate_off_diag_storage_len(N), n, n, device=dev, dtype=torch.float64)
    E[:, : N - 1] = torch.matmul(Ec, D[:, : N - 1].transpose(-1, -2).contiguous())
    b = torch.randn(B, N, n, m, generator=g, device=dev, dtype=torch.float64)
    w = torch.randn(B, N, n, m, generator=g, device=dev, dtype=torch.float64)
    L.requires_grad_(True)
    E.requires_grad_(True)
    b.requires_grad_(True)

    def step():
        x = block_tridiag_solve(L, E, b)
        torch.autograd.grad(x, (L, E, b), grad_outputs=w)

    def fwd():
        with torch.no_grad():
            block_tridiag_solve(L, E, b)

    per, _ = _time_steps(step, lambda: None, st, args.secondary_steps, 3)
    perf, _ = _time_steps(fwd, lambda: None, st, args.secondary_steps, 3)
    ms, msf = float(per.mean()), float(perf.mean())
    # the forward is one batched factor_and_solve: its compulsory bytes and
    # flops are B times those of one system
    ff, fsv = problems.nd_flops(N, n, m)
    res = {"workload": f"B={B} n={n} N={N} m={m} FP64, forward + backward (dL, dE, db)",
           "ms_per_step": round(ms, 4), "forward_ms": round(msf, 4),
           "backward_ms": round(ms - msf, 4), "systems_per_s": round(B / (ms * 1e-3), 1),
           "forward_path": path_fractions(B * (ff + fsv), B * problems.nd_bytes(N, n, m)[2], msf)}
    log(f"bench: adjoint {res}")
    return res


def path_fractions(flops, nbytes, ms):
    """Whole-path rates of one step: algorithmic flops and bytes over the step
    time, as fractions of MI355X's FP64, FP32 and dense BF16 MFMA peaks and of
    HBM bandwidth."""
    tf = flops / (ms * 1e-3) / 1e12
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"tflops": round(tf, 4), "frac_fp64_peak": round(tf / FP64_PEAK_TFLOPS, 5),
            "frac_fp32_peak": round(tf / FP32_PEAK_TFLOPS, 5), "frac_bf16_peak": round(tf / BF16_PEAK_TFLOPS, 6),
            "gb_per_s": round(gbs, 2), "frac_hbm": round(gbs / HBM_PEAK_GBS, 5)}


def _stats(per):
    return {"ms": round(float(per.mean()), 4), "p50_ms": round(float(np.median(per)), 4),
            "p99_ms": round(float(np.percentile(per, 99)), 4)}


def bench_fp64(L0, E0, b0, dev, st, steps, warmup, profile=True):
    """Plain FP64 factor_and_solve of one system, graph-captured; inputs
    restored from pristine device copies outside the timed region."""
    import torch

    from socu_amd import _lib, problems
    from socu_amd.block_tridiag_solver import create_cholesky_factor_and_solve_launch

    N, n, m = b0.shape
    Lp, Ep, bp = (torch.tensor(a, device=dev) for a in (L0, E0, b0))
    L, E, x = Lp.clone(), Ep.clone(), bp.clone()

    def restore():
        L.copy_(Lp)
        E.copy_(Ep)
        x.copy_(bp)

    fs = create_cholesky_factor_and_solve_launch(L, E, x, use_cuda_graph=True)
    per, _ = _time_steps(fs, restore, st, steps, warmup)
    res = problems.rel_residual(L0, E0, x.cpu().numpy(), b0)
    out = dict(_stats(per), rel_resid=res)
    if profile:
        lib = _lib.load()
        ncls = len(_lib.KERNEL_CLASSES)
        k64 = np.zeros(ncls)
        for _ in range(3):
            restore()
            km, _ = _profile(lib, lambda a, b: _lib.check(lib.socu_profile_op(
                _lib.OP_FACTOR_AND_SOLVE, L.data_ptr(), E.data_ptr(), x.data_ptr(), None, 1, N, n, m, _lib.DTYPE_F64,
                None, ctypes.c_void_p(st.cuda_stream), a, b), "profile"), ncls)
            k64 += km / 3
        out["kernel_ms"] = {k: round(float(k64[i]), 4) for i, k in enumerate(_lib.KERNEL_CLASSES)}
    return out


def bench_mixed(L0, E0, b0, Lp, Ep, dev, st, precision, tol, max_iter, steps, warmup):
    """Mixed-precision factor + FP64 refinement of one system (graph-captured,
    b restored outside the timed region); returns the timing, the refinement
    outcome and the launch (for profiling)."""
    import torch

    from socu_amd import problems
    from socu_amd.mixed import create_cholesky_mixed_factor_and_solve_launch

    bp = torch.tensor(b0, device=dev)
    xm = bp.clone()
    mx = create_cholesky_mixed_factor_and_solve_launch(Lp, Ep, xm, precision=precision, tol=tol, max_iter=max_iter,
                                                       use_cuda_graph=True)
    per, wall = _time_steps(mx, lambda: xm.copy_(bp), st, steps, warmup)
    rr, its = mx.stats()
    res = problems.rel_residual(L0, E0, xm.cpu().numpy(), b0)
    out = dict(_stats(per), precision=precision, rel_resid=res, device_rel_resid=float(rr[0]),
               correction_solves=int(its[0]), max_iter=max_iter,
               wall_ms_per_step_incl_restore=round(wall * 1e3 / steps, 4))
    return out, mx


def bench_c2(args, dev, st):
    """FP64 factor_and_solve at n=64, N=1024, m=16 (one system)."""
    from socu_amd import problems

    n, N, m = 64, 1024, 16
    L0, E0, b0 = problems.readme_problem(n, N, m, alpha=10.0, seed=0)
    r = bench_fp64(L0, E0, b0, dev, st, args.secondary_steps, 3, profile=False)
    ff, fsv = problems.nd_flops(N, n, m)
    res = {"workload": f"FP64 factor_and_solve, n={n} N={N} m={m} B=1, README generator alpha=10 seed 0",
           **r, "factor_and_solve_per_s": round(1e3 / r["ms"], 2),
           "gflops": round((ff + fsv) / (r["ms"] * 1e-3) / 1e9, 2),
           "path": path_fractions(ff + fsv, problems.nd_bytes(N, n, m)[2], r["ms"])}
    log(f"bench: C2 {res}")
    return res


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only; runs on several GPUs are not supported "
              "(independent systems would shard over GPUs with no communication).", file=sys.stderr)
        sys.exit(2)

    import torch

    from socu_amd import _lib, problems

    n, N, m = args.n, args.N, args.m
    dev = torch.device("cuda", 0)
    lib = _lib.load()
    ncls = len(_lib.KERNEL_CLASSES)
    st = torch.cuda.current_stream(dev)
    log(f"bench: generating n={n} N={N} m={m} alpha={args.alpha}")
    L0, E0, b0 = problems.readme_problem(n, N, m, alpha=args.alpha, seed=0)
    Lp = torch.tensor(L0, device=dev)
    Ep = torch.tensor(E0, device=dev)
    ff, fsv = problems.nd_flops(N, n, m)
    bytes64 = problems.nd_bytes(N, n, m)[2]

    # --- plain FP64 factor_and_solve: the denominator of the metric -------
    r64 = bench_fp64(L0, E0, b0, dev, st, args.steps, args.warmup)
    ms64 = r64["ms"]
    r64["path"] = path_fractions(ff + fsv, bytes64, ms64)
    r64["gflops"] = round((ff + fsv) / (ms64 * 1e-3) / 1e9, 2)
    log(f"bench: fp64 factor_and_solve {r64}")

    # --- mixed factor + FP64 refinement: the metric -----------------------
    cap = lambda prec: args.max_iter or DEFAULT_MAX_ITER[prec]  # noqa: E731
    rm, mx = bench_mixed(L0, E0, b0, Lp, Ep, dev, st, args.precision, args.tol, cap(args.precision), args.steps,
                         args.warmup)
    msm = rm["ms"]
    log(f"bench: mixed+refine ({args.precision}) {rm}")
    # the same graph captured with only the iterations this problem needs:
    # the difference is the cost of the converged (no-op) iterations
    need = max(rm["correction_solves"], 1)
    rneed, _ = bench_mixed(L0, E0, b0, Lp, Ep, dev, st, args.precision, args.tol, need, args.steps, args.warmup)
    log(f"bench: mixed+refine with max_iter={need}: {rneed['ms']:.4f} ms")
    km_mixed, kc_mixed = _profile(lib, lambda a, b: _lib.check(lib.socu_mixed_profile(
        mx._h, ctypes.c_void_p(st.cuda_stream), a, b), "mixed profile"), ncls)
    kernel_ms = {k: round(float(km_mixed[i]), 4) for i, k in enumerate(_lib.KERNEL_CLASSES)}
    kernel_launch = {k: int(kc_mixed[i]) for i, k in enumerate(_lib.KERNEL_CLASSES)}
    log(f"bench: mixed per-class kernel ms {kernel_ms} launches {kernel_launch}")
    # whole-path rates: the FP64 path's algorithmic flops and bytes plus, per
    # correction solve, an FP32 solve and an FP64 residual
    solves = max(rm["correction_solves"], 1)
    cb = problems.class_bytes(N, n, m, w=4)
    mixed_bytes = problems.nd_bytes(N, n, m, w=4)[0] + solves * (cb["forward"] + cb["backward"] + cb["refine"])
    mixed_flops = ff + solves * (fsv + problems.block_matvec_flops(N, n, m))
    rm["path"] = path_fractions(mixed_flops, mixed_bytes, msm)

    # --- the other factor precisions, same protocol (reported beside) -----
    variants = {}
    for prec in PRECISIONS:
        if prec == args.precision:
            continue
        rv, _ = bench_mixed(L0, E0, b0, Lp, Ep, dev, st, prec, args.tol, cap(prec), args.steps, args.warmup)
        rv["speedup_vs_fp64"] = round(ms64 / rv["ms"], 4)
        need_v = max(rv["correction_solves"], 1)
        if need_v < cap(prec):
            rvn, _ = bench_mixed(L0, E0, b0, Lp, Ep, dev, st, prec, args.tol, need_v, args.steps, args.warmup)
            rv["ms_with_max_iter_needed"] = rvn["ms"]
        variants[prec] = rv
        log(f"bench: mixed+refine ({prec}) {rv}")

    # --- C3 with eight right-hand sides ------------------------------------
    c3_m8 = None
    if not args.no_secondary and m == 1:
        L8, E8, b8 = problems.readme_problem(n, N, 8, alpha=args.alpha, seed=0)
        r8 = bench_fp64(L8, E8, b8, dev, st, args.secondary_steps, 3, profile=False)
        c3_m8 = {"workload": f"n={n} N={N} m=8 B=1, alpha={args.alpha}", "fp64": r8}
        for prec in PRECISIONS:
            rv, _ = bench_mixed(L8, E8, b8, Lp, Ep, dev, st, prec, args.tol, cap(prec), args.secondary_steps, 3)
            rv["speedup_vs_fp64"] = round(r8["ms"] / rv["ms"], 4)
            c3_m8[prec] = rv
        log(f"bench: C3 m=8 {c3_m8}")

    # --- the literal alphas of the configuration (numerically singular at n=128)
    literal = {}
    if n == 128:
        from socu_amd.block_tridiag_solver import create_cholesky_factor_and_solve_launch
        from socu_amd.mixed import create_cholesky_mixed_factor_and_solve_launch

        for alpha in (10.0, 2.0):
            L1, E1, b1 = problems.readme_problem(n, N, m, alpha=alpha, seed=0)
            L1t, E1t, x1 = torch.tensor(L1, device=dev), torch.tensor(E1, device=dev), torch.tensor(b1, device=dev)
            m1 = create_cholesky_mixed_factor_and_solve_launch(L1t, E1t, x1, tol=args.tol, max_iter=cap("float32"))
            m1()
            rr1, it1 = m1.stats()
            xf = torch.tensor(b1, device=dev)
            Lf, Ef = L1t.clone(), E1t.clone()
            create_cholesky_factor_and_solve_launch(Lf, Ef, xf)()
            diag = np.diagonal(Lf.cpu().numpy(), axis1=1, axis2=2)
            bad = np.where(~np.isfinite(diag).all(1))[0]
            literal[str(alpha)] = {"mixed_rel_resid": float(rr1[0]), "mixed_solves": int(it1[0]),
                                   "fp64_rel_resid": problems.rel_residual(L1, E1, xf.cpu().numpy(), b1),
                                   "fp64_breakdown_nodes": [int(v) for v in bad[:16]]}
        log(f"bench: literal alphas: {literal}")

    # --- roofline of the dominant kernel of the timed path -----------------
    # k_potrf_blk (the diagonal-block factor; one launch per ND level) is the
    # largest single kernel: algorithmic flops n^3/3 per node (the explicit
    # inverse it also forms is not counted), over its summed launch time
    # (HIP events around each launch in an eager run of the same solve).  With
    # FP64 pivots it computes in FP64.
    dom = "potrf"
    per_flops = problems.class_flops(N, n, m)[dom]
    achieved = per_flops / (kernel_ms[dom] * 1e-3) / 1e12 if kernel_ms[dom] > 0 else 0.0
    peak = FP32_PEAK_TFLOPS if args.precision == "float32_f32_pivots" else FP64_PEAK_TFLOPS
    roofline = {
        "kernel": "k_potrf_blk",
        "bound": "mfma",
        "achieved": round(achieved, 4),
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": round(achieved / peak, 5),
        "traffic": MEASURED_TRAFFIC.get((dom, n, N, m, args.precision)),
        "traffic_algorithmic_bytes": N * n * (n + 1) // 2 * 4 + N * n * n * 4,
        "algorithmic_per_step": per_flops,
        "launches_per_step": kernel_launch[dom],
        "avg_launch_us": round(kernel_ms[dom] * 1e3 / max(kernel_launch[dom], 1), 2),
        "ms_per_step": kernel_ms[dom],
    }

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline")
        rate, threads, creps = cpu_baseline(L0, E0, b0, args.cpu_seconds)
        cpu = {"value": round(rate, 4), "unit": "factor+solve/s", "cores": threads, "kind": "port",
               "sample": f"oracle/nd_oracle.c FP64 factor_and_solve (reaches rel-resid < 1e-13) through its "
                         f"C-ABI-shaped entry point ndo_cholesky_factor_and_solve, same n={n} N={N} m={m} "
                         f"problem, {creps} runs",
               "gpu_same_harness_fp64_ms": round(gpu_direct_ms(L0, E0, b0, dev, st), 4)}

    secondary = None
    if not args.no_secondary:
        secondary = {"c2_fp64": bench_c2(args, dev, st), "c3_m8": c3_m8,
                     "mpc_batched": bench_mpc_batch(args, dev, st), "adjoint": bench_adjoint(args, dev, st)}

    mixed_ok = rm["rel_resid"] <= args.tol
    value = 1e3 / msm if mixed_ok else 0.0
    out = {
        "metric": "factor+solve/s, mixed+refine vs FP64, n=128 N=512 @ rel-resid 1e-10",
        "value": round(value, 3),
        "unit": "factor+solve/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(msm, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"float32": "f32 Schur + f64 pivots, f64 refinement",
                  "float32_f32_pivots": "f32 factor (f32 pivots), f64 refinement",
                  "bfloat16": "bf16 Schur (f32 acc) + f64 pivots, f64 refinement"}[args.precision],
        "data": "synthetic (README generator, seed 0)",
        "config": {"workload": f"ND block-tridiagonal Cholesky, mixed factor ({PRECISIONS[args.precision]}) + FP64 "
                               f"refinement to rel-resid {args.tol:g}, n={n} N={N} m={m} B=1, alpha={args.alpha} "
                               "(the literal alpha=10 and alpha=2 are numerically singular at n=128: "
                               "literal_alpha)",
                   "n": n, "N": N, "m": m, "alpha": args.alpha, "precision": args.precision,
                   "mixed": rm, "mixed_ms_with_max_iter_needed": rneed["ms"],
                   "speedup_vs_fp64": round(ms64 / msm, 4),
                   "fp64_factor_and_solve_per_s": round(1e3 / ms64, 3), "fp64": r64,
                   "variants": variants, "kernel_ms": kernel_ms, "kernel_launches": kernel_launch,
                   "literal_alpha": literal},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
