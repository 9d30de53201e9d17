#!/usr/bin/env python3
"""Benchmark: time until ||b - A x|| / ||b|| <= 1e-12 with the spaND preconditioner
(factorization + PCG) on the 3D 7-point Dirichlet Laplacian 64^3 (N = 262,144), tol 1e-2,
12 levels, geometric MND, 4 right-hand sides b ~ N(0,1), on one MI355X.

One step = factorize (level-batched HIP kernels) + FP64 PCG on the 4 right-hand sides until
every column's recursive residual is below 1e-12 (b and x resident in HBM before the timed
region starts). The headline runs the FP32 factorization; the same steps with the FP64
factorization, and the CPU oracle (one core, FP64), are reported beside it. The true
residual ||b - A x|| / ||b|| of the last step is checked in FP64 on the host.

Kernel timing: during the warmup steps every library launch is bracketed by HIP events on
its stream (spand_set_profiling); the family with the largest summed time is the dominant
kernel. During the timed steps only that family carries events (the stepped RRQR's steps
per dependency wave: two events around all steps of a wave), so the instrumentation stays
out of the measured time; the roofline's achieved rate is the algorithmic bytes (or flops)
of those launches divided by their summed event time.

Warmup: the first factorization of a pattern grows the device arenas and the pinned
staging buffer, and the second still spends about 15 ms more host time between launches
than the later ones (the device work is the same); the default 2 warmup steps leave both
out, so the timed steps are the steady state of repeated factor + solve.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--n 64] [--nlevels 12]
Prints one JSON line on stdout (progress on stderr).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time to ‖r‖/‖b‖≤1e-12 (factor+PCG), 64³ Laplacian: FP32-batched vs FP64 GPU"
SEED = 2019
NRHS = 4
TOL = 1e-2
KRYLOV_TOL = 1e-12
MAXIT = 200
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}
# kernel families bound by memory traffic vs by arithmetic (all of them are reported)
FLOP_BOUND = {"potrf", "trsm", "schur"}
# HBM traffic per dispatch of the dominant kernel, measured with rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE on this workload and scaled by the same counters on a plain
# 256 MiB device copy (FETCH_SIZE reads 0.5x of the copied bytes, WRITE_SIZE 1.0x).
# Keyed by (family, factor dtype, grid edge n, levels) and valid only for the kernel source
# it was measured on (SHA-256 prefix of csrc/kernels_factor.hip): after any change to that
# file the run reports traffic null until the counters are measured again
# (profiles/rrqr_wave_laplace64_hbm_traffic.md gives the commands).
TRAFFIC_MEASURED = {
    ("rrqr_wave", "f32", 64, 12): {"fetch_kib": 16302.47, "write_kib": 29219.79, "fetch_factor": 2.0,
                                   "write_factor": 1.0, "src_sha16": "0c433369a10ce38a"},
    ("rrqr_wave", "f64", 64, 12): {"fetch_kib": 22373.32, "write_kib": 67373.68, "fetch_factor": 2.0,
                                   "write_factor": 1.0, "src_sha16": "0c433369a10ce38a"},
}


def kernel_source_sha16():
    import hashlib

    path = os.path.join(ROOT, "spand_public_amd", "csrc", "kernels_factor.hip")
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_tree(S, A, X, nlevels, prec):
    from spand_public_amd.problems import symmetric_graph

    t = S.Tree(nlevels)
    t.set_use_geo(True)
    t.set_Xcoo(X)
    t.set_tol(TOL)
    t.set_precision(prec)
    t0 = time.perf_counter()
    t.partition(symmetric_graph(A))
    t.assemble(A)
    return t, time.perf_counter() - t0


def run_steps(S, torch, t, Am, d_b, d_x, N, steps, warmup, profile):
    """warmup + timed steps of factorize + PCG; returns (s/step, iters, all-family warmup
    profile, timed profile of the dominant family)."""
    its = None

    def one_step():
        nonlocal its
        t.factorize()
        d_x.zero_()
        its, info = S.cg_device(Am, t, d_b.data_ptr(), d_x.data_ptr(), NRHS, N, iters=MAXIT, tol=KRYLOV_TOL)
        if info["status"] != 0:
            raise RuntimeError(f"PCG did not converge in {MAXIT} iterations: {its}")

    # warmup: every kernel family timed, to find the dominant one
    t.set_profiling(profile)
    t.profile_reset()
    for _ in range(warmup):
        one_step()
    torch.cuda.synchronize()
    wprof = t.kernel_profile() if profile else {}
    dominant = max(wprof.items(), key=lambda kv: kv[1]["ms"])[0] if wprof else "rrqr_step"
    # timed steps: only the dominant family carries events (the stepped RRQR is timed per
    # dependency wave, two events around all of its steps)
    t.set_profiling(dominant if profile else False)
    t.profile_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one_step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(steps, 1)
    tprof = t.kernel_profile() if profile else {}
    t.set_profiling(False)
    return dt, its.tolist(), wprof, tprof


def roofline(tprof, wprof, steps, dtype, n=64, nlevels=12):
    if not tprof:
        return None
    fam, k = max(tprof.items(), key=lambda kv: kv[1]["ms"])
    avg_s = k["ms"] / 1e3 / k["launches"]
    if fam in FLOP_BOUND:
        achieved = k["flops"] / k["launches"] / avg_s / 1e12
        peak, unit, bound = MFMA_PEAK_TFLOPS[dtype], "TFLOP/s", "mfma"
    else:
        achieved = k["bytes"] / k["launches"] / avg_s / 1e9
        peak, unit, bound = HBM_PEAK_GBS, "GB/s", "hbm"
    tm = TRAFFIC_MEASURED.get((fam, dtype, n, nlevels))
    traffic = None
    if tm and tm["src_sha16"] == kernel_source_sha16():
        traffic = round((tm["fetch_kib"] * tm["fetch_factor"] + tm["write_kib"] * tm["write_factor"]) * 1024.0)
    out = {
        "bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
        "frac": round(achieved / peak, 6), "traffic": traffic,
        "kernel": fam, "launches_per_step": round(k["launches"] / steps, 1),
        "avg_launch_us": round(avg_s * 1e6, 3),
        "algorithmic_bytes_per_launch": round(k["bytes"] / k["launches"], 1),
        "algorithmic_flops_per_launch": round(k["flops"] / k["launches"], 1),
    }
    if wprof:
        out["share_of_kernel_time"] = round(wprof[fam]["ms"] / sum(v["ms"] for v in wprof.values()), 4)
    return out


def cpu_baseline(A, X, B, nl):
    """The CPU oracle (FP64 restatement of the reference algorithm, one core) on the same
    problem: one factorization and PCG on the first right-hand side; the 4-RHS job time is
    factor + 4 x (one-RHS PCG), since the reference runs its Krylov solver one vector at a time."""
    import oracle

    t0 = time.perf_counter()
    o, _ = oracle.setup_tree(A, X, nl, TOL)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.factorize()
    t_fact = time.perf_counter() - t0
    t0 = time.perf_counter()
    x, it, _ = o.cg(A, B[:, 0], iters=MAXIT, tol=KRYLOV_TOL)
    t_cg = time.perf_counter() - t0
    res = float(np.linalg.norm(A.matvec(x) - B[:, 0]) / np.linalg.norm(B[:, 0]))
    log(f"cpu oracle: factor {t_fact:.2f} s, PCG (1 RHS) {t_cg:.2f} s in {it} iterations, residual {res:.1e}, "
        f"partition+assemble {t_setup:.2f} s")
    return {"value": round(t_fact + NRHS * t_cg, 3), "unit": "s", "cores": 1, "kind": "port",
            "sample": f"one oracle factorization of the same {A.n}-unknown problem + PCG on RHS 1 of {NRHS} "
                      f"({it} iterations, residual {res:.1e}); value = factor + {NRHS} x PCG",
            "factor_s": round(t_fact, 3), "pcg_one_rhs_s": round(t_cg, 3), "iters": int(it),
            "host": host_info()}


def host_info():
    """CPU model, clock, cores and the dense arithmetic of the CPU baseline."""
    model, mhz = None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if model is None and line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                elif mhz is None and line.startswith("cpu MHz"):
                    mhz = float(line.split(":", 1)[1])
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu": model or platform.processor() or platform.machine(), "cpu_mhz": mhz,
            "cores_available": avail, "cores_used": 1,
            "dense_backend": "none: the oracle's own unblocked C++ loops (no BLAS or LAPACK), "
                             "g++ -O3 -march=x86-64-v3 -ffp-contract=off, one thread; the reference "
                             "links sequential MKL or OpenBLAS, whose blocked kernels would be faster"}


def run_differentiable(args):
    """The differentiable-solve workload: 2D 512^2 variable-coefficient diffusion, kappa =
    exp(g) with g a Gaussian random field (sigma 1, correlation length 16 cells), 32 RHS
    N(0,1), loss 1/2 ||x - x*||^2 with x* ~ N(0,1). One step = re-assembly + factorization
    (FP64, tol 1e-2) + PCG forward + adjoint PCG + sampled outer product + the chain to
    kappa. Prints one JSON line with the step time and a central-difference check
    (h = 1e-6) of dL/dkappa and dL/db in random directions."""
    import torch

    import spand_public_amd.autograd as ad
    from spand_public_amd.problems import DiffusionFV, auto_nlevels, gaussian_field

    n, nrhs = 512, 32
    fv = DiffusionFV(n, 2)
    kappa = np.exp(gaussian_field(n, 2, 1.0, 16.0, SEED))
    solver = ad.fv_solver(fv, auto_nlevels(fv.N), tol=TOL, pcg_tol=1e-13)
    rng = np.random.default_rng(SEED + 1)
    dev = torch.device("cuda:0")
    b0 = torch.tensor(rng.standard_normal((fv.N, nrhs)), device=dev)
    xs = torch.tensor(rng.standard_normal((fv.N, nrhs)), device=dev)
    k0 = torch.tensor(kappa, device=dev)

    def loss(k, b):
        x = ad.spd_solve(ad.fv_values(fv, k), b, solver)
        return 0.5 * ((x - xs) ** 2).sum()

    def step():
        k = k0.clone().requires_grad_()
        b = b0.clone().requires_grad_()
        L = loss(k, b)
        L.backward()
        torch.cuda.synchronize()
        return k.grad, b.grad

    for _ in range(args.warmup):
        step()
    solver.iters.clear()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gk, gb = step()
    dt = (time.perf_counter() - t0) / max(args.steps, 1)
    its = solver.iters[:2]
    h, errs = 1e-6, []
    with torch.no_grad():
        for q in range(4):
            d = torch.tensor(rng.standard_normal(fv.N), device=dev) * k0
            fd = float((loss(k0 + h * d, b0) - loss(k0 - h * d, b0)) / (2 * h))
            errs.append(abs(fd - float((gk * d).sum())) / abs(fd))
            d = torch.tensor(rng.standard_normal((fv.N, nrhs)), device=dev)
            fd = float((loss(k0, b0 + h * d) - loss(k0, b0 - h * d)) / (2 * h))
            errs.append(abs(fd - float((gb * d).sum())) / abs(fd))
    print(json.dumps({"metric": "differentiable solve step (re-assembly + factorization + PCG + adjoint PCG)",
                      "value": round(dt, 6), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "higher_is_better": False, "dtype": "f64", "data": "synthetic",
                      "config": {"workload": f"2D {n}^2 variable-kappa diffusion, {nrhs} RHS, loss 1/2||x-x*||^2",
                                 "nlevels": auto_nlevels(fv.N), "tol": TOL},
                      "pcg_iters_forward": its[0] if its else None, "pcg_iters_adjoint": its[1] if len(its) > 1 else None,
                      "fd_max_rel_err": max(errs)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--nlevels", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-factor comparison")
    ap.add_argument("--fp64-only", action="store_true",
                    help="only the FP64-factor steps (for profiling its kernels); the JSON line carries them")
    ap.add_argument("--no-profile", action="store_true", help="no per-launch HIP events in the timed steps")
    ap.add_argument("--differentiable", action="store_true",
                    help="only the differentiable-solve workload (2D 512^2 variable kappa, 32 RHS)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one device copy of MIB MiB (the calibration kernel for HBM-traffic counters)")
    args = ap.parse_args()
    if args.copy_probe > 0:
        import torch

        n = args.copy_probe << 20
        src = torch.ones(n // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build", file=sys.stderr)
        sys.exit(2)
    if args.differentiable:
        run_differentiable(args)
        return

    import torch

    import spand_public_amd as S
    from spand_public_amd.problems import linspace_nd, neglapl

    n, nl = args.n, args.nlevels
    A, X = neglapl(n, 3), linspace_nd(n, 3)
    N = A.n
    B = np.random.default_rng(SEED).standard_normal((N, NRHS))
    dev = torch.device("cuda:0")
    d_b = torch.tensor(B.T.copy(), dtype=torch.float64, device=dev)  # column-major N x NRHS, ld N
    d_x = torch.zeros_like(d_b)
    Am = S.Matrix(A)

    results = {}
    for name, prec in (("f32", S.FP32), ("f64", S.FP64)):
        if name == "f64" and args.no_fp64:
            continue
        if name == "f32" and args.fp64_only:
            continue
        t, t_setup = make_tree(S, A, X, nl, prec)
        dt, its, wprof, tprof = run_steps(S, torch, t, Am, d_b, d_x, N, args.steps, args.warmup, not args.no_profile)
        x = d_x.cpu().numpy().T
        res = float((np.linalg.norm(A.matvec(x) - B, axis=0) / np.linalg.norm(B, axis=0)).max())
        fi = t.factor_info()
        results[name] = {"s": dt, "iters": its, "relres": res, "wprof": wprof, "tprof": tprof,
                         "factor_s": fi["seconds"]}
        results[name]["tune"] = None
        log(f"{name}: {dt * 1e3:.2f} ms/step (last factorize {fi['seconds'] * 1e3:.2f} ms), PCG iterations {its}, "
            f"true residual {res:.2e}, partition+assemble {t_setup:.2f} s")
        nw = max(args.warmup, 1)
        for fam, k in sorted(wprof.items(), key=lambda kv: -kv[1]["ms"]):
            log(f"   {fam:12s} {k['launches'] / nw:9.1f} launches/step {k['ms'] / nw:9.3f} ms/step "
                f"{k['bytes'] / max(k['ms'], 1e-9) / 1e6:9.1f} GB/s {k['flops'] / max(k['ms'], 1e-9) / 1e9:9.3f} TFLOP/s"
                "   (warmup, every launch timed)")
        tune = t.level_tune()
        log("   level  rows_med  cols_med  elim_med | reg  small_elems  step_waves  trsm_cw  (launch configuration)")
        for l, r in enumerate(tune):
            log(f"   {l:5d} {r['rows_med']:9d} {r['cols_med']:9d} {r['elim_med']:9d} | {r['rrqr_reg']:3d} "
                f"{r['small_elems']:12d} {r['step_waves']:11d} {r['trsm_cw']:8d}")
        results_tune = [{k: r[k] for k in ("rows_med", "cols_med", "elim_med", "rrqr_reg", "small_elems",
                                           "step_waves", "trsm_cw")} for r in tune]
        ls = t.level_stats()
        log("   level  plan_ms  device_ms  ranks_ms  merge_ms  solveplan_ms   (last factorization; host vs device)")
        for l in range(nl):
            log(f"   {l:5d} {ls[l, 10] * 1e3:8.2f} {ls[l, 11] * 1e3:10.2f} {ls[l, 12] * 1e3:9.2f} {ls[l, 13] * 1e3:9.2f}"
                f" {ls[l, 18] * 1e3:13.2f}")
        results[name]["tune"] = results_tune
        results[name]["status"] = t.status_flags()
        if results[name]["status"]["step_throttle_gave_up"]:
            log(f"   {name}: the stepped-RRQR launch throttle gave up after a 2 s wait (steps launched unthrottled)")
        if not res <= KRYLOV_TOL:
            raise RuntimeError(f"{name}: true residual {res:.2e} above {KRYLOV_TOL}")
        del t

    if args.fp64_only:
        r64 = results["f64"]
        print(json.dumps({"metric": METRIC + " (FP64 factor only)", "value": round(r64["s"], 6), "unit": "s",
                          "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": False,
                          "dtype": "f64", "data": "synthetic", "pcg_iters": r64["iters"], "relres": r64["relres"],
                          "roofline": roofline(r64["tprof"], r64["wprof"], args.steps, "f64", n, nl)}), flush=True)
        return
    head = results["f32"]
    out = {
        "metric": METRIC,
        "value": round(head["s"], 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["s"] * 1e3, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"3D 7-point Dirichlet Laplacian {n}^3 (N={N}), spaND tol {TOL}, {nl} levels, "
                        f"geometric MND, {NRHS} RHS N(0,1) seed {SEED}, FP32 factor + FP64 PCG to {KRYLOV_TOL}",
            "factor_dtype": "f32", "krylov_dtype": "f64", "nrhs": NRHS,
        },
        "pcg_iters": head["iters"],
        "relres": head["relres"],
        "roofline": roofline(head["tprof"], head["wprof"], args.steps, "f32", n, nl),
        "level_config": head["tune"],
        "status_flags": head["status"],
    }
    if "f64" in results:
        r64 = results["f64"]
        out["fp64_gpu"] = {"value": round(r64["s"], 6), "unit": "s", "pcg_iters": r64["iters"], "relres": r64["relres"],
                           "roofline": roofline(r64["tprof"], r64["wprof"], args.steps, "f64", n, nl),
                           "status_flags": r64["status"]}
        out["speedup_f32_vs_f64"] = round(r64["s"] / head["s"], 3)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(A, X, B, nl)
        out["speedup_vs_cpu"] = round(out["cpu_baseline"]["value"] / head["s"], 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
