#!/usr/bin/env python3
"""Benchmark of the assemble-then-iterate path on one MI355X.

Workload (the 128^3 Hex8 Poisson case): steady 3-D Poisson on a structured
Hex8 mesh of 128^3 elements (2,146,689 nodes, 57,066,625 nonzeros),
per-element conductivity lognormal(0, 1) from a seed, source 1, zero
Dirichlet on all faces, CG + Jacobi to relative residual 1e-10.

One step = one solve of the assembled system to 1e-10 from a zero initial
guess, in the headline form: FP32 inner CG (FP32 values, vectors and
accumulation) inside FP64 iterative refinement.  Plain FP64 CG in the
reference's loop order (true residual every iteration) and a fused FP64 CG
are timed beside it, as are GPU assembly and the SpMV kernels.

    python bench.py [--steps K] [--warmup W] [--n 128]

Prints one JSON line.  value = DOFs x inner iterations / solve time.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0


def spmv_bytes(nnz: int, n: int, fp32: bool = False) -> int:
    """Algorithmic bytes of one CSR SpMV (int32 indices, x read once, y written once)."""
    return (8 * nnz + 4 * (n + 1) + 8 * n) if fp32 else (12 * nnz + 4 * (n + 1) + 16 * n)


def fused_cg32_bytes(nnz: int, n: int) -> int:
    """Fused FP32 CG step 1 (q = A z + beta q, p = z + beta p, p.q):
    vals32 + cols + rows + z + q (r/w) + p (r/w)."""
    return 8 * nnz + 4 * (n + 1) + 4 * n + 16 * n


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(sys_, kw, iters: int):
    """The oracle's HUTI CG (serial C restatement of the reference loop with
    reference-BLAS ddot/dnrm2) on the same scaled system, for a bounded
    number of iterations on one host core."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import oracle_lib as O
    perm, rows, cols, diag = sys_.pattern()
    vals, rhs = sys_.values()

    class P:
        pass
    pat = P()
    pat.n, pat.rows, pat.cols, pat.diag = sys_.n, rows, cols, diag
    v, b, x0 = O.scaled_system(pat, vals, rhs)
    t0 = time.perf_counter()
    info, k, x, hist = O.huti_solve(rows, cols, diag, v, b, x0, method=0, precond=1, tol=1e-300, maxit=iters)
    dt = time.perf_counter() - t0
    done = min(k, iters)
    return {"value": sys_.n * done / dt, "unit": "DOF-iter/s", "cores": 1, "kind": "port",
            "sample": "%d HUTI-faithful CG iterations (2 SpMV + ddot/dnrm2 each) of the 128^3 system, "
                      "serial oracle restatement, %.1f s" % (done, dt),
            "seconds": dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--sigma", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-iters", type=int, default=150)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-forms", action="store_true", help="time only the headline form")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    from elmerfem_amd import _lib, workloads as W
    from elmerfem_amd.system import copy_bench

    _lib.init(0)
    t0 = time.perf_counter()
    w = W.poisson_hex8(args.n, sigma=args.sigma, seed=args.seed)
    m = w["mesh"]
    log("mesh %d^3: %d nodes, %d elements (%.1fs)" % (args.n, m.nnodes, m.nelem, time.perf_counter() - t0))
    t0 = time.perf_counter()
    sys_, t_asm_ms = W.build_system(w, optimize_bandwidth=True)
    t_setup = time.perf_counter() - t0
    n, nnz = sys_.n, sys_.nnz
    log("system n=%d nnz=%d, host pattern+maps+assembly %.1fs, GPU assembly %.2f ms" % (n, nnz, t_setup, t_asm_ms))
    # assembly again, warm
    t_asm = []
    for _ in range(3):
        t_asm.append(sys_.assemble(w["physics"], m.coords, elem_coef=w["coef"], elem_load=w["load"]))
    sys_.dirichlet_nodes(w["bc_nodes"])
    t_asm_ms = float(np.median(t_asm))
    npe = m.elements.shape[1]
    asm_bytes = 24 * m.nnodes + 8 * m.nelem + 4 * npe * m.nelem + 4 * npe * npe * m.nelem + 8 * nnz + 8 * n

    kw = w["kw"]
    kw.tol = 1e-10
    kw.maxit = 20000
    copy_gbs, _ = copy_bench(1 << 30, reps=20)
    forms = {}

    def run(name, precision, mode, reps, profile_every=0):
        out = []
        for _ in range(reps):
            t = time.perf_counter()
            x, rep = sys_.solve(kw, precision=precision, mode=mode, return_x=False, profile_every=profile_every)
            rep["wall_ms"] = 1000.0 * (time.perf_counter() - t)
            out.append(rep)
            log("%-14s info=%d iters=%d outer=%d res_scaled=%.3e res_true=%.3e wall=%.1f ms krylov=%.1f ms" % (
                name, rep["info"], rep["iters"], rep["outer_iters"], rep["res_scaled"], rep["res_true"],
                rep["wall_ms"], rep["t_krylov_ms"]))
        return out

    # FP64 SpMV (tuned) and the FP64 forms
    sys_.spmv_tune(0)
    spmv64_us = sys_.spmv_bench(0, reps=50)
    if not args.skip_forms:
        for name, mode in (("fp64_faithful", _lib.FAITHFUL), ("fp64_fast", _lib.FAST)):
            reps = run(name, _lib.FP64, mode, 2)
            r = reps[-1]
            forms[name] = {"iters": r["iters"], "ms": r["wall_ms"], "krylov_ms": r["t_krylov_ms"],
                           "res_scaled": r["res_scaled"], "res_true": r["res_true"], "info": r["info"],
                           "dof_iters_per_s": n * r["iters"] / (r["wall_ms"] / 1000.0)}
    # headline: mixed precision refined CG
    run("mixed_ir(warm)", _lib.MIXED_IR, _lib.FAST, 1)
    sys_.spmv_tune(1)
    spmv32_us = sys_.spmv_bench(1, reps=50)
    run("mixed_ir(warm)", _lib.MIXED_IR, _lib.FAST, max(0, args.warmup - 1))
    steps = run("mixed_ir", _lib.MIXED_IR, _lib.FAST, args.steps, profile_every=4)
    wall = np.array([r["wall_ms"] for r in steps])
    iters = np.array([r["iters"] for r in steps])
    ms_per_step = float(wall.mean())
    value = float(n * iters.mean() / (ms_per_step / 1000.0))
    last = steps[-1]
    forms["mixed_ir"] = {"iters": int(last["iters"]), "outer": int(last["outer_iters"]), "ms": ms_per_step,
                         "res_scaled": last["res_scaled"], "res_true": last["res_true"], "info": last["info"],
                         "dof_iters_per_s": value}
    if "fp64_faithful" in forms:
        forms["mixed_ir"]["speedup_vs_fp64_faithful"] = forms["fp64_faithful"]["ms"] / ms_per_step
        forms["mixed_ir"]["speedup_vs_fp64_fast"] = forms["fp64_fast"]["ms"] / ms_per_step

    # roofline of the dominant kernel: the fused FP32 SpMV of the inner CG,
    # its launches sampled with HIP events on the system's stream in the timed solves
    us = float(np.mean([r["spmv_us"] for r in steps if r["spmv_samples"] > 0]))
    bytes32 = fused_cg32_bytes(nnz, n)
    achieved = bytes32 / (us * 1e-6) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "kernel": "k_spmv_epi<float,256,2048,EpiCgFast1<float>> (fused FP32 SpMV + CG update)",
                "bytes_per_launch": bytes32, "mean_us": us,
                "samples": int(sum(r["spmv_samples"] for r in steps))}
    spmv = {"fp64_us": spmv64_us, "fp64_gbs": spmv_bytes(nnz, n) / (spmv64_us * 1e-6) / 1e9,
            "fp32_us": spmv32_us, "fp32_gbs": spmv_bytes(nnz, n, True) / (spmv32_us * 1e-6) / 1e9,
            "copy_gbs_measured": copy_gbs, "config64": sys_.spmv_config()}
    spmv["fp64_frac_peak"] = spmv["fp64_gbs"] / HBM_PEAK_GBS
    assembly = {"ms": t_asm_ms, "elements_per_s": m.nelem / (t_asm_ms / 1000.0),
                "gbs_algorithmic": asm_bytes / (t_asm_ms * 1e-3) / 1e9, "host_setup_s": t_setup}
    cpu = None
    if not args.skip_cpu:
        log("cpu baseline: %d oracle CG iterations ..." % args.cpu_iters)
        cpu = cpu_baseline(sys_, kw, args.cpu_iters)
        log("cpu baseline %.3e DOF-iter/s (%.1fs)" % (cpu["value"], cpu["seconds"]))
    rec = {
        "metric": "DOF-iters/s (mixed-precision refined CG to 1e-10) + SpMV HBM GB/s",
        "value": value, "unit": "DOF-iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement", "data": "synthetic",
        "config": {"workload": "poisson_hex8_%d_lognormal" % args.n, "elements": m.nelem, "dofs": n, "nnz": nnz,
                   "sigma": args.sigma, "seed": args.seed, "tol": 1e-10, "method": "cg", "precond": "diagonal"},
        "roofline": roofline, "cpu_baseline": cpu, "forms": forms, "spmv": spmv, "assembly": assembly,
    }
    print(json.dumps(rec), flush=True)
    sys_.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
