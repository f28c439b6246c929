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

Prints one JSON line.  value = DOFs x inner iterations / solve time.  After
the headline (outside its timed steps) a short run of the batching form (64
convection-diffusion systems) and the serving form at its full size (64
meshes, 1000 requests) add `batch_summary` and `serve_summary` to the same
line (`--no-summaries` leaves them out; `--form batch` runs the batching
form at its full size, 512 systems).
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


def index_bytes(nnz: int, bits: int, fmt: str = "csr", units: int = 0) -> int:
    """Column index bytes one SpMV streams (padding not counted): int32
    indices, or 16-bit offsets plus one int32 base per segment of 2048
    nonzeros (CSR stream) or per slice (SELL-64, which also reads one offset
    per slice); slice-shared offsets: one int32 offset per 64 nonzeros (the
    slice's list, read once per wave) and the slice bounds."""
    if fmt == "sell_shared":
        return 4 * ((nnz + 63) // 64) + 8 * units
    if fmt == "sell":
        return (4 if bits == 32 else 2) * nnz + (4 if bits == 32 else 8) * units
    return 4 * nnz if bits == 32 else 2 * nnz + 4 * ((nnz + 2047) // 2048 + 2)


def row_ptr_bytes(n: int, fp32: bool, fmt: str) -> int:
    """Row pointers, where the kernel reads them: the FP32 SELL-64 kernels
    sum the zero padding instead of reading row lengths; the slice-shared
    layout reads a 4-byte entry mask per row in FP64 and nothing in FP32."""
    if fmt == "sell_shared":
        return 0 if fp32 else 4 * n
    return 0 if (fp32 and fmt == "sell") else 4 * (n + 1)


def spmv_bytes(nnz: int, n: int, fp32: bool = False, bits: int = 32, fmt: str = "csr", units: int = 0) -> int:
    """Algorithmic bytes of one SpMV: values, column indices, row pointers
    (where read), x read once, y written once."""
    vb = 4 if fp32 else 8
    return vb * nnz + index_bytes(nnz, bits, fmt, units) + row_ptr_bytes(n, fp32, fmt) + 2 * vb * n


def fused_cg32_bytes(nnz: int, n: int, bits: int = 32, fmt: str = "csr", units: int = 0) -> int:
    """Fused FP32 CG step 1 (q = A z + beta q, p = z + beta p, p.q):
    vals32 + column indices + row pointers (where read) + z + q (r/w) + p (r/w)."""
    return 4 * nnz + index_bytes(nnz, bits, fmt, units) + row_ptr_bytes(n, True, fmt) + 4 * n + 16 * n


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by a 1 GiB copy kernel's
    reading), or None when the summary is for a different kernel."""
    path = os.path.join(HERE, "profiles", "hbm_traffic_poisson_hex8_128.json")
    try:
        rec = json.load(open(path))
    except (OSError, ValueError):
        return None
    norm = lambda k: k.replace(" ", "").replace("ehip::", "")
    return float(rec["traffic_bytes"]) if norm(rec.get("kernel", "")) == norm(kernel) else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Time of the oracle's HUTI CG over the reference's own HUTI CG + BLAS
# compiled from its sources, same system and threads: python
# oracle/time_ratio.py --n 64 --threads 8 in the build container (8-core
# Xeon @ 2.70 GHz) printed 2.078 s / 1.904 s.
PORT_VS_REFERENCE = 1.091


def cpu_baseline(sys_, kw, iters: int, asm_n: int):
    """The reference CPU path restated (the oracle: HUTI CG with
    reference-BLAS ddot/dnrm2 and CRS_MatrixVectorProd's row-parallel SpMV)
    on all host threads: the full CG + Jacobi solve of the same scaled
    system to 1e-10 (or `iters` iterations when > 0), and the serial element
    loop + glue of an asm_n^3 sample of the same mesh family."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import oracle_lib as O
    from elmerfem_amd import workloads as W
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    threads = O.lib().or_set_threads(threads)
    perm, rows, cols, diag = sys_.pattern()
    vals, rhs = sys_.values()

    class P:
        pass
    pat = P()
    pat.n, pat.rows, pat.cols, pat.diag = sys_.n, rows, cols, diag
    v, b, x0 = O.scaled_system(pat, vals, rhs)
    t0 = time.perf_counter()
    tol, maxit = (1e-300, iters) if iters > 0 else (1e-10, 20000)
    info, k, x, hist = O.huti_solve(rows, cols, diag, v, b, x0, method=0, precond=1, tol=tol, maxit=maxit)
    dt = time.perf_counter() - t0
    done = min(k, maxit)
    # CPU assembly on a smaller mesh of the same family
    w = W.poisson_hex8(asm_n, sigma=1.0, seed=20261202)
    m = w["mesh"]
    apat = O.Pattern(m.elements, m.nnodes, 1, optimize=True, extra_elements=[m.bnd_elements])
    t1 = time.perf_counter()
    O.assemble(0, m, apat, w["coef"], load=w["load"])
    ta = time.perf_counter() - t1
    what = ("full HUTI-faithful CG + Jacobi solve to 1e-10 (%d iterations, info %d)" % (done, info) if iters <= 0
            else "%d HUTI-faithful CG iterations" % done)
    return {"value": sys_.n * done / dt, "unit": "DOF-iter/s", "cores": threads, "kind": "port",
            "sample": "%s of the same scaled system, 2 SpMV (row-parallel) + reference-BLAS ddot/dnrm2 per "
                      "iteration, oracle restatement on %d host threads, %.1f s; port/reference time ratio %.3f "
                      "(oracle/time_ratio.py, 8 threads)" % (what, threads, dt, PORT_VS_REFERENCE),
            "seconds": dt, "iters": done, "port_vs_reference": PORT_VS_REFERENCE,
            "assembly": {"elements_per_s": m.nelem / ta, "seconds": ta, "threads": 1,
                         "sample": "serial element loop + CRS glue (ModelPDE Hex8, %d^3 elements)" % asm_n}}


def form_batch(args):
    rec = run_batch(args)
    print(json.dumps(rec), flush=True)
    return 0


def run_batch(args):
    """Batching form: independent steady 2-D convection-diffusion problems on
    structured Quad4 grids, side N ~ U{32..512} elements, velocity of random
    direction with Peclet ~ U[1, 100], BiCGStab + Jacobi to 1e-10, all in one
    graph-replayed multi-system launch sequence -- FP64 in HUTI's order, FP64
    with the recurrence residual and true-residual restarts, then mixed
    precision -- beside the same systems solved one at a time in FP64."""
    from elmerfem_amd import _lib, workloads as W
    from elmerfem_amd.batch import schedule, solve_batch
    from elmerfem_amd.mesh import uniforms
    _lib.init(0)
    k = args.systems
    u, _ = uniforms(args.seed, np.arange(1, k + 1, dtype=np.uint64), stream=11)
    sizes = (32 + np.minimum((u * 481).astype(int), 480)).tolist()
    t0 = time.perf_counter()
    systems = []
    for i, nn in enumerate(sizes):
        w = W.convdiff_quad4(nn, seed=args.seed, index=i + 1)
        systems.append(W.build_system(w, optimize_bandwidth=True)[0])
        if (i + 1) % 64 == 0:
            log("built %d/%d systems (%.1fs)" % (i + 1, k, time.perf_counter() - t0))
    dofs = np.array([s.n for s in systems], dtype=np.float64)
    nnz = int(sum(s.nnz for s in systems))
    kw = W._kw("bicgstab", "diagonal", 1e-10, 20000)
    out = {}
    sched = {}
    # FP64 in HUTI's loop order (its S-norm exit included), FP64 with the
    # recurrence residual and true-residual restarts, and mixed precision.
    # The warm-up solve of each form runs the schedule search (chunk size
    # per size class, then workgroup count) unless --no-tune; the timed
    # solve runs under the table it left.
    for name, prec, mode in (("fp64_faithful_batched", _lib.FP64, _lib.FAITHFUL),
                             ("fp64_fast_batched", _lib.FP64, _lib.FAST),
                             ("mixed_batched", _lib.MIXED_IR, _lib.FAST)):
        t = time.perf_counter()
        solve_batch(systems, kw, precision=prec, mode=mode, true_residual_stop=True, tune=0 if args.no_tune else 1)
        warm = time.perf_counter() - t
        t = time.perf_counter()
        xs, reps, ms = solve_batch(systems, kw, precision=prec, mode=mode, true_residual_stop=True)
        wall = time.perf_counter() - t
        it = np.array([r["iters"] for r in reps], dtype=np.float64)
        conv = sum(1 for r in reps if r["info"] == 1)
        out[name] = {"wall_s": wall, "device_ms": ms, "dof_iters_per_s": float((dofs * it).sum() / wall),
                     "iters_mean": float(it.mean()), "iters_max": int(it.max()), "converged": conv,
                     "res_true_max": float(max(r["res_true"] for r in reps)),
                     "warmup_with_search_s": warm}
        if mode == _lib.FAST and prec == _lib.FP64:
            out[name]["restarts_mean"] = float(np.mean([r["outer_iters"] for r in reps]))
        classes = sorted({int(s.nnz).bit_length() for s in systems})
        sched[name] = {str(c): schedule(kw, 1 << (c - 1), precision=prec, mode=mode) for c in classes}
        log("%-14s %.2f s, %d/%d converged, %.3e DOF-iter/s" % (name, wall, conv, k, out[name]["dof_iters_per_s"]))
    for name, mode in (("fp64_faithful_one_at_a_time", _lib.FAITHFUL), ("fp64_fast_one_at_a_time", _lib.FAST)):
        t = time.perf_counter()
        it_seq, conv = [], 0
        for s in systems:
            _, r = s.solve(kw, precision=_lib.FP64, mode=mode, return_x=False, true_residual_stop=True)
            it_seq.append(r["iters"])
            conv += r["info"] == 1
        wall = time.perf_counter() - t
        out[name] = {"wall_s": wall, "dof_iters_per_s": float((dofs * np.array(it_seq)).sum() / wall),
                     "converged": conv}
        log("%s %.2f s, %d/%d converged" % (name, wall, conv, k))
    head = out["mixed_batched"]
    rec = {"metric": "DOF-iters/s, batched BiCGStab to 1e-10 (mixed precision)", "value": head["dof_iters_per_s"],
           "unit": "DOF-iter/s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": 1000.0 * head["wall_s"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 inner / f64 refinement",
           "data": "synthetic",
           "config": {"workload": "convdiff_quad4_batch", "systems": k, "dofs_total": int(dofs.sum()),
                      "nnz_total": nnz, "side_min": int(min(sizes)), "side_max": int(max(sizes)),
                      "seed": args.seed, "method": "bicgstab", "precond": "diagonal", "tol": 1e-10},
           "forms": out,
           "schedule": {"note": "chunk nonzeros and workgroups per form, keyed by the bit length of the "
                                "systems' nonzero counts", "chosen": sched},
           "note": "HUTI-order FP64 BiCGStab stops with HUTI_BICGSTAB_SNORM (36) on some convection-dominated "
                   "systems, as the reference HUTI does on the same matrices (checked against the compiled "
                   "reference); the fast forms confirm convergence on the true residual and restart",
           "speedup_vs_fp64_fast_one_at_a_time": {n: out["fp64_fast_one_at_a_time"]["wall_s"] / out[n]["wall_s"]
                                                  for n in ("fp64_faithful_batched", "fp64_fast_batched",
                                                            "mixed_batched")}}
    for s in systems:
        s.free()
    return rec


def form_serve(args):
    rec = run_serve(args)
    print(json.dumps(rec), flush=True)
    return 0


def run_serve(args):
    """Serving form: Poisson arrivals of parameter-sweep requests -- steady
    3-D heat on jittered Tet4 cubes (n^3 hexes x 6, n log-uniform in
    [10, 100]) with a per-element lognormal(0, 0.5) conductivity field drawn
    per request -- each assembled on the GPU from its coefficients (the mesh,
    its pattern and coordinates resident in HBM), Dirichlet rows applied,
    and solved by mixed-precision refined CG to 1e-10 on the continuous-
    batching server with HBM-footprint admission; latency is arrival to
    report on the host.  Meshes come from a pool built before the loop
    (`--pool`, two resident systems each so a mesh can serve two requests at
    once); every request takes one of `--variants` coefficient fields of its
    mesh, generated beforehand in pinned host memory.  The reference rate is
    the same requests one at a time: assembly, Dirichlet rows and a plain
    FP64 CG solve, in HUTI's loop order and in the one-SpMV order."""
    from elmerfem_amd import _lib, workloads as W
    from elmerfem_amd.batch import Server
    from elmerfem_amd.mesh import uniforms
    from elmerfem_amd.system import pinned_empty
    _lib.init(0)
    P, V, REP = args.pool, max(1, args.variants), 2
    u, _ = uniforms(args.seed, np.arange(1, P + 1, dtype=np.uint64), stream=13)
    sizes = np.clip(np.floor(10.0 ** (1.0 + u)).astype(int), 10, 100).tolist()
    t0 = time.perf_counter()
    meshes, sysrep, coefs, bcs = [], [], [], []
    rng = np.random.default_rng(args.seed)
    for i, nn in enumerate(sizes):
        w = W.heat_tet4(nn, seed=args.seed, index=i + 1)
        m = w["mesh"]
        reps_i = []
        for _ in range(REP):
            sy = W.build_system(w, optimize_bandwidth=True)[0]
            reps_i.append(sy)
        # coefficient fields in pinned host memory: their uploads are DMA copies
        # that do not hold the request loop
        vs = []
        for v in range(V):
            buf = pinned_empty(m.nelem)
            buf[:] = np.exp(0.5 * rng.standard_normal(m.nelem))
            vs.append(buf)
        load = pinned_empty(m.nelem)
        load[:] = w["load"]
        meshes.append((w, load))
        sysrep.append(reps_i)
        coefs.append(vs)
        bcs.append(w["bc_nodes"])
        if (i + 1) % 16 == 0:
            log("built %d/%d pool meshes (%.1fs)" % (i + 1, P, time.perf_counter() - t0))
    kw = W._kw("cg", "diagonal", 1e-10, 20000)

    def prepare(sy, mi, v, wait):
        """Assembles request (mesh mi, field v) into system sy and applies its Dirichlet rows."""
        w, load = meshes[mi]
        sy.assemble(w["physics"], None, elem_coef=coefs[mi][v], elem_load=load, wait=wait)
        sy.dirichlet_nodes(bcs[mi])

    # every pool system is resident and warm in both precisions before any
    # timing (first-use work -- FP32 and SELL value copies, workspace, the
    # recorded iteration graphs -- happens once per system, not per request)
    for mi in range(P):
        for sy in sysrep[mi]:
            for prec, mode in ((_lib.FP64, _lib.FAITHFUL), (_lib.FP64, _lib.FAST), (_lib.MIXED_IR, _lib.FAST)):
                prepare(sy, mi, 0, True)
                sy.solve(kw, precision=prec, mode=mode, return_x=False, true_residual_stop=True)
    log("pool warm (%.1fs)" % (time.perf_counter() - t0))
    # one at a time: assembly + Dirichlet + plain FP64 CG, over one request per mesh
    seq = {}
    for name, mode in (("fp64_faithful", _lib.FAITHFUL), ("fp64_fast", _lib.FAST)):
        lat_s = []
        t = time.perf_counter()
        for mi in range(P):
            t1 = time.perf_counter()
            sy = sysrep[mi][0]
            prepare(sy, mi, mi % V, True)
            sy.solve(kw, precision=_lib.FP64, mode=mode, return_x=False, true_residual_stop=True)
            lat_s.append(time.perf_counter() - t1)
        seq[name] = {"solves_per_s": P / (time.perf_counter() - t),
                     "latency_ms_p50": 1000.0 * float(np.percentile(lat_s, 50)),
                     "latency_ms_p99": 1000.0 * float(np.percentile(lat_s, 99))}
        log("one-at-a-time %s: %.1f solves/s" % (name, seq[name]["solves_per_s"]))
    srv = Server(kw, precision=_lib.MIXED_IR, max_slots=args.slots, mem_budget=int(args.budget_gib * (1 << 30)),
                 check_every=args.check_every, true_residual_stop=True)
    if args.lane_workers > 0:
        srv.set_lane(args.lane_dofs, args.lane_workers)
    R = args.requests
    failures = []

    def serve(arrivals):
        """Runs the request stream; returns (elapsed s, latencies s, infos, res_true, max live)."""
        busy, queue, prepared = {}, [], {}
        free = [list(r) for r in sysrep]
        host = {"prepare_s": 0.0, "submit_s": 0.0, "step_s": 0.0}
        lat = np.zeros(R)
        infos = np.zeros(R, dtype=int)
        res_true = np.zeros(R)
        nxt = done = max_live = 0
        t0 = time.perf_counter()
        while done < R:
            now = time.perf_counter() - t0
            while nxt < R and arrivals[nxt] <= now:
                queue.append(nxt)
                nxt += 1
            # admit every waiting request that fits (a large one waiting for
            # the lane does not hold back small ones behind it); once a request
            # of one path is refused, later ones of that path are not tried
            still, full = [], set()
            for req in queue:
                mi = req % P
                sy0 = sysrep[mi][0]
                path = args.lane_workers > 0 and sy0.n >= args.lane_dofs
                if path in full or (req not in prepared and not free[mi]):
                    still.append(req)
                    continue
                if req not in prepared:
                    # assembled once, when it first gets a system; a refused
                    # submission keeps that system for the retry
                    sy = free[mi].pop()
                    tp = time.perf_counter()
                    prepare(sy, mi, (req // P) % V, False)
                    host["prepare_s"] += time.perf_counter() - tp
                    prepared[req] = sy
                sy = prepared[req]
                tp = time.perf_counter()
                ok = srv.submit(sy)
                host["submit_s"] += time.perf_counter() - tp
                if ok:
                    del prepared[req]
                    busy[sy._h.value] = (req, mi, sy)
                else:
                    full.add(path)
                    still.append(req)
            queue = still
            live, _ = srv.status()
            max_live = max(max_live, live)
            if live == 0:
                if nxt < R:
                    time.sleep(max(0.0, min(arrivals[nxt] - (time.perf_counter() - t0), 0.005)))
                continue
            tp = time.perf_counter()
            finished = srv.step()
            host["step_s"] += time.perf_counter() - tp
            for sy, rep in finished:
                req, mi, sy_ = busy.pop(sy._h.value)
                free[mi].append(sy_)
                lat[req] = time.perf_counter() - t0 - arrivals[req]
                infos[req] = rep["info"]
                res_true[req] = rep["res_true"]
                if rep["info"] != 1 and len(failures) < 20:
                    failures.append({"request": int(req), "mesh": int(mi), "dofs": int(sy.n),
                                     "info": int(rep["info"]), "iters": int(rep["iters"]),
                                     "outer": int(rep["outer_iters"]), "res_true": float(rep["res_true"]),
                                     "lane": bool(args.lane_workers > 0 and sy.n >= args.lane_dofs)})
                done += 1
        host_times.append(host)
        return time.perf_counter() - t0, lat, infos, res_true, max_live

    # capacity: every request waiting at t = 0
    host_times = []
    el, lat_sat, infos, rt, ml = serve(np.zeros(R))
    capacity = R / el
    ref = seq["fp64_faithful"]["solves_per_s"]
    log("server capacity %.1f solves/s (%.1fx one-at-a-time FP64 HUTI order, %.1fx one-SpMV order)" % (
        capacity, capacity / ref, capacity / seq["fp64_fast"]["solves_per_s"]))
    # latency under load: Poisson arrivals at a fixed fraction of that capacity
    rate = args.rate if args.rate > 0 else args.load * capacity
    arrivals = np.cumsum(np.random.default_rng(args.seed).exponential(1.0 / rate, R))
    el2, lat, infos2, rt2, max_live = serve(arrivals)
    srv.free()
    log("at %.1f arrivals/s: p50 %.2f ms, p99 %.2f ms" % (rate, 1000 * np.percentile(lat, 50),
                                                          1000 * np.percentile(lat, 99)))
    rec = {"metric": "solves/s (capacity) and p50/p99 latency at a fixed arrival rate, continuous-batching "
                     "mixed-precision CG to 1e-10 with per-request GPU assembly",
           "value": capacity, "unit": "solves/s", "n_gpus": 1, "steps": R, "warmup": 0,
           "ms_per_step": 1000.0 * el / R, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 inner / f64 refinement", "data": "synthetic",
           "config": {"workload": "heat_tet4_serving_sweep", "requests": R, "pool_meshes": P,
                      "systems_per_mesh": REP, "coefficient_fields_per_mesh": V,
                      "n_min": int(min(sizes)), "n_max": int(max(sizes)),
                      "dofs_mean": float(np.mean([r[0].n for r in sysrep])), "slots": args.slots,
                      "lane_min_dofs": args.lane_dofs, "lane_workers": args.lane_workers,
                      "check_every": args.check_every,
                      "mem_budget_gib": args.budget_gib, "seed": args.seed, "tol": 1e-10},
           "arrival_rate_per_s": rate, "served_rate_under_load": R / el2,
           "latency_ms": {"p50": 1000.0 * float(np.percentile(lat, 50)),
                          "p99": 1000.0 * float(np.percentile(lat, 99)), "max": 1000.0 * float(lat.max())},
           "one_at_a_time": {**seq, "request": "GPU assembly + Dirichlet rows + FP64 CG solve"},
           "speedup_solves_per_s": capacity / ref,
           "speedup_vs_fp64_fast_one_at_a_time": capacity / seq["fp64_fast"]["solves_per_s"],
           "converged": int((infos == 1).sum() + (infos2 == 1).sum()), "requests_total": 2 * R,
           "failures": failures,
           "res_true_max": float(max(rt.max(), rt2.max())), "max_live": max_live,
           "host_loop_s": {"capacity_run": host_times[0], "load_run": host_times[1]}}
    for r in sysrep:
        for sy in r:
            sy.free()
    return rec


def form_ilu(args):
    """FP64 CG preconditioned with ILU(k) (--ilu-order) on the n^3 Hex8
    Poisson case (lognormal(0, sigma) conductivity), to 1e-10 in HUTI's loop
    order, against the same solve with Jacobi."""
    from elmerfem_amd import _lib, workloads as W
    _lib.init(0)
    w = W.poisson_hex8(args.n, sigma=args.sigma, seed=args.seed)
    sys_, _ = W.build_system(w, optimize_bandwidth=True)
    out = {}
    for name, pc in (("ilu%d" % args.ilu_order, "ilu%d" % args.ilu_order), ("diagonal", "diagonal")):
        kw = W._kw("cg", pc, 1e-10, 20000)
        for rep in range(1 + args.steps):
            t = time.perf_counter()
            _, r = sys_.solve(kw, precision=_lib.FP64, mode=_lib.FAITHFUL, return_x=False, true_residual_stop=True)
            wall = 1000.0 * (time.perf_counter() - t)
            log("%-9s info=%d iters=%d res_true=%.3e wall=%.1f ms" % (name, r["info"], r["iters"], r["res_true"], wall))
        out[name] = {"iters": r["iters"], "ms": wall, "info": r["info"], "res_true": r["res_true"],
                     "dof_iters_per_s": sys_.n * r["iters"] / (wall / 1000.0)}
    lf, lb = sys_.ilu_levels()
    out["ilu%d" % args.ilu_order].update({"levels_forward": lf, "levels_backward": lb})
    head = out["ilu%d" % args.ilu_order]
    rec = {"metric": "time to 1e-10, FP64 CG + ILU(%d) (HUTI order)" % args.ilu_order, "value": head["ms"],
           "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": head["ms"],
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "poisson_hex8_%d_lognormal" % args.n, "dofs": sys_.n, "nnz": sys_.nnz,
                      "sigma": args.sigma, "seed": args.seed, "ilu_order": args.ilu_order},
           "forms": out}
    print(json.dumps(rec), flush=True)
    sys_.free()
    return 0


def form_elastic(args):
    """Precision form on 3-D linear elasticity: structured Hex8 cube (or Hex27
    with --quadratic), E = 1, nu = 0.3, x = 0 clamped, traction (0, 0, -1e-3)
    on x = 1; CG + Jacobi to 1e-10 in FP64 and in mixed precision (FP32 inner
    CG with reliable FP64 residual replacement).  The mixed solution's error
    is stated against the FP64 solution of the same system."""
    from elmerfem_amd import _lib, workloads as W
    _lib.init(0)
    t0 = time.perf_counter()
    w = W.elasticity_hex(args.n, quadratic=args.quadratic)
    sys_, t_asm = W.build_system(w, optimize_bandwidth=True)
    log("elasticity %s %d^3: n=%d nnz=%d, built in %.1fs" % ("hex27" if args.quadratic else "hex8", args.n, sys_.n,
                                                              sys_.nnz, time.perf_counter() - t0))
    t_asm = float(np.median([sys_.assemble(w["physics"], w["mesh"].coords, elem_coef=w["coef"],
                                           poisson=w["poisson"]) for _ in range(3)]))
    m = w["mesh"]
    sel = m.bnd_ids == w["traction_bnd"]
    sys_.assemble_traction(m.bnd_code, m.coords, m.bnd_elements[sel], w["traction"])
    sys_.dirichlet_nodes(w["bc_nodes"])
    vals, rhs = sys_.values()
    kw = w["kw"]
    kw.tol, kw.maxit = 1e-10, 50000
    out = {}
    xs = {}
    for name, prec, mode in (("fp64_fast", _lib.FP64, _lib.FAST), ("fp64_faithful", _lib.FP64, _lib.FAITHFUL),
                             ("mixed", _lib.MIXED_IR, _lib.FAST)):
        sys_.set_values(vals, rhs)
        t = time.perf_counter()
        x, r = sys_.solve(kw, precision=prec, mode=mode, inner_tol=args.inner_tol if prec == _lib.MIXED_IR else 0.0,
                          true_residual_stop=True)
        wall = time.perf_counter() - t
        xs[name] = x
        out[name] = {"iters": r["iters"], "ms": 1000.0 * wall, "info": r["info"], "res_true": r["res_true"],
                     "dof_iters_per_s": sys_.n * r["iters"] / wall, "replacements": r["outer_iters"]}
        log("%-14s info=%d iters=%d %.1f ms res_true=%.2e" % (name, r["info"], r["iters"], 1000 * wall,
                                                             r["res_true"]))
    ref = xs["fp64_fast"]
    out["mixed"]["rel_err_vs_fp64"] = float(np.abs(xs["mixed"] - ref).max() / np.abs(ref).max())
    head = out["mixed"]
    rec = {"metric": "DOF-iters/s, mixed-precision refined CG to 1e-10 (elasticity)", "value": head["dof_iters_per_s"],
           "unit": "DOF-iter/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": head["ms"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 inner / f64 refinement",
           "data": "synthetic",
           "config": {"workload": "elasticity_%s_%d" % ("hex27" if args.quadratic else "hex8", args.n),
                      "dofs": sys_.n, "nnz": sys_.nnz, "E": 1.0, "nu": 0.3, "tol": 1e-10},
           "forms": out, "assembly_ms": t_asm,
           "speedup_vs_fp64_faithful": out["fp64_faithful"]["ms"] / head["ms"],
           "speedup_vs_fp64_fast": out["fp64_fast"]["ms"] / head["ms"]}
    print(json.dumps(rec), flush=True)
    sys_.free()
    return 0


def form_assembly(args):
    """Assembly form: GPU assembly (element blocks + deterministic gather into
    CRS Values and RHS) of the n^3 Hex8 Poisson case with lognormal
    conductivity, timed over `--steps` calls after `--warmup`; then the same
    number of calls that keep the coordinates on the device and only enqueue
    (the re-assembly a parameter sweep makes), timed on the host including the
    coefficient upload."""
    from elmerfem_amd import _lib, workloads as W
    _lib.init(0)
    w = W.poisson_hex8(args.n, sigma=args.sigma, seed=args.seed)
    m = w["mesh"]
    t0 = time.perf_counter()
    s, _ = W.build_system(w, optimize_bandwidth=True)
    setup = time.perf_counter() - t0
    for _ in range(args.warmup):
        s.assemble(w["physics"], m.coords, elem_coef=w["coef"], elem_load=w["load"])
    ts = [s.assemble(w["physics"], None, elem_coef=w["coef"], elem_load=w["load"]) for _ in range(args.steps)]
    t = time.perf_counter()
    for _ in range(args.steps):
        s.assemble(w["physics"], None, elem_coef=w["coef"], elem_load=w["load"], wait=False)
    s.values()
    host_ms = 1000.0 * (time.perf_counter() - t) / args.steps
    npe = m.elements.shape[1]
    asm_bytes = 24 * m.nnodes + 8 * m.nelem + 4 * npe * m.nelem + 4 * npe * npe * m.nelem + 8 * s.nnz + 8 * s.n
    ms = float(np.median(ts))
    rec = {"metric": "GPU assembly time (element blocks + gather into CRS)", "value": ms, "unit": "ms",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "poisson_hex8_%d_lognormal" % args.n, "elements": int(m.nelem),
                      "nodes": int(m.nnodes), "nnz": int(s.nnz), "seed": args.seed},
           "elements_per_s": m.nelem / (ms * 1e-3), "alg_bytes": asm_bytes,
           "gbs_algorithmic": asm_bytes / (ms * 1e-3) / 1e9, "reassembly_host_ms": host_ms,
           "host_setup_s": setup}
    print(json.dumps(rec), flush=True)
    s.free()
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--form", choices=["headline", "batch", "serve", "elastic", "ilu", "assembly"], default="headline",
                    help="headline: the 128^3 solve; batch / serve: the batching and serving forms; "
                         "elastic: the precision form on elasticity (--n 96, or --quadratic --n 48)")
    ap.add_argument("--quadratic", action="store_true", help="elastic form: Hex27 elements")
    ap.add_argument("--ilu-order", type=int, default=0, help="ilu form: fill order k of ILU(k)")
    ap.add_argument("--inner-tol", type=float, default=0.0,
                    help="mixed CG: residual drop between FP64 replacements (0: the library default)")
    ap.add_argument("--systems", type=int, default=512, help="batch form: number of systems")
    ap.add_argument("--no-tune", action="store_true", help="batch form: no schedule search (default table)")
    ap.add_argument("--pool", type=int, default=64, help="serve form: distinct request meshes")
    ap.add_argument("--variants", type=int, default=4, help="serve form: coefficient fields per mesh")
    ap.add_argument("--requests", type=int, default=1000, help="serve form: requests")
    ap.add_argument("--rate", type=float, default=0.0, help="serve form: arrivals per second (0: load x capacity)")
    ap.add_argument("--load", type=float, default=0.7, help="serve form: arrival rate / measured capacity")
    ap.add_argument("--lane-dofs", type=int, default=100000, help="serve form: unknowns from which a request "
                                                                   "takes the concurrent lane")
    ap.add_argument("--lane-workers", type=int, default=4, help="serve form: concurrent-lane workers (0: off)")
    ap.add_argument("--slots", type=int, default=64, help="serve form: server slots")
    ap.add_argument("--check-every", type=int, default=0,
                    help="serve form: iterations per engine replay (0: the library's 8)")
    ap.add_argument("--budget-gib", type=float, default=64.0, help="serve form: HBM admission budget")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--sigma", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-iters", type=int, default=0, help="CPU baseline: 0 = full solve, else iterations")
    ap.add_argument("--cpu-asm-n", type=int, default=64, help="CPU baseline: assembly sample mesh (n^3)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-forms", action="store_true", help="time only the headline form")
    ap.add_argument("--no-summaries", action="store_true",
                    help="headline: no short batching and serving runs after it")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    if args.form == "batch":
        return form_batch(args)
    if args.form == "serve":
        return form_serve(args)
    if args.form == "ilu":
        return form_ilu(args)
    if args.form == "assembly":
        return form_assembly(args)
    if args.form == "elastic":
        if args.n == 128:
            args.n = 48 if args.quadratic else 96
        return form_elastic(args)
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

    sols = {}

    def run(name, precision, mode, reps, profile_every=0, keep_x=False):
        """Solves to 1e-10 on both the scaled (HUTI) and the unscaled FP64
        residual; keep_x keeps the last solution (its copy to the host is
        outside the timed region)."""
        out = []
        for _ in range(reps):
            t = time.perf_counter()
            x, rep = sys_.solve(kw, precision=precision, mode=mode, return_x=False, profile_every=profile_every,
                                inner_tol=args.inner_tol if precision == _lib.MIXED_IR else 0.0,
                                true_residual_stop=True)
            rep["wall_ms"] = 1000.0 * (time.perf_counter() - t)
            out.append(rep)
            if keep_x:
                sols[name] = sys_.solution()
            log("%-14s info=%d iters=%d outer=%d res_scaled=%.3e res_true=%.3e wall=%.1f ms krylov=%.1f ms" % (
                name, rep["info"], rep["iters"], rep["outer_iters"], rep["res_scaled"], rep["res_true"],
                rep["wall_ms"], rep["t_krylov_ms"]))
        return out

    # FP64 SpMV (tuned) and the FP64 forms
    sys_.spmv_tune(0)
    spmv64_us = sys_.spmv_bench(0, reps=50)
    if not args.skip_forms:
        for name, mode in (("fp64_faithful", _lib.FAITHFUL), ("fp64_fast", _lib.FAST)):
            reps = run(name, _lib.FP64, mode, 2, keep_x=True)
            r = reps[-1]
            forms[name] = {"iters": r["iters"], "ms": r["wall_ms"], "krylov_ms": r["t_krylov_ms"],
                           "res_scaled": r["res_scaled"], "res_true": r["res_true"], "info": r["info"],
                           "dof_iters_per_s": n * r["iters"] / (r["wall_ms"] / 1000.0)}
    # headline: mixed precision refined CG
    run("mixed_ir(warm)", _lib.MIXED_IR, _lib.FAST, 1)
    sys_.spmv_tune(1)
    spmv32_us = sys_.spmv_bench(1, reps=50)
    run("mixed_ir(warm)", _lib.MIXED_IR, _lib.FAST, max(0, args.warmup - 1))
    # timed steps run as a user would (graph-replayed iteration blocks); one
    # more solve samples every 4th fused SpMV launch with kernel-bound events
    steps = run("mixed_ir", _lib.MIXED_IR, _lib.FAST, args.steps, keep_x=True)
    sampled = run("mixed_ir(sampled)", _lib.MIXED_IR, _lib.FAST, 1, profile_every=4)
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
        # error of each solution against the FP64 (HUTI loop order) solution
        ref = sols["fp64_faithful"]
        for name in ("fp64_fast", "mixed_ir"):
            forms[name]["rel_err_vs_fp64"] = float(np.abs(sols[name] - ref).max() / np.abs(ref).max())

    # FP64 CG preconditioned with ILU(0) (HUTI order, level-scheduled LU solves)
    if not args.skip_forms:
        kw_ilu = W._kw("cg", "ilu0", 1e-10, 20000)
        reps = []
        for _ in range(2):
            t = time.perf_counter()
            _, r = sys_.solve(kw_ilu, precision=_lib.FP64, mode=_lib.FAITHFUL, return_x=False,
                              true_residual_stop=True)
            r["wall_ms"] = 1000.0 * (time.perf_counter() - t)
            reps.append(r)
            log("%-14s info=%d iters=%d res_true=%.3e wall=%.1f ms krylov=%.1f ms" % (
                "fp64_ilu0", r["info"], r["iters"], r["res_true"], r["wall_ms"], r["t_krylov_ms"]))
        r = reps[-1]
        lev = sys_.ilu_levels()
        forms["fp64_ilu0"] = {"iters": r["iters"], "ms": r["wall_ms"], "krylov_ms": r["t_krylov_ms"],
                              "res_scaled": r["res_scaled"], "res_true": r["res_true"], "info": r["info"],
                              "dof_iters_per_s": n * r["iters"] / (r["wall_ms"] / 1000.0),
                              "levels_forward": lev[0], "levels_backward": lev[1]}

    # roofline of the dominant kernel: the fused FP32 SpMV of the inner CG,
    # its launches sampled with HIP events on the system's stream in the timed solves
    c32 = sys_.spmv_config(1)
    bits = c32["index_bits"]
    # the SpMV step's epilogue: folded (the previous vector step's partials summed
    # in every workgroup, two launches per iteration) unless EHIP_CG_FOLD=0
    epi = "EpiCgFast1<float>" if os.environ.get("EHIP_CG_FOLD", "1")[:1] == "0" else "EpiCgFold<float>"
    if c32["format"] == "sell_shared":
        lay = sys_.spmv_shared_layout()
        # values streamed non-temporally only when the FP32 copy exceeds the 256 MiB Infinity Cache
        nt = "true" if 4 * lay["entries"] > (256 << 20) else "false"
        kname = "k_spmv_epi_sho<float,256,%d,%d,%s,%s>" % (lay["lane_unroll"], max(1, lay["node_block"]), nt, epi)
    elif c32["format"] == "sell":
        lay = sys_.spmv_sell_layout()
        if lay["node_block"]:
            kname = "k_spmv_epi_sellb<float,256,%d,%d,%s>" % (lay["lane_unroll"], lay["node_block"], epi)
        else:
            kname = "k_spmv_epi_sell<float,256,%d,%s,%s>" % (lay["lane_unroll"], "true" if bits == 16 else "false",
                                                            epi)
    else:
        kname = "k_spmv_epi<float,%d,%d,%s,%s>" % (c32["block"], c32["nnz_per_block"],
                                                  "ColsU16" if bits == 16 else "ColsI32", epi)
    us = float(np.mean([r["spmv_us"] for r in sampled if r["spmv_samples"] > 0]))
    # algorithmic bytes of one fused FP32 CG step: FP32 values, int32 column
    # indices, row pointers, z read, q and p read and written
    alg32 = 8 * nnz + 4 * (n + 1) + 20 * n
    # what this layout reads for the same work (16-bit column offsets or slice-shared
    # offsets, no row pointers)
    bytes32 = fused_cg32_bytes(nnz, n, bits, c32["format"], c32["chunks"])
    achieved = alg32 / (us * 1e-6) / 1e9
    traffic = measured_traffic(kname)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": kname + " (fused FP32 SpMV + CG update)",
                "bytes_per_launch": alg32, "mean_us": us,
                "samples": int(sum(r["spmv_samples"] for r in sampled)),
                "layout_bytes_per_launch": bytes32, "layout_gbs": bytes32 / (us * 1e-6) / 1e9,
                "physical_gbs": traffic / (us * 1e-6) / 1e9 if traffic else None,
                "physical_frac": traffic / (us * 1e-6) / 1e9 / HBM_PEAK_GBS if traffic else None,
                "traffic_source": "profiles/hbm_traffic_poisson_hex8_128.json" if traffic else None}
    c64 = sys_.spmv_config(0)
    b64 = spmv_bytes(nnz, n, False, c64["index_bits"], c64["format"], c64["chunks"])
    b32 = spmv_bytes(nnz, n, True, bits, c32["format"], c32["chunks"])
    alg64 = 12 * nnz + 4 * (n + 1) + 16 * n   # FP64 values, int32 indices, row pointers, x, y
    spmv = {"fp64_us": spmv64_us, "fp64_alg_bytes": alg64, "fp64_gbs": alg64 / (spmv64_us * 1e-6) / 1e9,
            "fp64_layout_bytes": b64, "fp64_layout_gbs": b64 / (spmv64_us * 1e-6) / 1e9,
            "fp32_us": spmv32_us, "fp32_alg_bytes": 8 * nnz + 4 * (n + 1) + 8 * n,
            "fp32_gbs": (8 * nnz + 4 * (n + 1) + 8 * n) / (spmv32_us * 1e-6) / 1e9,
            "fp32_layout_bytes": b32, "fp32_layout_gbs": b32 / (spmv32_us * 1e-6) / 1e9,
            "index_bits": bits, "copy_gbs_measured": copy_gbs,
            "config64": sys_.spmv_config(0), "config32": sys_.spmv_config(1)}
    spmv["fp64_frac_peak"] = spmv["fp64_gbs"] / HBM_PEAK_GBS
    spmv["fp64_layout_frac_peak"] = spmv["fp64_layout_gbs"] / HBM_PEAK_GBS
    assembly = {"ms": t_asm_ms, "elements_per_s": m.nelem / (t_asm_ms / 1000.0),
                "gbs_algorithmic": asm_bytes / (t_asm_ms * 1e-3) / 1e9, "host_setup_s": t_setup}
    cpu = None
    if not args.skip_cpu:
        log("cpu baseline: oracle CG (%s) ..." % ("full solve" if args.cpu_iters <= 0 else
                                                     "%d iterations" % args.cpu_iters))
        cpu = cpu_baseline(sys_, kw, args.cpu_iters, args.cpu_asm_n)
        log("cpu baseline %.3e DOF-iter/s on %d threads (%.1fs), assembly %.3e elements/s" % (
            cpu["value"], cpu["cores"], cpu["seconds"], cpu["assembly"]["elements_per_s"]))
    rec = {
        "metric": "DOF-iters/s (mixed-precision refined CG to 1e-10) + SpMV HBM GB/s",
        "value": value, "unit": "DOF-iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement", "data": "synthetic",
        "config": {"workload": "poisson_hex8_%d_lognormal" % args.n, "elements": m.nelem, "dofs": n, "nnz": nnz,
                   "sigma": args.sigma, "seed": args.seed, "tol": 1e-10, "method": "cg", "precond": "diagonal"},
        "roofline": roofline, "cpu_baseline": cpu, "forms": forms, "spmv": spmv, "assembly": assembly,
    }
    sys_.free()
    if not args.skip_forms and not args.no_summaries:
        # a short run of the batching form (--form batch runs it at full
        # size) and the serving form at the size of --form serve
        import copy
        a = copy.copy(args)
        a.systems = 64
        log("batch summary (%d systems) ..." % a.systems)
        b = run_batch(a)
        f = b["forms"]
        rec["batch_summary"] = {
            "config": b["config"], "metric": b["metric"], "value": b["value"], "unit": b["unit"],
            "wall_s": {k: f[k]["wall_s"] for k in f}, "converged": {k: f[k]["converged"] for k in f},
            "res_true_max": {k: f[k]["res_true_max"] for k in f if "res_true_max" in f[k]},
            "speedup_vs_fp64_fast_one_at_a_time": b["speedup_vs_fp64_fast_one_at_a_time"]}
        log("serve summary (%d meshes, %d requests) ..." % (a.pool, a.requests))
        v = run_serve(a)
        rec["serve_summary"] = {
            "config": v["config"], "metric": v["metric"], "capacity_solves_per_s": v["value"],
            "latency_ms": v["latency_ms"], "arrival_rate_per_s": v["arrival_rate_per_s"],
            "one_at_a_time": v["one_at_a_time"], "speedup_vs_fp64_faithful": v["speedup_solves_per_s"],
            "speedup_vs_fp64_fast": v["speedup_vs_fp64_fast_one_at_a_time"],
            "converged": v["converged"], "requests_total": v["requests_total"], "res_true_max": v["res_true_max"]}
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
