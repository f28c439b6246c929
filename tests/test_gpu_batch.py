"""Batched and served solves against the oracle and against themselves alone.

Every system in a batch keeps its own scalars, stopping test and chunk-order
reductions, so a system's result must not depend on its companions: the
same system alone and in a mix gives bit-identical solutions.
"""
import numpy as np
import pytest

import oracle_lib as O
from elmerfem_amd import _lib, workloads as W
from elmerfem_amd.batch import Server, solve_batch
from elmerfem_amd.system import System

pytestmark = pytest.mark.gpu


def oracle_system(w):
    m = w["mesh"]
    pat = O.Pattern(m.elements, m.nnodes, w["dofs"], optimize=True, extra_elements=[m.bnd_elements])
    vals, rhs = O.assemble(w["physics"], m, pat, w["coef"], conv=w.get("conv"), velo=w.get("velo"),
                           load=w.get("load"), poisson=w.get("poisson", 0.3))
    O.dirichlet(pat, vals, rhs, pat.perm[w["bc_nodes"] - 1])
    return pat, vals, rhs


def oracle_solution(w, method, tol=1e-12):
    pat, vals, rhs = oracle_system(w)
    info, k, x, hist, nrm = O.solve_linear_system(pat, vals.copy(), rhs, method=method, tol=tol, maxit=20000)
    assert info == 1
    return x


def host_residual(w, x):
    """||b - A x|| / ||b|| in FP64 on the host (the oracle's unscaled CRS)."""
    pat, vals, rhs = oracle_system(w)
    ax = np.zeros(pat.n)
    O.lib().or_spmv(pat.n, O._p(pat.rows), O._p(pat.cols), O._p(vals), O._p(np.ascontiguousarray(x)), O._p(ax))
    return np.linalg.norm(rhs - ax) / np.linalg.norm(rhs)


def make(kind, i):
    if kind == "convdiff":
        return W.convdiff_quad4(8 + 5 * i, index=i + 1)
    return W.heat_tet4(3 + i, index=i + 1)


@pytest.mark.parametrize("kind,method", [("convdiff", _lib.BICGSTAB), ("tet4", _lib.CG)])
@pytest.mark.parametrize("precision", [_lib.FP64, _lib.MIXED_IR])
def test_batch_matches_oracle(ehip, kind, method, precision):
    ws = [make(kind, i) for i in range(5)]
    systems = [W.build_system(w, optimize_bandwidth=True)[0] for w in ws]
    kw = ws[0]["kw"]
    kw.method, kw.tol, kw.maxit = method, 1e-10, 20000
    mode = _lib.FAITHFUL if precision == _lib.FP64 else _lib.FAST
    xs, reps, ms = solve_batch(systems, kw, precision=precision, mode=mode, true_residual_stop=True)
    for w, x, r in zip(ws, xs, reps):
        assert r["info"] == 1, r
        assert r["res_scaled"] < 1e-10 and r["res_true"] <= 1e-10
        assert host_residual(w, x) <= 1e-10
        xo = oracle_solution(w, method)
        assert np.abs(x - xo).max() <= 1e-8 * np.abs(xo).max()


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.MIXED_IR])
def test_batch_faithful_iterations_match_single(ehip, precision):
    """Batched CG takes the single-system iteration count (FP64 in HUTI order:
    within rounding) and reaches the same solution."""
    ws = [W.heat_tet4(4 + i, index=i + 1) for i in range(3)]
    kw = ws[0]["kw"]
    singles = []
    for w in ws:
        s = W.build_system(w, optimize_bandwidth=True)[0]
        x, r = s.solve(kw, precision=precision, mode=_lib.FAITHFUL if precision == _lib.FP64 else _lib.FAST)
        singles.append((x, r))
    systems = [W.build_system(w, optimize_bandwidth=True)[0] for w in ws]
    xs, reps, _ = solve_batch(systems, kw, precision=precision,
                              mode=_lib.FAITHFUL if precision == _lib.FP64 else _lib.FAST)
    for (xs1, r1), x, r in zip(singles, xs, reps):
        if precision == _lib.FP64:
            assert abs(r["iters"] - r1["iters"]) <= max(2, r1["iters"] // 10)
        else:
            # both engines run reliable-update CG with the same replacement
            # factor; their FP32 sums are grouped differently (per workgroup
            # vs per chunk), which moves the count by a few steps
            assert abs(r["iters"] - r1["iters"]) <= max(4, r1["iters"] // 4)
        assert np.abs(x - xs1).max() <= 1e-8 * np.abs(xs1).max()


def test_alone_equals_in_mix(ehip):
    ws = [make("tet4", i) for i in range(4)] + [make("tet4", 7)]
    kw = ws[0]["kw"]
    alone = []
    for w in ws:
        s = W.build_system(w, optimize_bandwidth=True)[0]
        xs, reps, _ = solve_batch([s], kw, precision=_lib.MIXED_IR, mode=_lib.FAST)
        alone.append((xs[0], reps[0]))
    systems = [W.build_system(w, optimize_bandwidth=True)[0] for w in ws]
    xs, reps, _ = solve_batch(systems, kw, precision=_lib.MIXED_IR, mode=_lib.FAST)
    for (xa, ra), x, r in zip(alone, xs, reps):
        assert np.array_equal(xa, x)
        assert ra["iters"] == r["iters"] and ra["outer_iters"] == r["outer_iters"]


@pytest.mark.parametrize("kind", ["tet4", "convdiff"])  # CG, BiCGStab
def test_refine_period_keeps_bits(ehip, monkeypatch, kind):
    """Mixed precision in the multi-system launches: the refinement ops run
    every k-th iteration (EHIP_BATCH_REFINE_EVERY); a slot whose inner solve
    ends in between only waits, so every k gives the same solutions, bit for
    bit, and the same inner and outer iteration counts."""
    ws = [make(kind, i) for i in range(5)]
    kw = ws[0]["kw"]
    runs = []
    for k in ("1", "3"):
        monkeypatch.setenv("EHIP_BATCH_REFINE_EVERY", k)
        systems = [W.build_system(w, optimize_bandwidth=True)[0] for w in ws]
        xs, reps, _ = solve_batch(systems, kw, precision=_lib.MIXED_IR, mode=_lib.FAST, true_residual_stop=True)
        assert all(r["info"] == 1 for r in reps)
        runs.append((xs, reps))
    (xa, ra), (xb, rb) = runs
    for w, x1, x3, r1, r3 in zip(ws, xa, xb, ra, rb):
        assert np.array_equal(x1, x3)
        assert r1["iters"] == r3["iters"] and r1["outer_iters"] == r3["outer_iters"]
        assert host_residual(w, x3) <= 1e-10


@pytest.mark.parametrize("precision,method", [(_lib.MIXED_IR, _lib.CG), (_lib.FP64, _lib.BICGSTAB)])
def test_batch_concurrent_lane(ehip, precision, method):
    """A batch whose larger systems take the concurrent lane: those results
    equal the system's own ehip_solve result bit for bit, the others their
    multi-system result alone."""
    kind = "tet4" if method == _lib.CG else "convdiff"
    ws = [make(kind, i) for i in range(5)]
    kw = ws[0]["kw"]
    kw.method, kw.tol, kw.maxit = method, 1e-10, 20000
    mode = _lib.FAST if precision == _lib.MIXED_IR else _lib.FAITHFUL
    sizes = sorted(W.build_system(w, optimize_bandwidth=True)[0].n for w in ws)
    cut = sizes[2]  # the three largest take the lane
    alone = []
    for w in ws:
        s = W.build_system(w, optimize_bandwidth=True)[0]
        if s.n >= cut:
            x, r = s.solve(kw, precision=precision, mode=mode)
        else:
            xs1, reps1, _ = solve_batch([s], kw, precision=precision, mode=mode, lane_workers=0)
            x, r = xs1[0], reps1[0]
        alone.append((x, r))
        s.free()
    systems = [W.build_system(w, optimize_bandwidth=True)[0] for w in ws]
    xs, reps, _ = solve_batch(systems, kw, precision=precision, mode=mode, lane_dofs=cut, lane_workers=2)
    for (xa, ra), x, r in zip(alone, xs, reps):
        assert r["info"] == ra["info"] == 1, (r, ra)
        assert ra["iters"] == r["iters"]
        assert np.array_equal(xa, x)


def test_server_continuous_batching(ehip):
    """Systems join while others run; each result equals the batch-alone result bit for bit."""
    ws = [make("tet4", i) for i in range(6)]
    kw = ws[0]["kw"]
    alone = []
    for w in ws:
        s = W.build_system(w, optimize_bandwidth=True)[0]
        xs, reps, _ = solve_batch([s], kw, precision=_lib.MIXED_IR, mode=_lib.FAST)
        alone.append(xs[0])
    srv = Server(kw, precision=_lib.MIXED_IR, max_slots=3)
    pending = [(i, W.build_system(w, optimize_bandwidth=True)[0]) for i, w in enumerate(ws)]
    ids = {}
    results = {}
    steps = 0
    while pending or srv.status()[0]:
        while pending:
            i, s = pending[0]
            if not srv.submit(s):
                break                       # no free slot: stays queued
            ids[s._h.value] = i
            pending.pop(0)
        for s, rep in srv.step():
            x = np.zeros(s.n)
            from elmerfem_amd import _lib as L
            L.check(L.load().ehip_get_solution(s._h, x.ctypes.data_as(__import__("ctypes").c_void_p)), "sol")
            results[ids[s._h.value]] = (x, rep)
        steps += 1
        assert steps < 10000
    assert srv.status() == (0, 0)
    for i in range(len(ws)):
        x, rep = results[i]
        assert rep["info"] == 1
        assert np.array_equal(x, alone[i])
    srv.free()


def test_server_admission_budget(ehip):
    w = make("tet4", 0)
    s1 = W.build_system(w)[0]
    s2 = W.build_system(make("tet4", 1))[0]
    srv = Server(w["kw"], precision=_lib.MIXED_IR, max_slots=8, mem_budget=1)   # 1 byte: nothing fits
    assert not srv.submit(s1)
    srv.free()
    srv = Server(w["kw"], precision=_lib.MIXED_IR, max_slots=1)
    assert srv.submit(s1)
    assert not srv.submit(s2)          # the only slot is taken
    while srv.status()[0]:
        srv.step()
    assert srv.submit(s2)
    srv.free()


def test_server_concurrent_lane(ehip):
    """Large systems go to the concurrent lane (whole single-system solves on
    worker threads) while small ones are batched; every result equals its
    path's result alone, bit for bit."""
    ws = [make("tet4", i) for i in range(6)]
    kw = ws[0]["kw"]
    systems = [W.build_system(w, optimize_bandwidth=True)[0] for w in ws]
    sizes = sorted(s.n for s in systems)
    cut = sizes[len(sizes) // 2]          # the larger half takes the lane
    alone = []
    for w, s0 in zip(ws, systems):
        s = W.build_system(w, optimize_bandwidth=True)[0]
        if s.n >= cut:
            x, rep = s.solve(kw, precision=_lib.MIXED_IR, mode=_lib.FAST)
        else:
            xs, reps, _ = solve_batch([s], kw, precision=_lib.MIXED_IR, mode=_lib.FAST)
            x = xs[0]
        alone.append(x)
        s.free()
    srv = Server(kw, precision=_lib.MIXED_IR, max_slots=2)
    srv.set_lane(cut, workers=2)
    pending = list(enumerate(systems))
    ids, results, steps = {}, {}, 0
    while pending or srv.status()[0]:
        while pending:
            i, s = pending[0]
            if not srv.submit(s):
                break
            ids[s._h.value] = i
            pending.pop(0)
        for s, rep in srv.step():
            x = np.zeros(s.n)
            _lib.check(_lib.load().ehip_get_solution(s._h, x.ctypes.data_as(__import__("ctypes").c_void_p)), "sol")
            results[ids[s._h.value]] = (x, rep)
        steps += 1
        assert steps < 100000
    assert srv.status() == (0, 0)
    for i in range(len(ws)):
        x, rep = results[i]
        assert rep["info"] == 1
        assert np.array_equal(x, alone[i]), i
    srv.free()
