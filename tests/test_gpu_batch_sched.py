"""Batched FP64 fast mode (true-residual restarts) and the multi-system
schedule table and its search.

Bars: every FP64 fast batched solve reaches a host-recomputed unscaled
residual <= 1e-10 and the oracle's solution within 1e-8, including systems
on which HUTI's BiCGStab stops with its S-norm exit (code 36); a system's
result depends on its own size class's chunk size only -- alone and in a
mix, and under any workgroup count, the bits are the same; a solve after a
schedule search gives the same bits as a later solve under the table the
search left (the search restores the admitted state exactly).
"""
import numpy as np
import pytest

import oracle_lib as O
from elmerfem_amd import _lib, workloads as W
from elmerfem_amd.batch import schedule, set_schedule, solve_batch

pytestmark = pytest.mark.gpu

# (side, index) of Quad4 convection-diffusion problems on which HUTI's
# BiCGStab (the oracle, and the compiled reference it is pinned to) stops
# with HUTI_BICGSTAB_SNORM at tol 1e-10 with the unscaled residual above
# 1e-10 (5.4e-8, 2.9e-9, 2.5e-10)
SNORM_CASES = [(32, 42), (32, 49), (32, 13)]
CLASS_NNZ = [0, 1 << 15, 1 << 17, 1 << 19, 1 << 21, 1 << 23]


def oracle_system(w):
    m = w["mesh"]
    pat = O.Pattern(m.elements, m.nnodes, w["dofs"], optimize=True, extra_elements=[m.bnd_elements])
    vals, rhs = O.assemble(w["physics"], m, pat, w["coef"], conv=w.get("conv"), velo=w.get("velo"),
                           load=w.get("load"), poisson=w.get("poisson", 0.3))
    O.dirichlet(pat, vals, rhs, pat.perm[w["bc_nodes"] - 1])
    return pat, vals, rhs


def host_residual(w, x):
    pat, vals, rhs = oracle_system(w)
    ax = np.zeros(pat.n)
    O.lib().or_spmv(pat.n, O._p(pat.rows), O._p(pat.cols), O._p(vals), O._p(np.ascontiguousarray(x)), O._p(ax))
    return np.linalg.norm(rhs - ax) / np.linalg.norm(rhs)


def make(kind, i):
    if kind == "convdiff":
        return W.convdiff_quad4(8 + 5 * i, index=i + 1)
    return W.heat_tet4(3 + i, index=i + 1)


@pytest.fixture
def default_table(ehip):
    """Leaves the process-wide schedule table at its defaults afterwards."""
    yield
    kw = W._kw("cg")
    for prec, mode in ((_lib.FP64, _lib.FAITHFUL), (_lib.FP64, _lib.FAST), (_lib.MIXED_IR, _lib.FAST)):
        for method in (_lib.CG, _lib.BICGSTAB):
            kw.method = method
            for nnz in CLASS_NNZ:
                set_schedule(kw, nnz, 0, 0, precision=prec, mode=mode)


@pytest.mark.parametrize("kind,method", [("convdiff", _lib.BICGSTAB), ("tet4", _lib.CG)])
def test_batch_fp64_fast_matches_oracle(ehip, kind, method):
    ws = [make(kind, i) for i in range(5)]
    print("fp64 fast batch %s" % kind, flush=True)
    kw = ws[0]["kw"]
    kw.method, kw.tol, kw.maxit = method, 1e-10, 20000
    systems = [W.build_system(w, optimize_bandwidth=True)[0] for w in ws]
    xs, reps, _ = solve_batch(systems, kw, precision=_lib.FP64, mode=_lib.FAST, true_residual_stop=True)
    for w, x, r in zip(ws, xs, reps):
        assert r["info"] == 1, r
        assert r["res_scaled"] < 1e-10 and r["res_true"] <= 1e-10
        assert host_residual(w, x) <= 1e-10
        pat, vals, rhs = oracle_system(w)
        info, k, xo, hist, nrm = O.solve_linear_system(pat, vals.copy(), rhs, method=method, tol=1e-12, maxit=20000)
        assert info == 1
        assert np.abs(x - xo).max() <= 1e-8 * np.abs(xo).max()
    # alone == in the mix, bit for bit
    s = W.build_system(ws[2], optimize_bandwidth=True)[0]
    xa, ra, _ = solve_batch([s], kw, precision=_lib.FP64, mode=_lib.FAST, true_residual_stop=True)
    assert np.array_equal(xa[0], xs[2]) and ra[0]["iters"] == reps[2]["iters"]
    for s in systems:
        s.free()


def test_batch_fp64_fast_restarts_past_snorm(ehip):
    """HUTI order stops at the S-norm exit like the oracle; the fast mode
    confirms on the true residual, restarts, and reaches 1e-10."""
    ws = [W.convdiff_quad4(n, index=i) for n, i in SNORM_CASES]
    kw = ws[0]["kw"]
    kw.method, kw.tol, kw.maxit = _lib.BICGSTAB, 1e-10, 20000
    for w in ws:
        pat, vals, rhs = oracle_system(w)
        info, k, xo, hist, nrm = O.solve_linear_system(pat, vals.copy(), rhs, method=1, tol=1e-10, maxit=20000)
        assert info == 36
    systems = [W.build_system(w, optimize_bandwidth=True)[0] for w in ws]
    print("snorm cases faithful", flush=True)
    _, reps, _ = solve_batch(systems, kw, precision=_lib.FP64, mode=_lib.FAITHFUL)
    assert all(r["info"] == 36 for r in reps), reps
    print("snorm cases fast", flush=True)
    xs, reps, _ = solve_batch(systems, kw, precision=_lib.FP64, mode=_lib.FAST, true_residual_stop=True)
    for w, s, x, r in zip(ws, systems, xs, reps):
        assert r["info"] == 1, r
        assert r["outer_iters"] >= 1          # at least one restart from the true residual
        assert host_residual(w, x) <= 1e-10
        x1, r1 = s.solve(kw, precision=_lib.FP64, mode=_lib.FAST, true_residual_stop=True)
        assert r1["info"] == 1
        assert np.abs(x - x1).max() <= 1e-8 * np.abs(x1).max()
    for s in systems:
        s.free()


@pytest.mark.parametrize("precision,mode", [(_lib.MIXED_IR, _lib.FAST), (_lib.FP64, _lib.FAITHFUL)])
def test_schedule_chunk_size_is_per_system(default_table, precision, mode):
    """Under a table with different chunk sizes per size class, a system's
    bits alone equal its bits in the mix, and the workgroup count changes
    nothing."""
    ws = [make("tet4", i) for i in range(2)] + [make("tet4", 13)]   # nnz 622, 1333, 66961
    kw = ws[0]["kw"]

    def fresh():
        # a solve scales the values in place and scales them back, which is
        # not an exact round trip: every comparison starts from new systems
        return [W.build_system(w, optimize_bandwidth=True)[0] for w in ws]

    systems = fresh()
    classes = sorted({s.nnz for s in systems})
    assert len({_class(n) for n in classes}) >= 2
    set_schedule(kw, classes[0], 1024, 1024, precision=precision, mode=mode)
    set_schedule(kw, classes[-1], 4096, -1, precision=precision, mode=mode)
    assert schedule(kw, classes[0], precision=precision, mode=mode) == (1024, 1024)
    assert schedule(kw, classes[-1], precision=precision, mode=mode) == (4096, 1024)
    print("schedule mix %d" % precision, flush=True)
    xs, reps, _ = solve_batch(systems, kw, precision=precision, mode=mode, true_residual_stop=True)
    for i, s in enumerate(fresh()):
        xa, ra, _ = solve_batch([s], kw, precision=precision, mode=mode, true_residual_stop=True)
        assert ra[0]["info"] == reps[i]["info"] == 1
        assert np.array_equal(xa[0], xs[i]), i
        assert host_residual(ws[i], xs[i]) <= 1e-10
        s.free()
    set_schedule(kw, 0, -1, 4096, precision=precision, mode=mode)
    xg, rg, _ = solve_batch(fresh(), kw, precision=precision, mode=mode, true_residual_stop=True)
    for a, b in zip(xs, xg):
        assert np.array_equal(a, b)
    # another chunk size regroups the partial sums: same solution to rounding
    set_schedule(kw, classes[-1], 1024, -1, precision=precision, mode=mode)
    xc, rc, _ = solve_batch(fresh(), kw, precision=precision, mode=mode, true_residual_stop=True)
    for a, b, r in zip(xs, xc, rc):
        assert r["info"] == 1
        assert np.abs(a - b).max() <= 1e-8 * np.abs(a).max()
    for s in systems:
        s.free()


def _class(nnz):
    c = 0
    lim = 1 << 15
    while c < 5 and nnz >= lim:
        c += 1
        lim <<= 2
    return c


@pytest.mark.parametrize("precision,mode,method", [(_lib.MIXED_IR, _lib.FAST, _lib.BICGSTAB),
                                                   (_lib.FP64, _lib.FAST, _lib.CG)])
def test_schedule_search_restores_state(default_table, precision, mode, method):
    """The search fills the table with compiled choices and leaves every slot
    in its admitted state: the tuned solve equals a later untuned solve
    under the same table, bit for bit."""
    kind = "convdiff" if method == _lib.BICGSTAB else "tet4"
    ws = [make(kind, i) for i in (0, 3, 6)]
    kw = ws[0]["kw"]
    kw.method, kw.tol, kw.maxit = method, 1e-10, 20000
    systems = [W.build_system(w, optimize_bandwidth=True)[0] for w in ws]
    print("schedule search %d %d" % (precision, method), flush=True)
    xt, rt, _ = solve_batch(systems, kw, precision=precision, mode=mode, true_residual_stop=True, tune=2,
                            tune_replays=1)
    for s in systems:
        nb, grid = schedule(kw, s.nnz, precision=precision, mode=mode)
        assert nb in (1024, 2048, 4096) and grid in (1024, 2048, 4096)
    # new systems: the tuned solve's scale round trip moved the values' last bits
    systems2 = [W.build_system(w, optimize_bandwidth=True)[0] for w in ws]
    xs, rs, _ = solve_batch(systems2, kw, precision=precision, mode=mode, true_residual_stop=True)
    for s in systems2:
        s.free()
    for w, a, b, ra, rb in zip(ws, xt, xs, rt, rs):
        assert ra["info"] == rb["info"] == 1
        assert ra["iters"] == rb["iters"]
        assert np.array_equal(a, b)
        assert host_residual(w, a) <= 1e-10
    for s in systems:
        s.free()
