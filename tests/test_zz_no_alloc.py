"""The no-allocation solve (ehip_solve_csr_noalloc) through the C ABI.

Its results must match ehip_solve_csr's on the same input (same HUTI
status, iterations within 2 -- its SpMV is the CSR stream layout, whose
per-workgroup partial sums differ from the SELL layout's --, solutions
within 1e-8 relative of each other and of the oracle, and the unscaled FP64
residual recomputed on the host <= 1e-10 with the unscaled stop test on).
Every case is two calls in a row, the second on a smaller system, so stale
contents of the shared scratch array from the first call would show; limits
and unsupported options come back as status codes before any device work.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from elmerfem_amd import _lib
from elmerfem_amd.system import solve_csr
from golden.golden_cases import CASES, assembled_case

pytestmark = pytest.mark.gpu

RES_BAR = 1e-10


def host_residual(pat, vals, rhs, x):
    ax = np.zeros(pat.n)
    O.lib().or_spmv(pat.n, O._p(pat.rows), O._p(pat.cols), O._p(vals), O._p(np.ascontiguousarray(x)), O._p(ax))
    return np.linalg.norm(rhs - ax) / np.linalg.norm(rhs)


def check_against_allocating(name, mode):
    spec = CASES[name]
    w, pat, vals, rhs, _ = assembled_case(spec)
    print("no-alloc case %s n=%d nnz=%d mode=%d" % (name, pat.n, len(pat.cols), mode), flush=True)
    xa, ra = solve_csr(pat.rows, pat.cols, pat.diag, vals, rhs, w["kw"], mode=mode, true_residual_stop=True)
    xn, rn = solve_csr(pat.rows, pat.cols, pat.diag, vals, rhs, w["kw"], mode=mode, true_residual_stop=True,
                       noalloc=True)
    info, k, xo, hist, nrm = O.solve_linear_system(pat, vals.copy(), rhs, method=spec["method"],
                                                   precond=spec["precond"], tol=spec["tol"], maxit=spec["maxit"])
    assert rn["info"] == ra["info"] == 1, (rn, ra)
    assert abs(rn["iters"] - ra["iters"]) <= 2, (rn["iters"], ra["iters"])
    scale = np.abs(xa).max()
    assert np.abs(xn - xa).max() <= 1e-8 * scale
    assert np.abs(xn - xo).max() <= 1e-8 * np.abs(xo).max()
    assert host_residual(pat, vals, rhs, xn) <= RES_BAR
    assert abs(rn["norm"] - ra["norm"]) <= 1e-8 * ra["norm"]
    return xn, rn


@pytest.mark.parametrize("first,second,mode", [
    ("cg_poisson32", "cg_lognormal8", _lib.FAITHFUL),
    ("cg_poisson32", "cg_none_lognormal8", _lib.FAST),
    ("bicgstab_poisson8", "bicgstab_convdiff16", _lib.FAST),
])
def test_noalloc_matches_solve_csr_twice(ehip, first, second, mode):
    check_against_allocating(first, mode)
    check_against_allocating(second, mode)


def test_noalloc_repeatable(ehip):
    """The same input twice gives the same bits (nothing carried over)."""
    spec = CASES["cg_lognormal8"]
    w, pat, vals, rhs, _ = assembled_case(spec)
    x1, r1 = solve_csr(pat.rows, pat.cols, pat.diag, vals, rhs, w["kw"], noalloc=True)
    spec2 = CASES["cg_poisson8"]
    w2, pat2, vals2, rhs2, _ = assembled_case(spec2)
    solve_csr(pat2.rows, pat2.cols, pat2.diag, vals2, rhs2, w2["kw"], noalloc=True)
    x2, r2 = solve_csr(pat.rows, pat.cols, pat.diag, vals, rhs, w["kw"], noalloc=True)
    assert np.array_equal(x1, x2) and r1["iters"] == r2["iters"]


def _identity(n):
    rows = np.arange(1, n + 2, dtype=np.int32)
    cols = np.arange(1, n + 1, dtype=np.int32)
    return rows, cols, cols.copy(), np.ones(n), np.ones(n)


def _call(n, rows, cols, diag, vals, b, p):
    L = _lib.init()
    x = np.zeros(n)
    rep = _lib.Report()
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    return L.ehip_solve_csr_noalloc(n, ptr(rows), ptr(cols), ptr(diag), ptr(vals), ptr(b), ptr(x), C.byref(p),
                                    C.byref(rep)), x, rep


def test_noalloc_limits_and_options(ehip):
    from elmerfem_amd.keywords import from_solver_section
    kw = from_solver_section({"Linear System Iterative Method": "CG", "Linear System Preconditioning": "Diagonal",
                              "Linear System Convergence Tolerance": 1e-10, "Linear System Max Iterations": 100})
    n = _lib.NOALLOC_MAX_N + 1
    rc, _, _ = _call(n, *_identity(n), kw.to_params(_lib.FP64, _lib.FAST))
    assert rc == _lib.ECAP
    n = 64
    rc, _, _ = _call(n, *_identity(n), kw.to_params(_lib.MIXED_IR, _lib.FAST))
    assert rc == _lib.EUNSUP
    # ILU(k) would allocate its factors: refused before any device work
    kw_ilu = from_solver_section({"Linear System Iterative Method": "CG", "Linear System Preconditioning": "ILU0",
                                  "Linear System Convergence Tolerance": 1e-10, "Linear System Max Iterations": 100})
    rc, _, _ = _call(n, *_identity(n), kw_ilu.to_params(_lib.FP64, _lib.FAST))
    assert rc == _lib.EUNSUP
    # within the limits the identity system solves at once
    rc, x, rep = _call(n, *_identity(n), kw.to_params(_lib.FP64, _lib.FAST))
    assert rc == 0 and rep.info == 1 and np.allclose(x, 1.0, rtol=1e-12)
