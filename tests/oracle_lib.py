"""ctypes access to the checker libraries (oracle/liboracle.so and, when it
has been built in this container, oracle/_ref/libhuti_ref.so).

Test infrastructure only: the oracle checks the product, it is never the
thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libhuti_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lib = None
_ref = None


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = C.CDLL(ORACLE_SO)
        vp = C.c_void_p
        L.or_gauss1d.argtypes = [C.c_int, vp, vp]
        L.or_gauss_points.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
        L.or_element_info.argtypes = [C.c_int, C.c_int, vp, vp, vp, C.c_double, C.c_double, C.c_double, vp, vp, vp]
        L.or_local_modelpde.argtypes = [C.c_int, C.c_int] + [vp] * 10
        L.or_local_elasticity.argtypes = [C.c_int] + [vp] * 7
        L.or_local_traction.argtypes = [C.c_int] + [vp] * 5
        L.or_pattern_build.restype = vp
        L.or_pattern_build.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int]
        L.or_pattern_n.argtypes = [vp]
        L.or_pattern_nnz.argtypes = [vp]
        L.or_pattern_nnz.restype = C.c_longlong
        L.or_pattern_bandwidth.argtypes = [vp, vp, vp, vp]
        L.or_pattern_copy.argtypes = [vp, vp, vp, vp, vp]
        L.or_pattern_free.argtypes = [vp]
        L.or_glue_local.argtypes = [C.c_int, vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        L.or_assemble.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int, C.c_int, vp, vp, vp, vp, vp,
                                  C.c_double, vp, C.c_int, vp, vp, vp, C.c_int, vp, vp]
        L.or_assemble_traction.argtypes = [C.c_int, C.c_int, vp, C.c_int, C.c_int, vp, vp, vp, C.c_int, vp]
        L.or_dirichlet.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_int, vp, vp, C.c_int, C.c_int]
        L.or_ddot.argtypes = [C.c_int, vp, vp]
        L.or_ddot.restype = C.c_double
        L.or_dnrm2.argtypes = [C.c_int, vp]
        L.or_dnrm2.restype = C.c_double
        L.or_spmv.argtypes = [C.c_int, vp, vp, vp, vp, vp]
        L.or_huti_solve.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, C.c_double, C.c_int,
                                    C.c_int, vp, vp]
        L.or_solve_linear_system.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp,
                                             C.c_double, C.c_int, vp, vp, vp]
        L.or_compute_norm.argtypes = [C.c_int, vp]
        L.or_compute_norm.restype = C.c_double
        L.or_set_threads.argtypes = [C.c_int]
        L.or_ilu_pattern.argtypes = [C.c_int, vp, vp, vp, C.c_int, vp, vp, vp]
        L.or_free.argtypes = [vp]
        L.or_ilu_factor.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, vp]
        L.or_lu_solve.argtypes = [C.c_int, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_SO)
        vp = C.c_void_p
        R.ref_huti_solve.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, C.c_double, C.c_int, vp, vp]
        R.ref_ddot.argtypes = [C.c_int, vp, vp]
        R.ref_ddot.restype = C.c_double
        R.ref_dnrm2.argtypes = [C.c_int, vp]
        R.ref_dnrm2.restype = C.c_double
        R.ref_set_ilu.argtypes = [vp, vp, vp, vp]
        _ref = R
    return _ref


# ---------------------------------------------------------------- wrappers

def gauss1d(n):
    p = np.zeros(n)
    w = np.zeros(n)
    assert lib().or_gauss1d(n, _p(p), _p(w)) == 0
    return p, w


def gauss_points(code, np_):
    u, v, w, s = (np.zeros(64) for _ in range(4))
    k = lib().or_gauss_points(code, np_, _p(u), _p(v), _p(w), _p(s))
    assert k > 0
    return u[:k], v[:k], w[:k], s[:k]


def element_info(code, cdim, xyz, u, v, w):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    x, y, z = (np.ascontiguousarray(xyz[:, i]) for i in range(3))
    basis = np.zeros(n)
    db = np.zeros((n, 3))
    detj = np.zeros(1)
    ok = lib().or_element_info(code, cdim, _p(x), _p(y), _p(z), u, v, w, _p(basis), _p(db), _p(detj))
    return ok, basis, db, float(detj[0])


def local_modelpde(code, cdim, xyz, diff, conv=None, react=None, velo=None, load=None):
    xyz = np.asarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    x, y, z = (np.ascontiguousarray(xyz[:, i]) for i in range(3))
    f = lambda a: np.ascontiguousarray(np.zeros(n) if a is None else np.broadcast_to(a, (n,)), dtype=np.float64)
    vel = np.ascontiguousarray(np.zeros((n, 3)) if velo is None else np.broadcast_to(velo, (n, 3)), dtype=np.float64)
    stiff = np.zeros(n * n)
    force = np.zeros(n)
    rc = lib().or_local_modelpde(code, cdim, _p(x), _p(y), _p(z), _p(f(diff)), _p(f(conv)), _p(f(react)),
                                 _p(vel), _p(f(load)), _p(stiff), _p(force))
    assert rc == 0, rc
    return stiff.reshape(n, n, order="F"), force


def local_elasticity(code, xyz, young, poisson):
    xyz = np.asarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    x, y, z = (np.ascontiguousarray(xyz[:, i]) for i in range(3))
    E = np.ascontiguousarray(np.broadcast_to(young, (n,)), dtype=np.float64)
    nu = np.ascontiguousarray(np.broadcast_to(poisson, (n,)), dtype=np.float64)
    stiff = np.zeros(9 * n * n)
    force = np.zeros(3 * n)
    rc = lib().or_local_elasticity(code, _p(x), _p(y), _p(z), _p(E), _p(nu), _p(stiff), _p(force))
    assert rc == 0, rc
    return stiff.reshape(3 * n, 3 * n, order="F"), force


class Pattern:
    def __init__(self, mesh_elements, nnodes, dofs=1, optimize=False, extra_elements=()):
        blocks = [np.asarray(mesh_elements, np.int32)] + [np.asarray(b, np.int32) for b in extra_elements]
        maxnpe = max(b.shape[1] for b in blocks)
        el = np.zeros((sum(len(b) for b in blocks), maxnpe), np.int32)
        npe = np.zeros(len(el), np.int32)
        o = 0
        for b in blocks:
            el[o:o + len(b), :b.shape[1]] = b
            npe[o:o + len(b)] = b.shape[1]
            o += len(b)
        L = lib()
        h = L.or_pattern_build(nnodes, len(el), maxnpe, _p(el), _p(npe), dofs, 1 if optimize else 0)
        self.n = L.or_pattern_n(h)
        self.nnz = L.or_pattern_nnz(h)
        self.perm = np.zeros(nnodes, np.int32)
        self.rows = np.zeros(self.n + 1, np.int32)
        self.cols = np.zeros(self.nnz, np.int32)
        self.diag = np.zeros(self.n, np.int32)
        L.or_pattern_copy(h, _p(self.perm), _p(self.rows), _p(self.cols), _p(self.diag))
        b, a, acc = (np.zeros(1, np.int32) for _ in range(3))
        L.or_pattern_bandwidth(h, _p(b), _p(a), _p(acc))
        self.bw_before, self.bw_after, self.accepted = int(b[0]), int(a[0]), bool(acc[0])
        L.or_pattern_free(h)
        self.dofs = dofs


def assemble(kind, mesh, pat, coef, conv=None, velo=None, load=None, poisson=0.3):
    n = pat.n
    values = np.zeros(pat.nnz)
    rhs = np.zeros(n)
    coords = np.ascontiguousarray(mesh.coords, np.float64)
    el = np.ascontiguousarray(mesh.elements, np.int32)
    coef = np.ascontiguousarray(np.broadcast_to(coef, (mesh.nelem,)), np.float64)
    conv_a = None if conv is None else np.ascontiguousarray(np.broadcast_to(conv, (mesh.nelem,)), np.float64)
    velo_a = None if velo is None else np.ascontiguousarray(np.broadcast_to(velo, (mesh.nelem, 3)), np.float64)
    load_a = None if load is None else np.ascontiguousarray(np.broadcast_to(load, (mesh.nelem,)), np.float64)
    rc = lib().or_assemble(kind, mesh.elem_code, 3 if mesh.dim == 3 else 2, mesh.nnodes, _p(coords), mesh.nelem,
                           el.shape[1], _p(el), _p(coef), _p(conv_a), _p(velo_a), _p(load_a), float(poisson),
                           _p(pat.perm), n, _p(pat.rows), _p(pat.cols), _p(pat.diag), pat.dofs, _p(values), _p(rhs))
    assert rc == 0, rc
    return values, rhs


def assemble_traction(mesh, pat, bnd_sel, load3, rhs):
    coords = np.ascontiguousarray(mesh.coords, np.float64)
    be = np.ascontiguousarray(mesh.bnd_elements[bnd_sel], np.int32)
    L3 = np.ascontiguousarray(load3, np.float64)
    rc = lib().or_assemble_traction(mesh.bnd_code, mesh.nnodes, _p(coords), len(be), be.shape[1], _p(be),
                                    _p(L3), _p(pat.perm), pat.dofs, _p(rhs))
    assert rc == 0, rc


def dirichlet(pat, values, rhs, bc_rows, bc_vals=None, symmetric=False, scaling=True):
    bc_rows = np.ascontiguousarray(bc_rows, np.int32)
    bc_vals = np.ascontiguousarray(np.zeros(len(bc_rows)) if bc_vals is None else bc_vals, np.float64)
    lib().or_dirichlet(pat.n, _p(pat.rows), _p(pat.cols), _p(pat.diag), _p(values), _p(rhs), len(bc_rows),
                       _p(bc_rows), _p(bc_vals), int(symmetric), int(scaling))


def solve_linear_system(pat, values, b, method=0, precond=1, scaling=True, tol=1e-10, maxit=10000, x=None):
    n = pat.n
    x = np.zeros(n) if x is None else np.ascontiguousarray(x, np.float64).copy()
    it = np.zeros(1, np.int32)
    hist = np.zeros(maxit)
    nrm = np.zeros(1)
    bb = np.ascontiguousarray(b, np.float64).copy()
    info = lib().or_solve_linear_system(method, precond, int(scaling), n, _p(pat.rows), _p(pat.cols),
                                        _p(pat.diag), _p(values), _p(bb), _p(x), tol, maxit, _p(it), _p(hist),
                                        _p(nrm))
    k = int(it[0])
    return info, k, x, hist[:k], float(nrm[0])


def huti_solve(rows, cols, diag, values, b, x0, method=0, precond=1, tol=1e-10, maxit=10000, stopc=1):
    n = len(rows) - 1
    x = np.ascontiguousarray(x0, np.float64).copy()
    it = np.zeros(1, np.int32)
    hist = np.zeros(maxit)
    info = lib().or_huti_solve(method, precond, n, _p(rows), _p(cols), _p(diag), _p(values), _p(b), _p(x), tol,
                               maxit, stopc, _p(it), _p(hist))
    k = int(it[0])
    return info, k, x, hist[:k]


def ilu_pattern(rows, cols, diag, order):
    """1-based ILU(order) pattern (irows, icols, idiag) of a CRS pattern."""
    n = len(rows) - 1
    pr, pc, pd = C.c_void_p(), C.c_void_p(), C.c_void_p()
    rc = lib().or_ilu_pattern(n, _p(rows), _p(cols), _p(diag), order, C.byref(pr), C.byref(pc), C.byref(pd))
    assert rc == 0
    ir = np.ctypeslib.as_array(C.cast(pr, C.POINTER(C.c_int)), (n + 1,)).copy()
    ic = np.ctypeslib.as_array(C.cast(pc, C.POINTER(C.c_int)), (int(ir[n]) - 1,)).copy()
    idg = np.ctypeslib.as_array(C.cast(pd, C.POINTER(C.c_int)), (n,)).copy()
    for q in (pr, pc, pd):
        lib().or_free(q)
    return ir, ic, idg


def ilu_factor(rows, cols, values, irows, icols, idiag):
    """CRS_IncompleteLU factors on the given ILU pattern (U(i,i) stored inverted)."""
    n = len(rows) - 1
    ilu = np.zeros(int(irows[n]) - 1)
    rc = lib().or_ilu_factor(n, _p(rows), _p(cols), _p(np.ascontiguousarray(values, np.float64)), _p(irows),
                             _p(icols), _p(idiag), _p(ilu))
    assert rc == 0
    return ilu


def lu_solve(irows, icols, idiag, ilu, b):
    x = np.ascontiguousarray(b, np.float64).copy()
    lib().or_lu_solve(len(irows) - 1, _p(irows), _p(icols), _p(idiag), _p(ilu), _p(x))
    return x


def ref_huti_solve(rows, cols, diag, values, b, x0, method=0, precond=1, tol=1e-10, maxit=10000):
    n = len(rows) - 1
    x = np.ascontiguousarray(x0, np.float64).copy()
    it = np.zeros(1, np.int32)
    hist = np.zeros(maxit)
    keep = None
    if precond >= 2:
        # the factors IterSolver would compute, handed to the driver's CRS_LUPrecondition
        ir, ic, idg = ilu_pattern(rows, cols, diag, precond - 2)
        keep = (ir, ic, idg, ilu_factor(rows, cols, values, ir, ic, idg))
        ref().ref_set_ilu(*[_p(a) for a in keep])
    info = ref().ref_huti_solve(method, precond, n, _p(rows), _p(cols), _p(diag), _p(values), _p(b), _p(x), tol,
                                maxit, _p(it), _p(hist))
    k = int(it[0])
    return info, k, x, hist[:k]


def scaled_system(pat, values, b):
    """ScaleLinearSystem of a copy (SolverUtils.src:6470-6640), as IterSolver sees it.
    Returns (scaled values, scaled b, x0=1e-8)."""
    n = pat.n
    d = values[pat.diag - 1].copy()
    tiny = np.finfo(np.float64).tiny
    if np.any(np.abs(d) <= tiny):
        for i in np.nonzero(np.abs(d) <= tiny)[0]:
            d[i] = np.abs(values[pat.rows[i] - 1:pat.rows[i + 1] - 1]).sum()
    D = np.where(np.abs(d) > tiny, 1.0 / np.sqrt(np.abs(d)), 1.0)
    rowid = np.repeat(np.arange(n), np.diff(pat.rows))
    v = values * (D[rowid] * D[pat.cols - 1])
    bb = b * D
    bn = np.sqrt(np.sum(bb * bb))
    return v, bb / bn, np.full(n, 1e-8)
