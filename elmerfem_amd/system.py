"""Host-side mirror of the reference's Default* solver API over libehip.

A `System` is a device-resident Matrix_t (CRS Rows/Cols/Diag/Values/RHS)
with the entry points a solver module calls (fem/src/DefUtils.src):

    DefaultInitialize + element loop + DefaultUpdateEquations -> assemble()
    DefaultDirichletBCs                                      -> dirichlet()
    DefaultSolve                                             -> solve()

All work runs in libehip (HIP on the GPU); this module only moves host
arrays across the C ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .keywords import LinearSystemKeywords


def pinned_empty(n, dtype=np.float64):
    """A numpy array in pinned host memory (freed with the array): uploads
    from it are DMA copies that do not hold the caller."""
    L = _lib.init()
    dt = np.dtype(dtype)
    nbytes = max(int(n) * dt.itemsize, 1)
    p = C.c_void_p()
    _lib.check(L.ehip_host_alloc(nbytes, C.byref(p)), "ehip_host_alloc")
    buf = (C.c_char * nbytes).from_address(p.value)
    arr = np.frombuffer(buf, dtype=dt, count=int(n))
    import weakref
    weakref.finalize(buf, L.ehip_host_free, C.c_void_p(p.value))
    return arr


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, n=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError("expected %d values, got %d" % (n, a.size))
    return a


class System:
    def __init__(self, handle: C.c_void_p):
        self._h = handle
        L = _lib.load()
        n = C.c_int()
        nnz = C.c_int64()
        dofs = C.c_int()
        bb, ba, acc = C.c_int(), C.c_int(), C.c_int()
        _lib.check(L.ehip_system_info(self._h, C.byref(n), C.byref(nnz), C.byref(dofs), C.byref(bb), C.byref(ba),
                                      C.byref(acc)), "ehip_system_info")
        self.n, self.nnz, self.dofs = n.value, nnz.value, dofs.value
        self.bandwidth = (bb.value, ba.value, bool(acc.value))
        self.nnodes = self.n // self.dofs

    # ------------------------------------------------------------ creation
    @classmethod
    def from_csr(cls, rows, cols, diag):
        L = _lib.init()
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        diag = np.ascontiguousarray(diag, np.int32)
        h = C.c_void_p()
        _lib.check(L.ehip_system_create(len(rows) - 1, len(cols), _ptr(rows), _ptr(cols), _ptr(diag), C.byref(h)),
                   "ehip_system_create")
        return cls(h)

    @classmethod
    def from_mesh(cls, mesh, dofs=1, optimize_bandwidth=True, with_boundary=True):
        L = _lib.init()
        el = np.ascontiguousarray(mesh.elements, np.int32)
        be = np.ascontiguousarray(mesh.bnd_elements, np.int32) if with_boundary and len(mesh.bnd_elements) else None
        if be is not None:
            # boundary elements of a smaller type (0-padded rows, e.g. 101 points
            # read from a mesh file) add only pairs the bulk elements already have
            be = np.ascontiguousarray(be[(be > 0).all(axis=1)])
            be = be if len(be) else None
        h = C.c_void_p()
        _lib.check(L.ehip_system_from_mesh(mesh.nnodes, el.shape[0], el.shape[1], _ptr(el),
                                           0 if be is None else be.shape[0], 0 if be is None else be.shape[1],
                                           _ptr(be), dofs, int(bool(optimize_bandwidth)), C.byref(h)),
                   "ehip_system_from_mesh")
        s = cls(h)
        s.mesh = mesh
        return s

    def free(self):
        if self._h:
            _lib.load().ehip_system_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # ------------------------------------------------------------ pattern / values
    def pattern(self):
        perm = np.zeros(self.nnodes, np.int32)
        rows = np.zeros(self.n + 1, np.int32)
        cols = np.zeros(self.nnz, np.int32)
        diag = np.zeros(self.n, np.int32)
        _lib.check(_lib.load().ehip_system_get_pattern(self._h, _ptr(perm), _ptr(rows), _ptr(cols), _ptr(diag)),
                   "ehip_system_get_pattern")
        return perm, rows, cols, diag

    def set_values(self, values=None, rhs=None):
        v = _f64(values, self.nnz)
        b = _f64(rhs, self.n)
        _lib.check(_lib.load().ehip_system_set_values(self._h, _ptr(v), _ptr(b)), "ehip_system_set_values")

    def values(self):
        v = np.zeros(self.nnz)
        b = np.zeros(self.n)
        _lib.check(_lib.load().ehip_system_get_values(self._h, _ptr(v), _ptr(b)), "ehip_system_get_values")
        return v, b

    def device_bytes(self) -> int:
        return int(_lib.load().ehip_system_bytes(self._h))

    # ------------------------------------------------------------ assembly
    def assemble(self, physics, coords, elem_coef=None, elem_conv=None, elem_velo=None, elem_load=None,
                 poisson=0.3, elem_code=None, cdim=None, wait=True) -> float:
        """Element loop + DefaultUpdateEquations; returns device time in ms.
        coords None: the coordinates of the previous call (kept on the
        device).  wait False: the call only enqueues the work and returns 0;
        a degenerate element is then reported by the next call that waits on
        the system (solve, values, batch or server admission)."""
        mesh = getattr(self, "mesh", None)
        code = elem_code if elem_code is not None else mesh.elem_code
        cd = cdim if cdim is not None else (mesh.dim if mesh is not None else 3)
        ne = mesh.nelem
        xyz = None if coords is None else _f64(coords)
        t = C.c_double()
        arr = lambda a, m: None if a is None else _f64(np.broadcast_to(a, m).copy() if np.ndim(a) == 0 or np.shape(a) != m else a)
        coef = arr(elem_coef, (ne,))
        conv = arr(elem_conv, (ne,))
        velo = arr(elem_velo, (ne, 3))
        load = arr(elem_load, (ne,))
        _lib.check(_lib.load().ehip_assemble(self._h, physics, code, cd, _ptr(xyz), _ptr(coef), _ptr(conv), _ptr(velo),
                                             _ptr(load), float(poisson), C.byref(t) if wait else None),
                   "ehip_assemble")
        return t.value if wait else 0.0

    def assemble_traction(self, bnd_code, coords, belems, load3):
        be = np.ascontiguousarray(belems, np.int32)
        _lib.check(_lib.load().ehip_assemble_traction(self._h, bnd_code, _ptr(_f64(coords)), be.shape[0], be.shape[1],
                                                      _ptr(be), _ptr(_f64(load3, 3))), "ehip_assemble_traction")

    def dirichlet(self, rows, vals=None, symmetric=False, scaling=True):
        r = np.ascontiguousarray(rows, np.int32)
        v = None if vals is None else _f64(vals, r.size)
        _lib.check(_lib.load().ehip_dirichlet(self._h, r.size, _ptr(r), _ptr(v), int(symmetric), int(scaling)),
                   "ehip_dirichlet")

    def dirichlet_nodes(self, nodes, dof=None, vals=None, symmetric=False, scaling=True):
        """Rows of the given 1-based mesh nodes (all dofs, or one 1-based dof)."""
        perm, _, _, _ = self.pattern() if not hasattr(self, "_perm") else (self._perm, None, None, None)
        self._perm = perm
        k = perm[np.asarray(nodes) - 1].astype(np.int64)
        dofs = range(1, self.dofs + 1) if dof is None else [dof]
        rows = np.concatenate([self.dofs * (k - 1) + d for d in dofs]).astype(np.int32)
        self.dirichlet(rows, vals, symmetric, scaling)
        return rows

    # ------------------------------------------------------------ solve
    def solve(self, kw: LinearSystemKeywords | _lib.Params, x0=None, precision=_lib.FP64, mode=_lib.FAITHFUL,
              inner_tol=0.0, inner_maxit=0, check_every=0, history=False, return_x=True, profile_every=0,
              true_residual_stop=False):
        p = kw if isinstance(kw, _lib.Params) else kw.to_params(precision, mode, inner_tol, inner_maxit, check_every,
                                                              history, profile_every, true_residual_stop)
        if isinstance(kw, _lib.Params):
            p.record_history = int(history)
            if true_residual_stop:
                p.true_residual_stop = 1
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        rep = _lib.Report()
        hist = np.zeros(p.maxit + 2) if history else None
        _lib.check(_lib.load().ehip_solve(self._h, C.byref(p), _ptr(x) if return_x or x0 is not None else None,
                                          C.byref(rep), _ptr(hist)), "ehip_solve")
        out = rep.as_dict()
        if history:
            k = rep.outer_iters + 1 if p.precision == _lib.MIXED_IR else min(rep.iters, p.maxit)
            out["history"] = hist[:k].copy()
        return (x if return_x else None), out

    def solution(self):
        """The solution currently on the device (the last solve's, back-scaled)."""
        x = np.zeros(self.n)
        _lib.check(_lib.load().ehip_get_solution(self._h, _ptr(x)), "ehip_get_solution")
        return x

    # ------------------------------------------------------------ ILU(k)
    def ilu_factor(self, order=0):
        """CRS_IncompleteLU of the current Values: (irows, icols, idiag, factors), 1-based."""
        L = _lib.load()
        m = C.c_int64()
        _lib.check(L.ehip_ilu_factor(self._h, order, C.byref(m), None, None, None, None), "ehip_ilu_factor")
        ir = np.zeros(self.n + 1, np.int32)
        ic = np.zeros(m.value, np.int32)
        idg = np.zeros(self.n, np.int32)
        f = np.zeros(m.value)
        _lib.check(L.ehip_ilu_factor(self._h, order, C.byref(m), _ptr(ir), _ptr(ic), _ptr(idg), _ptr(f)),
                   "ehip_ilu_factor")
        return ir, ic, idg, f

    def ilu_levels(self):
        """(forward, backward) dependency levels of the current ILU factors."""
        lf, lb = C.c_int(), C.c_int()
        _lib.check(_lib.load().ehip_ilu_info(self._h, None, None, C.byref(lf), C.byref(lb)), "ehip_ilu_info")
        return lf.value, lb.value

    def lu_solve(self, b):
        """CRS_LUPrecondition with the last factors: (LU)^-1 b."""
        b = _f64(b, self.n)
        x = np.zeros(self.n)
        _lib.check(_lib.load().ehip_lu_solve(self._h, _ptr(b), _ptr(x)), "ehip_lu_solve")
        return x

    # ------------------------------------------------------------ SpMV
    def spmv(self, u, precision=0):
        u = _f64(u, self.n)
        v = np.zeros(self.n)
        _lib.check(_lib.load().ehip_spmv(self._h, precision, _ptr(u), _ptr(v)), "ehip_spmv")
        return v

    def spmv_bench(self, precision=0, reps=50) -> float:
        us = C.c_double()
        _lib.check(_lib.load().ehip_spmv_bench(self._h, precision, reps, C.byref(us)), "ehip_spmv_bench")
        return us.value

    def spmv_tune(self, precision=0):
        _lib.check(_lib.load().ehip_spmv_tune(self._h, precision), "ehip_spmv_tune")
        return self.spmv_config(precision)

    def spmv_set_schedule(self, block, nnz_per_block, precision=0):
        _lib.check(_lib.load().ehip_spmv_set_schedule(self._h, precision, block, nnz_per_block),
                   "ehip_spmv_set_schedule")

    def spmv_set_format(self, fmt, precision=0):
        _lib.check(_lib.load().ehip_spmv_set_format(self._h, precision, fmt), "ehip_spmv_set_format")

    def spmv_sell_layout(self):
        """{"sort_window": 0 or the length-sort window, "node_block": 0 or DOFs per node,
        "lane_unroll": nonzeros (blocks, with node_block) each lane loads at once,
        "pair_unroll": FP64 value pairs each lane loads at once (0: not pair-packed)}"""
        w, b, u, pu = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _lib.check(_lib.load().ehip_spmv_sell_layout(self._h, C.byref(w), C.byref(b), C.byref(u), C.byref(pu)),
                   "ehip_spmv_sell_layout")
        return {"sort_window": w.value, "node_block": b.value, "lane_unroll": u.value, "pair_unroll": pu.value}

    def spmv_shared_layout(self):
        """{"node_block": 0 or DOFs per node block, "lane_unroll": entries (blocks) each lane
        loads at once, "entries": stored values} of the slice-shared offset layout
        (EhipError when the pattern has none)."""
        b, u, e = C.c_int(), C.c_int(), C.c_int64()
        _lib.check(_lib.load().ehip_spmv_shared_layout(self._h, C.byref(b), C.byref(u), C.byref(e)),
                   "ehip_spmv_shared_layout")
        return {"node_block": b.value, "lane_unroll": u.value, "entries": e.value}

    def spmv_config(self, precision=0):
        f, b, nb, ch, ib = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _lib.check(_lib.load().ehip_spmv_config(self._h, precision, C.byref(f), C.byref(b), C.byref(nb),
                                                C.byref(ch), C.byref(ib)), "ehip_spmv_config")
        name = {_lib.SPMV_SELL: "sell", _lib.SPMV_SELL_SHARED: "sell_shared"}.get(f.value, "csr")
        return {"format": name, "block": b.value,
                "nnz_per_block": nb.value, "chunks": ch.value, "index_bits": ib.value}


def copy_bench(nbytes: int, reps: int = 20):
    """Plain device copy of nbytes: (GB/s counting read + write, us per call)."""
    L = _lib.init()
    g, us = C.c_double(), C.c_double()
    _lib.check(L.ehip_copy_bench(int(nbytes), reps, C.byref(g), C.byref(us)), "ehip_copy_bench")
    return g.value, us.value


def solve_csr(rows, cols, diag, values, b, kw, x0=None, precision=_lib.FP64, mode=_lib.FAITHFUL,
              true_residual_stop=False, noalloc=False):
    """One-shot drop-in (the Before-Linsolve hook path): host CRS in, x out.
    noalloc=True calls ehip_solve_csr_noalloc (no device allocation; FP64 only,
    within EHIP_NOALLOC_MAX_N / EHIP_NOALLOC_MAX_NNZ)."""
    L = _lib.init()
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    diag = np.ascontiguousarray(diag, np.int32)
    v = _f64(values)
    bb = _f64(b)
    x = np.zeros(len(rows) - 1) if x0 is None else np.array(x0, dtype=np.float64)
    p = kw if isinstance(kw, _lib.Params) else kw.to_params(precision, mode,
                                                          true_residual_stop=true_residual_stop)
    rep = _lib.Report()
    fn, name = (L.ehip_solve_csr_noalloc, "ehip_solve_csr_noalloc") if noalloc else (L.ehip_solve_csr,
                                                                                     "ehip_solve_csr")
    _lib.check(fn(len(rows) - 1, _ptr(rows), _ptr(cols), _ptr(diag), _ptr(v), _ptr(bb), _ptr(x), C.byref(p),
                  C.byref(rep)), name)
    return x, rep.as_dict()
