"""Batching and serving on the multi-system engine of libehip.

`solve_batch` runs many independent systems in one graph-replayed launch
sequence (each system keeps its own scalars and stopping test).  `Server`
is the continuous-batching form: systems are admitted while a slot and the
HBM budget allow, every `step()` advances all live systems by one block of
iterations, and finished systems leave with their reports.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def solve_batch(systems, kw, precision=_lib.FP64, mode=_lib.FAITHFUL, inner_tol=0.0, inner_maxit=0,
                check_every=0, lane_dofs=None, lane_workers=None, true_residual_stop=False, tune=0,
                tune_replays=0):
    """Returns (solutions, reports, ms).  With lane_workers > 0, systems of
    at least lane_dofs unknowns (default 40000) are solved whole on that many
    concurrent threads instead of in the multi-system launches.

    FP64 runs HUTI's loop order with mode FAITHFUL and the recurrence
    residual with true-residual restarts with mode FAST.  tune = 1 runs the
    schedule search for the size classes the process-wide schedule table
    does not hold yet (tune = 2: for all) before solving; the chosen chunk
    sizes and workgroup count stay in the table (see `schedule`)."""
    L = _lib.init()
    arr = (C.c_void_p * len(systems))(*[s._h for s in systems])
    h = C.c_void_p()
    _lib.check(L.ehip_batch_create(len(systems), arr, C.byref(h)), "ehip_batch_create")
    try:
        if lane_dofs is not None or lane_workers is not None:
            _lib.check(L.ehip_batch_set_lane(h, 40000 if lane_dofs is None else int(lane_dofs),
                                             0 if lane_workers is None else int(lane_workers)),
                       "ehip_batch_set_lane")
        p = kw if isinstance(kw, _lib.Params) else kw.to_params(precision, mode, inner_tol, inner_maxit, check_every,
                                                                true_residual_stop=true_residual_stop)
        if tune:
            _lib.check(L.ehip_batch_set_tuning(h, int(tune), int(tune_replays)), "ehip_batch_set_tuning")
        reps = (_lib.Report * len(systems))()
        ms = C.c_double()
        _lib.check(L.ehip_batch_solve(h, C.byref(p), reps, C.byref(ms)), "ehip_batch_solve")
        xs = []
        for i, s in enumerate(systems):
            x = np.zeros(s.n)
            _lib.check(L.ehip_batch_get_solution(h, i, x.ctypes.data_as(C.c_void_p)), "ehip_batch_get_solution")
            xs.append(x)
        return xs, [r.as_dict() for r in reps], ms.value
    finally:
        L.ehip_batch_free(h)


def _params(kw, precision, mode):
    return kw if isinstance(kw, _lib.Params) else kw.to_params(precision, mode, 0.0, 0, 0)


def schedule(kw, nnz, precision=_lib.FP64, mode=_lib.FAITHFUL):
    """(chunk nonzeros, workgroups) the multi-system launches use for a
    system of `nnz` nonzeros in this solver form."""
    L = _lib.init()
    p = _params(kw, precision, mode)
    nb, grid = C.c_int(), C.c_int()
    _lib.check(L.ehip_batch_get_schedule(C.byref(p), int(nnz), C.byref(nb), C.byref(grid)), "ehip_batch_get_schedule")
    return nb.value, grid.value


def set_schedule(kw, nnz, chunk_nnz=-1, grid=-1, precision=_lib.FP64, mode=_lib.FAITHFUL):
    """Sets the table entry for the size class of `nnz` (chunk_nnz 1024,
    2048 or 4096; 0 = default; -1 = unchanged) and the form's workgroup
    count (a multiple of 8; 0 = default; -1 = unchanged)."""
    L = _lib.init()
    p = _params(kw, precision, mode)
    _lib.check(L.ehip_batch_set_schedule(C.byref(p), int(nnz), int(chunk_nnz), int(grid)), "ehip_batch_set_schedule")


class Server:
    """Continuous batching over device-resident systems."""

    def __init__(self, kw, precision=_lib.MIXED_IR, max_slots=256, mem_budget=0, inner_tol=0.0, inner_maxit=0,
                 check_every=0, true_residual_stop=False):
        L = _lib.init()
        self.p = kw if isinstance(kw, _lib.Params) else kw.to_params(precision, _lib.FAST, inner_tol, inner_maxit,
                                                                     check_every,
                                                                     true_residual_stop=true_residual_stop)
        self._h = C.c_void_p()
        _lib.check(L.ehip_server_create(C.byref(self.p), max_slots, int(mem_budget), C.byref(self._h)),
                   "ehip_server_create")
        self.max_slots = max_slots
        self._live = {}

    def set_lane(self, min_dofs: int, workers: int = 4):
        """Systems of at least min_dofs unknowns are solved whole on `workers`
        host threads (overlapping on the device) instead of being batched."""
        _lib.check(_lib.load().ehip_server_set_lane(self._h, int(min_dofs), int(workers)), "ehip_server_set_lane")

    def submit(self, system) -> bool:
        """Admits the system; False when no slot or budget is free (keep it queued)."""
        slot = C.c_int()
        rc = _lib.load().ehip_server_submit(self._h, system._h, C.byref(slot))
        if rc == -5:  # EHIP_ECAP
            return False
        _lib.check(rc, "ehip_server_submit")
        self._live[system._h.value] = system
        return True

    def step(self):
        """One block of iterations for every live system; returns [(system, report)] of finished ones."""
        k = self.max_slots + 64  # batched slots plus the concurrent lane's workers
        nd = C.c_int()
        sys_out = (C.c_void_p * k)()
        rep_out = (_lib.Report * k)()
        _lib.check(_lib.load().ehip_server_step(self._h, C.byref(nd), sys_out, rep_out, k), "ehip_server_step")
        done = []
        for i in range(nd.value):
            s = self._live.pop(sys_out[i])
            done.append((s, rep_out[i].as_dict()))
        return done

    def status(self):
        live, nbytes = C.c_int(), C.c_int64()
        _lib.check(_lib.load().ehip_server_status(self._h, C.byref(live), C.byref(nbytes)), "ehip_server_status")
        return live.value, nbytes.value

    def free(self):
        if self._h:
            _lib.load().ehip_server_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
