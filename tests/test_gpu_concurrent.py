"""Concurrent single-system solves: several host threads, each solving its
own system on its own stream, the way the server's concurrent lane runs
them.  Every solve must converge and repeat its own result exactly, however
the solves overlap on the device.  (A CG variant that wrote its stop flag
from inside the SpMV launch that also tests the flag on entry -- once per
wavefront -- corrupted a few per cent of such solves, iteration-limit stops
or NaN, while sequential solves were unaffected: a late wavefront returned
while its siblings summed its unwritten LDS slots.  This is its regression
test.)"""
import queue
import threading

import numpy as np
import pytest

from elmerfem_amd import _lib, workloads as W

pytestmark = pytest.mark.gpu


def test_concurrent_mixed_cg_solves_converge_and_repeat(ehip):
    sizes = (60, 66, 72, 78)  # 0.23-0.49 M unknowns: the size where it showed
    systems = [W.build_system(W.heat_tet4(n, seed=20261202, index=i + 1), optimize_bandwidth=True)[0]
               for i, n in enumerate(sizes)]
    kw = W._kw("cg", "diagonal", 1e-10, 20000)
    # each system's reference: the same solve sequence run alone (every
    # solve scales and back-scales the matrix in place, so the k-th solve of
    # a system is compared with its own k-th solve)
    rounds = 20
    ref = {}
    for i, s in enumerate(systems):
        for k in range(rounds):
            _, r = s.solve(kw, precision=_lib.MIXED_IR, mode=_lib.FAST, return_x=False, true_residual_stop=True)
            ref[i, k] = (r["info"], r["iters"], r["res_true"])
    systems2 = [W.build_system(W.heat_tet4(n, seed=20261202, index=i + 1), optimize_bandwidth=True)[0]
                for i, n in enumerate(sizes)]
    q = queue.Queue()
    for i in range(len(systems2)):
        q.put(i)
    got, errors = {}, []

    def work():
        while True:
            try:
                i = q.get_nowait()
            except queue.Empty:
                return
            try:
                for k in range(rounds):
                    _, r = systems2[i].solve(kw, precision=_lib.MIXED_IR, mode=_lib.FAST, return_x=False,
                                            true_residual_stop=True)
                    got[i, k] = (r["info"], r["iters"], r["res_true"])
            except Exception as e:  # reported below
                errors.append(repr(e))

    th = [threading.Thread(target=work) for _ in range(len(systems2))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    bad = [(key, got[key], ref[key]) for key in ref if got[key] != ref[key]]
    assert not bad, bad[:5]
    assert all(v[0] == 1 and v[2] <= 1e-10 for v in got.values())
    for s in systems + systems2:
        s.free()
