"""The oracle's HUTI CG/BiCGStab and BLAS-1 against the reference's own
compiled sources: golden vectors always, the live oracle/_ref build when it
is present in this container."""
import os

import numpy as np
import pytest

import oracle_lib as O
from golden.golden_cases import CASES, scaled_case

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_golden(name):
    spec = CASES[name]
    g = np.load(os.path.join(GOLD, "huti_%s.npz" % name))
    pat, v, b, x0 = scaled_case(spec)
    assert pat.n == int(g["n"]) and pat.nnz == int(g["nnz"])
    assert np.sum(b) == float(g["bsum"]) and np.sum(v) == float(g["vsum"])   # the same input system
    info, k, x, hist = O.huti_solve(pat.rows, pat.cols, pat.diag, v, b, x0, method=spec["method"],
                                    precond=spec["precond"], tol=spec["tol"], maxit=spec["maxit"])
    assert info == int(g["info"]) and k == int(g["iters"])
    assert np.array_equal(hist, g["hist"])                      # bit-exact residual history
    xs = x if bool(g["full_x"]) else x[::97]
    assert np.array_equal(xs, g["x"])                            # bit-exact solution


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here (needs the reference tree)")
def test_blas_against_reference_sources():
    rng = np.random.default_rng(3)
    for n in [1, 2, 4, 5, 6, 11, 1000, 12347]:
        x = rng.standard_normal(n) * 10.0 ** rng.integers(-150, 150, n)
        y = rng.standard_normal(n)
        x[::7] = 0.0
        assert O.lib().or_ddot(n, O._p(x), O._p(y)) == O.ref().ref_ddot(n, O._p(x), O._p(y))
        assert O.lib().or_dnrm2(n, O._p(x)) == O.ref().ref_dnrm2(n, O._p(x))
    assert O.lib().or_dnrm2(0, O._p(np.zeros(1))) == 0.0


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here (needs the reference tree)")
@pytest.mark.parametrize("name", ["cg_lognormal8", "bicgstab_convdiff16", "cg_elastic3"])
def test_oracle_matches_live_reference(name):
    spec = dict(CASES[name])
    pat, v, b, x0 = scaled_case(spec)
    for maxit in (spec["maxit"], 3):
        a = O.huti_solve(pat.rows, pat.cols, pat.diag, v, b, x0, method=spec["method"], precond=spec["precond"],
                         tol=spec["tol"], maxit=maxit)
        r = O.ref_huti_solve(pat.rows, pat.cols, pat.diag, v, b, x0, method=spec["method"],
                             precond=spec["precond"], tol=spec["tol"], maxit=maxit)
        assert a[0] == r[0] and a[1] == r[1]
        assert np.array_equal(a[2], r[2]) and np.array_equal(a[3], r[3])


def test_ddot_known_answer():
    x = np.arange(1, 8, dtype=float)
    assert O.lib().or_ddot(7, O._p(x), O._p(x)) == 140.0
    # dnrm2 avoids overflow where the plain sum of squares would not
    big = np.array([1e200, 1e200])
    assert O.lib().or_dnrm2(2, O._p(big)) == pytest.approx(np.sqrt(2) * 1e200, rel=1e-15)
