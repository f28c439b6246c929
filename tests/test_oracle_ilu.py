"""Pins the oracle's ILU(n) restatement (or_ilu_pattern, or_ilu_factor,
or_lu_solve; CRSMatrix.src:2157-2500, 3293-3397) by what the mathematics
fixes, not by re-deriving its loops:

* incomplete LU on a pattern P reproduces A on P: (L U)_ij = a_ij for every
  (i, j) in P, fills included (a_ij = 0 there) -- a dropped update, a wrong
  index or a transposed factor breaks it;
* where no fill is dropped (tridiagonal, or a fill order that closes the
  band) the factors are the exact LU and the LU solve inverts A;
* the ILU(1) pattern of the 5-point Laplacian on an N x N grid adds exactly
  the two diagonals at offsets +-(N - 1): 2 (N - 1)^2 fills;
* the reference's special cases: no division by a pivot with |U_kk| <= AEPS
  and the stored inverse 1 for |U_ii| < AEPS.

The compiled reference cannot build CRSMatrix.src here (it needs Elmer's
generated build configuration), so these properties are the pin; the HUTI
loops driven by these factors are compared bit for bit with the reference's
own compiled HUTI through the golden vectors (test_oracle_ref.py).
"""
import numpy as np
import pytest

import oracle_lib as O

AEPS = 10 * np.finfo(float).eps


def csr_from_dense(A):
    n = A.shape[0]
    rows, cols, vals, diag = [1], [], [], []
    for i in range(n):
        for j in range(n):
            if A[i, j] != 0 or i == j:
                cols.append(j + 1)
                vals.append(A[i, j])
                if i == j:
                    diag.append(len(cols))
        rows.append(len(cols) + 1)
    return (np.array(rows, np.int32), np.array(cols, np.int32), np.array(diag, np.int32), np.array(vals))


def dense_factors(ir, ic, idg, ilu):
    n = len(ir) - 1
    L = np.eye(n)
    U = np.zeros((n, n))
    for i in range(n):
        for k in range(ir[i] - 1, ir[i + 1] - 1):
            j = ic[k] - 1
            if k < idg[i] - 1:
                L[i, j] = ilu[k]
            elif k == idg[i] - 1:
                U[i, j] = 1.0 / ilu[k]
            else:
                U[i, j] = ilu[k]
    return L, U


def laplace5(N, rng=None, nonsym=0.0):
    n = N * N
    A = np.zeros((n, n))
    for y in range(N):
        for x in range(N):
            i = x + N * y
            A[i, i] = 4.0 + (rng.uniform(0, 1) if rng is not None else 0.0)
            for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                if 0 <= x + dx < N and 0 <= y + dy < N:
                    A[i, (x + dx) + N * (y + dy)] = -1.0 + nonsym * (dx + 2 * dy)
    return A


def random_sparse(n, density, rng):
    A = (rng.uniform(size=(n, n)) < density) * rng.standard_normal((n, n))
    A = A + A.T * (rng.uniform(size=(n, n)) < 0.5)   # some structural asymmetry
    A[np.arange(n), np.arange(n)] = np.abs(A).sum(1) + 1.0
    return A


def pattern_of(ir, ic):
    n = len(ir) - 1
    return [(i, ic[k] - 1) for i in range(n) for k in range(ir[i] - 1, ir[i + 1] - 1)]


@pytest.mark.parametrize("order", [0, 1, 2, 3])
@pytest.mark.parametrize("kind", ["laplace", "nonsym", "random"])
def test_lu_reproduces_a_on_the_pattern(kind, order):
    rng = np.random.default_rng(7 + order)
    A = {"laplace": lambda: laplace5(6), "nonsym": lambda: laplace5(5, rng, 0.3),
         "random": lambda: random_sparse(40, 0.08, rng)}[kind]()
    rows, cols, diag, vals = csr_from_dense(A)
    ir, ic, idg = O.ilu_pattern(rows, cols, diag, order)
    ilu = O.ilu_factor(rows, cols, vals, ir, ic, idg)
    L, U = dense_factors(ir, ic, idg, ilu)
    LU = L @ U
    scale = np.abs(A).max()
    for i, j in pattern_of(ir, ic):
        assert abs(LU[i, j] - A[i, j]) <= 1e-13 * scale, (i, j, LU[i, j], A[i, j])
    # the pattern holds every entry of A and stays sorted with its diagonal
    for i in range(A.shape[0]):
        row = ic[ir[i] - 1:ir[i + 1] - 1]
        assert np.all(np.diff(row) > 0) and ic[idg[i] - 1] == i + 1
        assert set(np.nonzero(A[i])[0] + 1) <= set(row)
    # the LU solve inverts the product of the factors
    b = rng.standard_normal(A.shape[0])
    x = O.lu_solve(ir, ic, idg, ilu, b)
    assert np.allclose(LU @ x, b, rtol=0, atol=1e-12 * np.abs(b).max())


def test_tridiagonal_is_exact_lu():
    rng = np.random.default_rng(1)
    n = 50
    A = np.diag(rng.uniform(3, 4, n)) + np.diag(rng.uniform(-1, 0, n - 1), 1) + np.diag(rng.uniform(-1, 0, n - 1), -1)
    rows, cols, diag, vals = csr_from_dense(A)
    ir, ic, idg = O.ilu_pattern(rows, cols, diag, 0)
    ilu = O.ilu_factor(rows, cols, vals, ir, ic, idg)
    b = rng.standard_normal(n)
    assert np.allclose(O.lu_solve(ir, ic, idg, ilu, b), np.linalg.solve(A, b), rtol=1e-13, atol=1e-13)
    # no fill exists, so every order gives the same pattern
    assert np.array_equal(O.ilu_pattern(rows, cols, diag, 3)[1], ic)


def test_closed_band_is_exact_lu():
    """On a 4 x 4 grid every fill of the exact LU lies in the band |i - j| <= 4;
    enough fill rounds close it and the LU solve is then exact."""
    A = laplace5(4)
    rows, cols, diag, vals = csr_from_dense(A)
    ir, ic, idg = O.ilu_pattern(rows, cols, diag, 9)
    ilu = O.ilu_factor(rows, cols, vals, ir, ic, idg)
    b = np.arange(1.0, 17.0)
    assert np.allclose(O.lu_solve(ir, ic, idg, ilu, b), np.linalg.solve(A, b), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("N", [3, 5, 8])
def test_ilu1_fill_of_the_five_point_laplacian(N):
    A = laplace5(N)
    rows, cols, diag, vals = csr_from_dense(A)
    ir, ic, idg = O.ilu_pattern(rows, cols, diag, 1)
    assert ir[-1] - rows[-1] == 2 * (N - 1) ** 2
    fills = set(pattern_of(ir, ic)) - set(pattern_of(rows, cols))
    assert all(abs(i - j) == N - 1 for i, j in fills)


def test_small_pivot_special_cases():
    # [[0, 1], [1, 1]]: U_11 = 0 is not divided by (L_21 stays 1), U_22 = 1 - 1*1 = 0,
    # and both stored inverses become 1
    rows = np.array([1, 3, 5], np.int32)
    cols = np.array([1, 2, 1, 2], np.int32)
    diag = np.array([1, 4], np.int32)
    vals = np.array([0.0, 1.0, 1.0, 1.0])
    ir, ic, idg = O.ilu_pattern(rows, cols, diag, 0)
    ilu = O.ilu_factor(rows, cols, vals, ir, ic, idg)
    assert np.array_equal(ilu, [1.0, 1.0, 1.0, 1.0])
    # a pivot exactly at AEPS is inverted (the prescale tests |U_ii| < AEPS) but
    # not divided by during elimination (that tests |U_kk| > AEPS)
    vals = np.array([AEPS, 1.0, 2.0, 3.0])
    ilu = O.ilu_factor(rows, cols, vals, ir, ic, idg)
    assert ilu[0] == 1.0 / AEPS and ilu[2] == 2.0 and ilu[3] == 1.0 / (3.0 - 2.0 * 1.0)
    # a stored zero in the lower part is skipped: the row is left as is
    vals = np.array([2.0, 1.0, 0.0, 3.0])
    ilu = O.ilu_factor(rows, cols, vals, ir, ic, idg)
    assert np.array_equal(ilu, [0.5, 1.0, 0.0, 1.0 / 3.0])
