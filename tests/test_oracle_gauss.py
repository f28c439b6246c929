"""Oracle Gauss rules against closed forms and exactness (Integration.src)."""
import math

import numpy as np
import pytest

import oracle_lib as O


def test_gauss1d_closed_forms():
    p, w = O.gauss1d(1)
    assert p[0] == 0.0 and w[0] == 2.0
    p, w = O.gauss1d(2)
    # ComputeGaussPoints1D orders the roots +sqrt, -sqrt
    assert p[0] > 0 > p[1]
    assert abs(p[0] - 1 / math.sqrt(3)) < 2e-16 and abs(p[1] + 1 / math.sqrt(3)) < 2e-16
    assert np.allclose(w, [1.0, 1.0], atol=1e-15, rtol=0)
    p, w = O.gauss1d(3)
    assert abs(p[0] - math.sqrt(0.6)) < 2e-16 and abs(p[1] + math.sqrt(0.6)) < 2e-16 and p[2] == 0.0
    assert np.allclose(w, [5 / 9, 5 / 9, 8 / 9], atol=1e-15, rtol=0)
    assert abs(w.sum() - 2.0) < 1e-15


@pytest.mark.parametrize("n", [2, 3])
def test_gauss1d_exact_degree(n):
    p, w = O.gauss1d(n)
    for deg in range(2 * n):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.sum(w * p ** deg) - exact) < 1e-14
    # and not one degree beyond
    deg = 2 * n
    assert abs(np.sum(w * p ** deg) - 2.0 / (deg + 1)) > 1e-4


@pytest.mark.parametrize("code,np_,nexp", [(404, 4, 4), (409, 9, 9), (808, 8, 8), (827, 27, 27), (504, 4, 4),
                                           (510, 27, 27), (303, 3, 3)])
def test_default_rules(code, np_, nexp):
    u, v, w, s = O.gauss_points(code, np_)
    assert len(u) == nexp
    fam = code // 100
    if fam == 4:
        assert abs(s.sum() - 4.0) < 1e-14                      # area of [-1,1]^2
        assert abs(np.sum(s * u ** 2 * v ** 2) - 4 / 9) < 1e-14
    elif fam == 8:
        assert abs(s.sum() - 8.0) < 1e-14
        assert abs(np.sum(s * u ** 2 * v ** 2 * w ** 2) - 8 / 27) < 1e-14
        # brick ordering: w outer, v, u inner (GaussPointsBrick)
        m = round(len(u) ** (1 / 3))
        assert np.all(u[:m] == u[m:2 * m]) and np.all(w[:m * m] == w[0])
    elif fam == 3:
        assert abs(s.sum() - 0.5) < 1e-15                       # area of the unit triangle
        assert abs(np.sum(s * u ** 2) - 1 / 12) < 1e-15         # degree-2 exactness
        assert abs(np.sum(s * u * v) - 1 / 24) < 1e-15
    else:
        assert abs(s.sum() - 1 / 6) < 1e-14                     # volume of the unit tetrahedron
        # degree-2 exactness of the 4-point rule: int x^2 over the tet = 1/60
        assert abs(np.sum(s * u ** 2) - 1 / 60) < 1e-10
        assert abs(np.sum(s * u * v) - 1 / 120) < 1e-10
        assert np.all(u + v + w <= 1 + 1e-15) and np.all(u >= 0) and np.all(v >= 0) and np.all(w >= 0)
