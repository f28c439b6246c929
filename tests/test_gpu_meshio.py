"""The path from a mesh file to a solution on the GPU, through the C ABI.

The reference's own heateq mesh (tests/golden/meshes/heateq_tri3, from
fem/tests/heateq/Mesh: 734 Tri3, 80 Line2 + 6 point boundary elements) is
read with ehip_mesh_read; the CRS pattern with OptimizeBandwidth must equal
the oracle's bit for bit, the assembled Values and RHS match the oracle
within 1e-12, and the solves give the reference test's Target NRM
0.768016492512E-01 (fem/tests/heateq/TempDist.sif: conductivity 1, source 1,
T = 0 on boundaries 1-6) within its default tolerance 1e-5, in each solver
form; with no source and T = 5 on the boundary (fem/tests/linearsolvers)
NRM is 5.  Tet10 meshes (generated) solve to the oracle's solution.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from elmerfem_amd import _lib, meshio, workloads as W
from elmerfem_amd.system import System

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
HEATEQ = os.path.join(HERE, "golden", "meshes", "heateq_tri3")
TARGET = 0.768016492512E-01


def heateq_system(bc_value, source):
    m = meshio.read_mesh(HEATEQ)
    s = System.from_mesh(m, dofs=1, optimize_bandwidth=True)
    s.assemble(0, m.coords, elem_coef=np.ones(m.nelem), elem_load=np.full(m.nelem, source))
    nodes = m.boundary_nodes(range(1, 7))
    vals = np.full(len(nodes), bc_value)
    return m, s, nodes, vals


def test_heateq_pattern_and_values(ehip):
    m = meshio.read_mesh(HEATEQ)
    lines = m.bnd_elements[m.info["bnd_types"] == 202]
    pat = O.Pattern(m.elements, m.nnodes, 1, optimize=True, extra_elements=[lines])
    s = System.from_mesh(m, dofs=1, optimize_bandwidth=True)
    perm, rows, cols, diag = s.pattern()
    assert np.array_equal(perm, pat.perm) and np.array_equal(rows, pat.rows)
    assert np.array_equal(cols, pat.cols) and np.array_equal(diag, pat.diag)
    s.assemble(0, m.coords, elem_coef=np.ones(m.nelem), elem_load=np.ones(m.nelem))
    gv, gb = s.values()
    ov, ob = O.assemble(0, m, pat, np.ones(m.nelem), load=np.ones(m.nelem))
    assert np.abs(gv - ov).max() <= 1e-12 * np.abs(ov).max()
    assert np.abs(gb - ob).max() <= 1e-12 * np.abs(ob).max()
    s.free()


@pytest.mark.parametrize("method,precond,precision,mode,tol", [
    ("bicgstab", "ilu0", _lib.FP64, _lib.FAITHFUL, 1e-8),        # the sif's solver
    ("cg", "diagonal", _lib.FP64, _lib.FAITHFUL, 1e-10),
    ("cg", "diagonal", _lib.FP64, _lib.FAST, 1e-10),
    ("cg", "diagonal", _lib.MIXED_IR, _lib.FAST, 1e-10),
    ("cg", "ilu1", _lib.FP64, _lib.FAITHFUL, 1e-10),
])
def test_heateq_target_nrm(ehip, method, precond, precision, mode, tol):
    m, s, nodes, vals = heateq_system(0.0, 1.0)
    s.dirichlet_nodes(nodes, vals=vals)
    kw = W._kw(method, precond, tol, 350 if method == "bicgstab" else 2000)
    x, rep = s.solve(kw, precision=precision, mode=mode)
    assert rep["info"] == 1, rep
    assert abs(rep["norm"] - TARGET) <= 1e-5 * TARGET, rep["norm"]
    s.free()


@pytest.mark.parametrize("method", ["cg", "bicgstab"])
def test_linearsolvers_constant(ehip, method):
    m, s, nodes, vals = heateq_system(5.0, 0.0)
    s.dirichlet_nodes(nodes, vals=vals)
    x, rep = s.solve(W._kw(method, "ilu0", 1e-12, 3500), precision=_lib.FP64, mode=_lib.FAITHFUL)
    assert rep["info"] == 1 and abs(rep["norm"] - 5.0) <= 1e-9
    assert np.abs(x - 5.0).max() <= 1e-9
    s.free()


@pytest.mark.parametrize("wl", ["heat", "elastic"])
def test_tet10_solve_matches_oracle(ehip, wl):
    w = W.heat_tet10(3) if wl == "heat" else W.elasticity_tet10(2)
    m = w["mesh"]
    print("tet10 %s: %d nodes, %d elements" % (wl, m.nnodes, m.nelem), flush=True)
    pat = O.Pattern(m.elements, m.nnodes, w["dofs"], extra_elements=[m.bnd_elements])
    vals, rhs = O.assemble(w["physics"], m, pat, w["coef"], load=w.get("load"), poisson=w.get("poisson", 0.3))
    if wl == "elastic":
        rhs[2::3] = -1e-3 * np.ones(len(rhs) // 3) / len(rhs)   # a body load in z on every node
    k = pat.perm[w["bc_nodes"] - 1].astype(np.int64)
    rows = np.concatenate([w["dofs"] * (k - 1) + d for d in range(1, w["dofs"] + 1)]).astype(np.int32)
    rhs0 = rhs.copy()
    O.dirichlet(pat, vals, rhs, rows)
    info, kk, xo, hist, nrm = O.solve_linear_system(pat, vals.copy(), rhs, tol=1e-10, maxit=5000)
    s = System.from_mesh(m, dofs=w["dofs"], optimize_bandwidth=False)
    s.assemble(w["physics"], m.coords, elem_coef=w["coef"], elem_load=w.get("load"), poisson=w.get("poisson", 0.3))
    gv, _ = s.values()
    s.set_values(gv, rhs0)
    s.dirichlet(rows)
    x, rep = s.solve(w["kw"], precision=_lib.FP64, mode=_lib.FAST, true_residual_stop=True)
    assert rep["info"] == info == 1
    assert np.abs(x - xo).max() <= 1e-8 * np.abs(xo).max()
    s.free()
