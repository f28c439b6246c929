"""Decoding of the 'Linear System' keywords (fem/src/IterSolve.src:221-411,
fem/src/SolverUtils.src:7256-7257) into ehip_params."""
import pytest

from elmerfem_amd import _lib
from elmerfem_amd.keywords import KeywordError, from_solver_section, parse_sif_solver

BASE = {"Linear System Convergence Tolerance": "1e-10", "Linear System Max Iterations": "500"}


def sec(**kw):
    d = dict(BASE)
    d.update({k.replace("_", " "): v for k, v in kw.items()})
    return d


def test_defaults_follow_itersolver():
    k = from_solver_section(BASE)
    assert k.method == _lib.BICGSTAB  # no method keyword -> BiCGStab (IterSolve.src:221-226)
    assert k.precond == _lib.PRECOND_NONE
    assert k.scaling is True and k.abort_not_converged is True and k.optimize_bandwidth is True
    assert k.symmetric is False and k.residual_output == 1
    assert k.tol == 1e-10 and k.maxit == 500


@pytest.mark.parametrize("name,code", [("CG", _lib.CG), ("cg", _lib.CG), ("BiCGStab", _lib.BICGSTAB),
                                       ("bicgstab", _lib.BICGSTAB), ("something else", _lib.BICGSTAB)])
def test_method_prefix_decoding(name, code):
    k = from_solver_section(sec(**{"Linear System Iterative Method": name}))
    assert k.method == code


@pytest.mark.parametrize("name", ["BiCGStabl", "BiCGStab2", "GMRES", "TFQMR", "CGS", "Jacobi", "GCR"])
def test_other_methods_refused(name):
    with pytest.raises(KeywordError):
        from_solver_section(sec(**{"Linear System Iterative Method": name}))


def test_preconditioning():
    assert from_solver_section(sec(**{"Linear System Preconditioning": "Diagonal"})).precond == _lib.PRECOND_DIAGONAL
    assert from_solver_section(sec(**{"Linear System Preconditioning": "None"})).precond == _lib.PRECOND_NONE
    # unknown names fall back to none (IterSolve.src:408-410)
    assert from_solver_section(sec(**{"Linear System Preconditioning": "foo"})).precond == _lib.PRECOND_NONE
    for name in ("ILUT", "Multigrid", "Vanka"):
        with pytest.raises(KeywordError):
            from_solver_section(sec(**{"Linear System Preconditioning": name}))


@pytest.mark.parametrize("name,extra,dofs,k", [
    ("ILU0", {}, 1, 0), ("ILU1", {}, 1, 1), ("ilu3", {}, 1, 3), ("ILU9", {}, 1, 9),
    ("ILU", {}, 1, 0),                                   # str(4:4) = ' ' -> -16 -> 0
    ("ILUx", {}, 1, 0),                                  # not a digit -> 0
    ("ILU", {"Linear System ILU Order": "2"}, 1, 2),     # the order keyword wins over the name
    ("ILU1", {"Linear System ILU Order": "12"}, 1, 0),   # out of 0..9 -> 0
    ("BILU2", {}, 1, 2),                                 # block ILU of one DOF is ILU
])
def test_ilu_orders(name, extra, dofs, k):
    kw = from_solver_section(sec(**dict({"Linear System Preconditioning": name}, **extra)), dofs=dofs)
    assert kw.precond == _lib.PRECOND_ILU0 + k


def test_ilu_outside_path():
    with pytest.raises(KeywordError):
        from_solver_section(sec(**{"Linear System Preconditioning": "BILU0"}), dofs=3)
    with pytest.raises(KeywordError):
        from_solver_section(sec(**{"Linear System Preconditioning": "ILU0", "Linear System Symmetric ILU": "True"}))


def test_required_keywords_and_direct_solver():
    with pytest.raises(KeywordError):
        from_solver_section({"Linear System Max Iterations": "10"})
    with pytest.raises(KeywordError):
        from_solver_section({"Linear System Convergence Tolerance": "1e-8"})
    with pytest.raises(KeywordError):
        from_solver_section(sec(**{"Linear System Max Iterations": "0"}))
    with pytest.raises(KeywordError):
        from_solver_section(sec(**{"Linear System Solver": "Direct"}))


def test_logicals_and_params():
    k = from_solver_section(sec(**{"Linear System Scaling": "False", "Linear System Symmetric": "True",
                                   "Linear System Iterative Method": "CG",
                                   "Linear System Preconditioning": "diagonal",
                                   "Linear System Divergence Tolerance": "1e6"}))
    assert k.scaling is False and k.symmetric is True
    p = k.to_params(precision=_lib.MIXED_IR, mode=_lib.FAST, inner_tol=1e-3)
    assert (p.method, p.precond, p.scaling, p.precision, p.mode) == (_lib.CG, _lib.PRECOND_DIAGONAL, 0,
                                                                      _lib.MIXED_IR, _lib.FAST)
    assert p.tol == 1e-10 and p.maxit == 500 and p.inner_tol == 1e-3 and p.divtol == 1e6


def test_parse_sif_block():
    text = """
Solver 1
  Equation = Heat Equation
  Linear System Solver = "Iterative"   ! comment
  Linear System Iterative Method = CG
  Linear System Preconditioning = Diagonal
  Linear System Convergence Tolerance = Real 1.0e-10
  Linear System Max Iterations = Integer 10000
  Linear System Residual Output = 0
  Optimize Bandwidth = Logical False
End
"""
    k = from_solver_section(parse_sif_solver(text))
    assert k.method == _lib.CG and k.precond == _lib.PRECOND_DIAGONAL
    assert k.tol == 1e-10 and k.maxit == 10000 and k.residual_output == 0
    assert k.optimize_bandwidth is False
