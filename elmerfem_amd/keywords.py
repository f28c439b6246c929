"""The .sif 'Linear System ...' keywords the drop-in honours, with the
reference's defaults and decoding rules.

Decoding follows IterSolver (fem/src/IterSolve.src:221-366, 372-411) and
SolveLinearSystem (fem/src/SolverUtils.src:7256-7257, 7402-7415):

- 'Linear System Iterative Method' is matched by prefix in the reference's
  order; a missing keyword means BiCGStab, and an unrecognised name also
  falls back to BiCGStab (IterSolve.src:251-252).
- 'Linear System Preconditioning' defaults to 'none'; 'diagonal' selects
  CRS_DiagPrecondition; anything else is not part of this path.
- 'Linear System Scaling' defaults to True.
- 'Linear System Abort Not Converged' defaults to True.
- 'Linear System Residual Output' defaults to 1 (printing is a host concern).
"""
from __future__ import annotations

from dataclasses import dataclass

from . import _lib

SUPPORTED_METHODS = {"cg": _lib.CG, "bicgstab": _lib.BICGSTAB}


class KeywordError(ValueError):
    pass


@dataclass
class LinearSystemKeywords:
    solver: str = "iterative"
    method: int = _lib.BICGSTAB
    method_name: str = "bicgstab"
    precond: int = _lib.PRECOND_NONE
    tol: float = 1e-8
    maxit: int = 500
    scaling: bool = True
    symmetric: bool = False
    abort_not_converged: bool = True
    residual_output: int = 1
    divtol: float = 0.0
    optimize_bandwidth: bool = True

    def to_params(self, precision: int = _lib.FP64, mode: int = _lib.FAITHFUL, inner_tol: float = 0.0,
                  inner_maxit: int = 0, check_every: int = 0, record_history: bool = False,
                  profile_every: int = 0, true_residual_stop: bool = False) -> _lib.Params:
        return _lib.Params(self.method, self.precond, self.tol, self.maxit, int(self.scaling), precision, mode,
                           inner_tol, inner_maxit, check_every, self.divtol, int(record_history), profile_every,
                           int(true_residual_stop))


def _method(name: str | None) -> tuple[int, str]:
    """Prefix decoding of IterSolve.src:229-253."""
    if name is None:
        return _lib.BICGSTAB, "bicgstab"
    s = name.strip().lower()
    order = [("bicgstab2", None), ("bicgstabl", None), ("bicgstab", _lib.BICGSTAB), ("tfqmr", None),
             ("cgs", None), ("cg", _lib.CG), ("gmres", None), ("sgs", None), ("jacobi", None),
             ("richardson", None), ("gcr", None)]
    for prefix, code in order:
        if s.startswith(prefix):
            if code is None:
                raise KeywordError("Linear System Iterative Method = %s is outside the accelerated path "
                                   "(CG and BiCGStab)" % name)
            return code, prefix
    return _lib.BICGSTAB, "bicgstab"  # unknown names run BiCGStab in the reference


def _ilu_digit(s: str, pos: int) -> int:
    """ICHAR(str(pos:pos)) - ICHAR('0') clamped to 0 outside 0..9 (IterSolve.src:387-395)."""
    c = s[pos - 1] if len(s) >= pos else " "
    k = ord(c) - ord("0")
    return k if 0 <= k <= 9 else 0


def _precond(name: str | None, ilu_order=None, dofs: int = 1, symmetric_ilu=False) -> int:
    """Prefix decoding of IterSolve.src:372-411."""
    if name is None:
        return _lib.PRECOND_NONE
    s = name.strip().lower()
    if s.startswith("none"):
        return _lib.PRECOND_NONE
    if s.startswith("diagonal"):
        return _lib.PRECOND_DIAGONAL
    if s.startswith("ilut") or s.startswith("multigrid") or s.startswith("vanka"):
        raise KeywordError("Linear System Preconditioning = %s is outside the accelerated path "
                           "(none, diagonal, ILU0..ILU9)" % name)
    if s.startswith("ilu") or s.startswith("bilu"):
        if s.startswith("bilu"):
            if dofs > 1:
                raise KeywordError("Linear System Preconditioning = %s (block ILU for %d DOFs) is outside "
                                   "the accelerated path" % (name, dofs))
            k = _ilu_digit(s, 5)  # BILU for one DOF is ILU (IterSolve.src:396-405)
        elif ilu_order is not None:
            k = int(round(float(ilu_order)))
            k = k if 0 <= k <= 9 else 0
        else:
            k = _ilu_digit(s, 4)
        if symmetric_ilu:
            raise KeywordError("Linear System Symmetric ILU (incomplete Cholesky) is outside the accelerated path")
        return _lib.PRECOND_ILU0 + k
    return _lib.PRECOND_NONE  # unknown names: no preconditioning, as the reference warns and does


def _bool(v) -> bool:
    if isinstance(v, str):
        return v.strip().lower() in ("true", "t", "1", "yes", "logical true")
    return bool(v)


def from_solver_section(sec: dict, dofs: int = 1) -> LinearSystemKeywords:
    """Reads a Solver section (keys matched case-insensitively); dofs is the
    solver variable's DOFs per node (BILU decoding)."""
    d = {k.strip().lower(): v for k, v in sec.items()}
    k = LinearSystemKeywords()
    k.solver = str(d.get("linear system solver", "iterative")).strip().lower()
    if k.solver != "iterative":
        raise KeywordError("Linear System Solver = %s is outside the accelerated path (iterative)" % k.solver)
    k.method, k.method_name = _method(d.get("linear system iterative method"))
    k.precond = _precond(d.get("linear system preconditioning"), d.get("linear system ilu order"), dofs,
                         _bool(d.get("linear system symmetric ilu", False)))
    if "linear system convergence tolerance" not in d:
        raise KeywordError("Linear System Convergence Tolerance is required (ListGetCReal, IterSolve.src:361)")
    k.tol = float(d["linear system convergence tolerance"])
    if "linear system max iterations" not in d:
        raise KeywordError("Linear System Max Iterations is required (ListGetInteger minv=1, IterSolve.src:349)")
    k.maxit = int(d["linear system max iterations"])
    if k.maxit < 1:
        raise KeywordError("Linear System Max Iterations must be >= 1")
    k.scaling = _bool(d.get("linear system scaling", True))
    k.symmetric = _bool(d.get("linear system symmetric", False))
    k.abort_not_converged = _bool(d.get("linear system abort not converged", True))
    k.residual_output = int(d.get("linear system residual output", 1))
    k.divtol = float(d.get("linear system divergence tolerance", 0.0))
    k.optimize_bandwidth = _bool(d.get("optimize bandwidth", True))
    return k


def parse_sif_solver(text: str) -> dict:
    """Minimal reader for one 'Solver n ... End' block of a .sif file:
    'Keyword = value' lines, quotes stripped, '!' comments dropped."""
    out = {}
    for raw in text.splitlines():
        line = raw.split("!")[0].strip()
        if not line or "=" not in line:
            continue
        key, val = line.split("=", 1)
        key = key.strip()
        val = val.strip().strip('"')
        for typ in ("Real ", "Integer ", "Logical ", "String "):
            if val.startswith(typ):
                val = val[len(typ):].strip().strip('"')
        out[key] = val
    return out
