"""The small systems whose HUTI runs are stored as golden vectors."""
import numpy as np

import oracle_lib as O
from elmerfem_amd import workloads as W

CASES = {
    "cg_poisson8": {"wl": ("poisson", 8, 0.0), "method": 0, "precond": 1, "tol": 1e-10, "maxit": 1000},
    "cg_lognormal8": {"wl": ("poisson", 8, 1.0), "method": 0, "precond": 1, "tol": 1e-10, "maxit": 1000},
    "cg_none_lognormal8": {"wl": ("poisson", 8, 1.0), "method": 0, "precond": 0, "tol": 1e-10, "maxit": 1000},
    "cg_poisson32": {"wl": ("poisson", 32, 0.0), "method": 0, "precond": 1, "tol": 1e-10, "maxit": 2000},
    "bicgstab_convdiff16": {"wl": ("convdiff", 16), "method": 1, "precond": 1, "tol": 1e-10, "maxit": 1000},
    "bicgstab_poisson8": {"wl": ("poisson", 8, 1.0), "method": 1, "precond": 1, "tol": 1e-10, "maxit": 1000},
    "cg_elastic3": {"wl": ("elastic", 3), "method": 0, "precond": 1, "tol": 1e-10, "maxit": 2000},
    "cg_tet4_4": {"wl": ("tet4", 4), "method": 0, "precond": 1, "tol": 1e-10, "maxit": 1000},
    "cg_maxit_lognormal8": {"wl": ("poisson", 8, 1.0), "method": 0, "precond": 1, "tol": 1e-10, "maxit": 7},
    # ILU(k) preconditioning: precond = 2 + k
    "cg_ilu0_lognormal8": {"wl": ("poisson", 8, 1.0), "method": 0, "precond": 2, "tol": 1e-10, "maxit": 1000},
    "cg_ilu1_poisson8": {"wl": ("poisson", 8, 0.0), "method": 0, "precond": 3, "tol": 1e-10, "maxit": 1000},
    "cg_ilu0_poisson32": {"wl": ("poisson", 32, 0.0), "method": 0, "precond": 2, "tol": 1e-10, "maxit": 2000},
    "cg_ilu0_elastic3": {"wl": ("elastic", 3), "method": 0, "precond": 2, "tol": 1e-10, "maxit": 2000},
    "bicgstab_ilu0_convdiff16": {"wl": ("convdiff", 16), "method": 1, "precond": 2, "tol": 1e-10, "maxit": 1000},
    "bicgstab_ilu2_convdiff16": {"wl": ("convdiff", 16), "method": 1, "precond": 4, "tol": 1e-10, "maxit": 1000},
}


def workload(spec):
    kind = spec["wl"][0]
    if kind == "poisson":
        return W.poisson_hex8(spec["wl"][1], sigma=spec["wl"][2])
    if kind == "convdiff":
        return W.convdiff_quad4(spec["wl"][1])
    if kind == "elastic":
        return W.elasticity_hex(spec["wl"][1])
    return W.heat_tet4(spec["wl"][1])


def assembled_case(spec):
    """Oracle pattern, assembly and Dirichlet rows of a case (unscaled)."""
    w = workload(spec)
    m = w["mesh"]
    pat = O.Pattern(m.elements, m.nnodes, w["dofs"])
    vals, rhs = O.assemble(w["physics"], m, pat, w["coef"], conv=w.get("conv"), velo=w.get("velo"),
                           load=w.get("load"), poisson=w.get("poisson", 0.3))
    if w.get("traction_bnd"):
        sel = m.bnd_ids == w["traction_bnd"]
        O.assemble_traction(m, pat, sel, w["traction"], rhs)
    k = pat.perm[w["bc_nodes"] - 1].astype(np.int64)
    rows = np.concatenate([w["dofs"] * (k - 1) + d for d in range(1, w["dofs"] + 1)]).astype(np.int32)
    O.dirichlet(pat, vals, rhs, rows)
    return w, pat, vals, rhs, rows


def scaled_case(spec):
    w, pat, vals, rhs, rows = assembled_case(spec)
    v, b, x0 = O.scaled_system(pat, vals, rhs)
    return pat, v, b, x0
