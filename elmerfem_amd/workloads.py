"""The synthetic workloads of the benchmark and parity suites.

Each builder returns a dict with the mesh, the physics and its per-element
coefficients, the Dirichlet nodes and the Solver-section keywords, so the
GPU path and the oracle consume identical inputs.

  poisson_hex8(n, sigma)      steady 3-D Poisson/heat on n^3 Hex8, k = 1 (sigma = 0)
                              or k_e ~ lognormal(0, sigma) per element, f = 1, T = 0 on all faces
  elasticity_hex(n, quadratic) 3-D isotropic elasticity E = 1, nu = 0.3 on n^3 Hex8 (Hex27),
                              x = 0 clamped, traction (0, 0, -1e-3) on x = 1
  convdiff_quad4(n, seed, i)  steady 2-D convection-diffusion on n^2 Quad4, D = 1, constant
                              velocity of random direction with Peclet 1..100, f = 1, zero Dirichlet
  heat_tet4(n, seed, i)       steady 3-D heat on a jittered n^3 x 6 Tet4 mesh, k_e ~ lognormal(0, 0.5)
"""
from __future__ import annotations

import math

import numpy as np

from . import mesh as M
from .keywords import LinearSystemKeywords

DEFAULT_SEED = M.DEFAULT_SEED


def _kw(method="cg", precond="diagonal", tol=1e-10, maxit=10000):
    from .keywords import from_solver_section
    return from_solver_section({
        "Linear System Solver": "Iterative",
        "Linear System Iterative Method": method,
        "Linear System Preconditioning": precond,
        "Linear System Convergence Tolerance": tol,
        "Linear System Max Iterations": maxit,
        "Linear System Residual Output": 0,
    })


def poisson_hex8(n: int, sigma: float = 0.0, seed: int = DEFAULT_SEED, method="cg") -> dict:
    m = M.hex8_cube(n)
    k = np.ones(m.nelem) if sigma == 0 else M.lognormal(seed, m.nelem, 0.0, sigma)
    return {
        "name": "poisson_hex8_%d" % n, "mesh": m, "physics": 0, "dofs": 1,
        "coef": k, "conv": None, "velo": None, "load": np.ones(m.nelem),
        "bc_nodes": m.boundary_nodes(), "kw": _kw(method),
        "info": {"n": n, "sigma": sigma, "seed": seed},
    }


def poisson_hex27(n: int, sigma: float = 0.0, seed: int = DEFAULT_SEED) -> dict:
    m = M.hex27_cube(n)
    k = np.ones(m.nelem) if sigma == 0 else M.lognormal(seed, m.nelem, 0.0, sigma)
    return {
        "name": "poisson_hex27_%d" % n, "mesh": m, "physics": 0, "dofs": 1,
        "coef": k, "conv": None, "velo": None, "load": np.ones(m.nelem),
        "bc_nodes": m.boundary_nodes(), "kw": _kw("cg"),
        "info": {"n": n, "sigma": sigma, "seed": seed},
    }


def elasticity_hex(n: int, quadratic: bool = False) -> dict:
    m = M.hex27_cube(n) if quadratic else M.hex8_cube(n)
    return {
        "name": "elasticity_%s_%d" % ("hex27" if quadratic else "hex8", n), "mesh": m, "physics": 1, "dofs": 3,
        "coef": np.ones(m.nelem), "poisson": 0.3,
        "bc_nodes": m.boundary_nodes([1]),          # x = 0 clamped, all three components
        "traction_bnd": 2, "traction": np.array([0.0, 0.0, -1e-3]),
        "kw": _kw("cg"),
        "info": {"n": n, "E": 1.0, "nu": 0.3},
    }


def convdiff_quad4(n: int, seed: int = DEFAULT_SEED, index: int = 1) -> dict:
    m = M.quad4_square(n)
    u1, u2 = M.uniforms(seed, np.array([index], dtype=np.uint64), stream=7)
    theta = 2.0 * math.pi * float(u2[0])
    pe = 1.0 + 99.0 * float(u1[0] if u1[0] < 1.0 else 0.0)
    # Peclet number on the unit square with D = 1: |a| * L / D
    a = pe * np.array([math.cos(theta), math.sin(theta), 0.0])
    return {
        "name": "convdiff_quad4_%d" % n, "mesh": m, "physics": 0, "dofs": 1,
        "coef": np.ones(m.nelem), "conv": np.ones(m.nelem), "velo": np.tile(a, (m.nelem, 1)),
        "load": np.ones(m.nelem), "bc_nodes": m.boundary_nodes(), "kw": _kw("bicgstab"),
        "info": {"n": n, "peclet": pe, "theta": theta},
    }


def heat_tet4(n: int, seed: int = DEFAULT_SEED, index: int = 1, jitter: float = 0.2) -> dict:
    m = M.tet4_cube(n, jitter=jitter, seed=seed + index)
    k = M.lognormal(seed + index, m.nelem, 0.0, 0.5, stream=3)
    return {
        "name": "heat_tet4_%d" % n, "mesh": m, "physics": 0, "dofs": 1,
        "coef": k, "conv": None, "velo": None, "load": np.ones(m.nelem),
        "bc_nodes": m.boundary_nodes(), "kw": _kw("cg"),
        "info": {"n": n, "seed": seed, "index": index},
    }


def heat_tet10(n: int, seed: int = DEFAULT_SEED, index: int = 1, jitter: float = 0.2) -> dict:
    m = M.tet10_cube(n, jitter=jitter, seed=seed + index)
    k = M.lognormal(seed + index, m.nelem, 0.0, 0.5, stream=3)
    return {
        "name": "heat_tet10_%d" % n, "mesh": m, "physics": 0, "dofs": 1,
        "coef": k, "conv": None, "velo": None, "load": np.ones(m.nelem),
        "bc_nodes": m.boundary_nodes(), "kw": _kw("cg"),
        "info": {"n": n, "seed": seed, "index": index},
    }


def elasticity_tet10(n: int) -> dict:
    m = M.tet10_cube(n, jitter=0.0)
    return {
        "name": "elasticity_tet10_%d" % n, "mesh": m, "physics": 1, "dofs": 3,
        "coef": np.ones(m.nelem), "poisson": 0.3,
        "bc_nodes": m.boundary_nodes([1]), "kw": _kw("cg"),
        "info": {"n": n, "E": 1.0, "nu": 0.3},
    }


def build_system(w: dict, optimize_bandwidth: bool = False):
    """Creates the device system, assembles and applies the Dirichlet rows.
    Returns (system, assembly ms)."""
    from .system import System
    m = w["mesh"]
    s = System.from_mesh(m, dofs=w["dofs"], optimize_bandwidth=optimize_bandwidth)
    t = s.assemble(w["physics"], m.coords, elem_coef=w["coef"], elem_conv=w.get("conv"), elem_velo=w.get("velo"),
                   elem_load=w.get("load"), poisson=w.get("poisson", 0.3))
    if w.get("traction_bnd"):
        sel = m.bnd_ids == w["traction_bnd"]
        s.assemble_traction(m.bnd_code, m.coords, m.bnd_elements[sel], w["traction"])
    s.dirichlet_nodes(w["bc_nodes"], symmetric=w["kw"].symmetric)
    return s, t
