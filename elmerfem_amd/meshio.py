"""Mesh and result files: ElmerGrid mesh directories (mesh.header, mesh.nodes,
mesh.elements, mesh.boundary) as LoadMesh reads them
(fem/src/MeshUtils.src:313-1009), and SaveResult's ASCII result format
(fem/src/ModelDescription.src:2442-2865).  The parsing and writing run in
libehip (csrc/meshio.cpp); this module moves the arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .mesh import Mesh


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def read_mesh(directory: str) -> Mesh:
    """Reads an ElmerGrid mesh directory.  Nodes are numbered 1..N in file
    order; every bulk element must have the same type.  Boundary elements
    keep their own types in info["bnd_types"] (rows of smaller types are
    0-padded in bnd_elements) and their parents in info["bnd_parents"]."""
    L = _lib.load()
    h = C.c_void_p()
    _lib.check(L.ehip_mesh_read(directory.encode(), C.byref(h)), "ehip_mesh_read")
    try:
        nn, ne, nb, mb, mf = (C.c_int() for _ in range(5))
        _lib.check(L.ehip_mesh_sizes(h, C.byref(nn), C.byref(ne), C.byref(nb), C.byref(mb), C.byref(mf)),
                   "ehip_mesh_sizes")
        nn, ne, nb, mb, mf = nn.value, ne.value, nb.value, mb.value, mf.value
        xyz = np.zeros((nn, 3))
        bt, bb = np.zeros(ne, np.int32), np.zeros(ne, np.int32)
        bn = np.zeros((ne, mb), np.int32)
        ft, fg = np.zeros(nb, np.int32), np.zeros(nb, np.int32)
        fp = np.zeros((nb, 2), np.int32)
        fn = np.zeros((nb, max(mf, 1)), np.int32)
        _lib.check(L.ehip_mesh_get(h, _p(xyz), _p(bt), _p(bb), _p(bn), _p(ft), _p(fg), _p(fp), _p(fn)),
                   "ehip_mesh_get")
    finally:
        L.ehip_mesh_free(h)
    types = np.unique(bt)
    if len(types) != 1:
        raise _lib.EhipError("mixed bulk element types %s: one type per mesh on this path" % types.tolist())
    # MeshDim: the number of coordinates that vary (MeshUtils.src:481-485)
    dim = int(sum(np.any(xyz[:, k] != xyz[0, k]) for k in range(3)))
    bcode = int(np.bincount(ft).argmax()) if nb else 0
    return Mesh(dim=dim, coords=xyz, elem_code=int(types[0]), elements=bn, bnd_code=bcode,
                bnd_elements=fn[:, :max(mf, 1)], bnd_ids=fg,
                info={"bodies": bb, "bnd_types": ft, "bnd_parents": fp, "source": directory})


def write_mesh(directory: str, mesh: Mesh, bodies=None, bnd_parents=None) -> None:
    """Writes mesh as an ElmerGrid mesh directory (serial: partition -1)."""
    L = _lib.load()
    el = np.ascontiguousarray(mesh.elements, np.int32)
    be = np.ascontiguousarray(mesh.bnd_elements, np.int32)
    bodies = np.ones(len(el), np.int32) if bodies is None else np.ascontiguousarray(bodies, np.int32)
    ids = np.ascontiguousarray(mesh.bnd_ids, np.int32)
    par = (np.zeros((len(be), 2), np.int32) if bnd_parents is None
           else np.ascontiguousarray(bnd_parents, np.int32))
    xyz = np.ascontiguousarray(mesh.coords, np.float64)
    _lib.check(L.ehip_mesh_write(directory.encode(), len(xyz), _p(xyz), len(el), int(mesh.elem_code), _p(el),
                                 _p(bodies), len(be), int(mesh.bnd_code), _p(be), _p(ids), _p(par)),
               "ehip_mesh_write")


def write_result(path: str, variable: str, values, perm=None, equation: str = "no equation", append=False,
                 step: int = 1, time: float = 0.0) -> None:
    """One scalar nodal variable in SaveResult's ASCII format 3; perm (1-based,
    0 = no value) maps node i to its entry in values, as Variable % Perm."""
    L = _lib.load()
    v = np.ascontiguousarray(values, np.float64)
    nn = len(v) if perm is None else len(perm)
    pp = None if perm is None else np.ascontiguousarray(perm, np.int32)
    _lib.check(L.ehip_result_write(path.encode(), variable.encode(), equation.encode(), nn,
                                   _p(pp) if pp is not None else None, 1, _p(v), int(append), step, time),
               "ehip_result_write")


def read_result(path: str) -> dict:
    """Reads back an ASCII format 3 result file: {variable: [values per step]}
    with values placed by node (NaN where Perm is 0)."""
    out: dict = {}
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    i = 0
    nnodes = None
    while i < len(lines):
        ln = lines[i].strip()
        if ln.startswith("Number Of Nodes:"):
            nnodes = int(ln.split(":")[1])
        if ln.startswith("Time:"):
            name = lines[i + 1].strip()
            perm_line = lines[i + 2].split()
            i += 3
            if perm_line[1] == "NULL":
                idx = np.arange(1, nnodes + 1)
            else:
                cnt = int(perm_line[2])
                pairs = np.array([lines[i + k].split() for k in range(cnt)], dtype=np.int64).reshape(-1, 2)
                i += cnt
                idx = pairs[:, 0]
            vals = np.array([float(lines[i + k].replace("D", "E")) for k in range(len(idx))])
            i += len(idx)
            full = np.full(nnodes, np.nan)
            full[idx - 1] = vals
            out.setdefault(name, []).append(full)
            continue
        i += 1
    return out
