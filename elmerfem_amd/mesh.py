"""Structured meshes and per-element random fields for the solver workloads.

Conventions follow what Elmer reads from an ElmerGrid mesh
(fem/src/MeshUtils.src:313-1009): 1-based node numbers, bulk elements first,
element-local node order of fem/src/elements.def (404 Quad4, 409 Quad9,
504 Tet4, 808 Hex8, 827 Hex27), and boundary elements tagged with a
boundary id.  Unit cube [0,1]^3 (unit square in 2-D), lexicographic node
numbering with x fastest; boundary ids x0=1, x1=2, y0=3, y1=4, z0=5, z1=6.

Per-element random coefficients use a counter-based generator keyed on
(seed, 1-based element index): splitmix64 -> two uniforms -> Box-Muller.
The same generator gives identical values on every host, so an element's
coefficient does not depend on how the mesh is traversed.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

DEFAULT_SEED = 20261202

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniforms(seed: int, index: np.ndarray, stream: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """Two independent uniforms per index: u1 in (0, 1], u2 in [0, 1)."""
    idx = np.asarray(index, dtype=np.uint64)
    s = _splitmix64(np.full(idx.shape, np.uint64(seed) ^ np.uint64(stream * 0x5851F42D4C957F2D & 0xFFFFFFFFFFFFFFFF), dtype=np.uint64))
    with np.errstate(over="ignore"):
        h1 = _splitmix64(s ^ _splitmix64(np.uint64(2) * idx))
        h2 = _splitmix64(s ^ _splitmix64(np.uint64(2) * idx + np.uint64(1)))
    scale = 1.0 / 9007199254740992.0  # 2**-53
    u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 1.0) * scale
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * scale
    return u1, u2


def lognormal(seed: int, count: int, mu: float = 0.0, sigma: float = 1.0, stream: int = 0) -> np.ndarray:
    """k_e = exp(mu + sigma * N(0,1)) for elements e = 1..count."""
    u1, u2 = uniforms(seed, np.arange(1, count + 1, dtype=np.uint64), stream)
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return np.exp(mu + sigma * z)


@dataclass
class Mesh:
    """A mesh in Elmer's terms: coordinates, bulk and boundary elements."""

    dim: int
    coords: np.ndarray          # (nnodes, 3) float64
    elem_code: int              # elements.def type code of the bulk elements
    elements: np.ndarray        # (nelem, npe) int32, 1-based node ids
    bnd_code: int               # type code of the boundary elements
    bnd_elements: np.ndarray    # (nbnd, npe_b) int32, 1-based
    bnd_ids: np.ndarray         # (nbnd,) int32 boundary id per boundary element
    info: dict = field(default_factory=dict)

    @property
    def nnodes(self) -> int:
        return int(self.coords.shape[0])

    @property
    def nelem(self) -> int:
        return int(self.elements.shape[0])

    def boundary_nodes(self, ids=None) -> np.ndarray:
        """Sorted unique 1-based nodes of the boundary elements with the given ids."""
        sel = np.ones(len(self.bnd_ids), bool) if ids is None else np.isin(self.bnd_ids, list(ids))
        nodes = np.unique(self.bnd_elements[sel].ravel())
        return nodes[nodes > 0].astype(np.int32)  # rows of smaller element types are 0-padded


def _grid3(nx: int) -> np.ndarray:
    """Node ids (1-based) of an (nx+1)^3 lexicographic grid, indexed [k, j, i]."""
    m = nx + 1
    return (np.arange(m ** 3, dtype=np.int64).reshape(m, m, m) + 1).astype(np.int32)


def _coords3(m: int, h: float) -> np.ndarray:
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    return np.stack([i.ravel() * h, j.ravel() * h, k.ravel() * h], axis=1).astype(np.float64)


def hex8_cube(n: int) -> Mesh:
    """n^3 Hex8 elements on the unit cube with Quad4 boundary faces."""
    g = _grid3(n)
    c = _coords3(n + 1, 1.0 / n)
    # elements.def 808 local order: (-1,-1,-1) (1,-1,-1) (1,1,-1) (-1,1,-1), then w = +1
    n0 = g[:-1, :-1, :-1]
    n1 = g[:-1, :-1, 1:]
    n2 = g[:-1, 1:, 1:]
    n3 = g[:-1, 1:, :-1]
    n4 = g[1:, :-1, :-1]
    n5 = g[1:, :-1, 1:]
    n6 = g[1:, 1:, 1:]
    n7 = g[1:, 1:, :-1]
    el = np.stack([a.ravel() for a in (n0, n1, n2, n3, n4, n5, n6, n7)], axis=1).astype(np.int32)
    faces, ids = _quad_faces(g, 1)
    return Mesh(3, c, 808, el, 404, faces, ids, {"n": n, "kind": "hex8"})


def _quad_faces(g: np.ndarray, step: int, order9: bool = False):
    """Boundary quads of a structured brick grid g[k, j, i]; step = 1 (Quad4) or 2 (Quad9)."""
    out, ids = [], []
    m = g.shape[0]
    last = m - 1

    def quads(face2d):
        # face2d indexed [b, a]; produce 4- or 9-node quads, a fastest
        if step == 1:
            q = [face2d[:-1, :-1], face2d[:-1, 1:], face2d[1:, 1:], face2d[1:, :-1]]
        else:
            s = slice
            q = [face2d[0:-2:2, 0:-2:2], face2d[0:-2:2, 2::2], face2d[2::2, 2::2], face2d[2::2, 0:-2:2],
                 face2d[0:-2:2, 1:-1:2], face2d[1:-1:2, 2::2], face2d[2::2, 1:-1:2], face2d[1:-1:2, 0:-2:2],
                 face2d[1:-1:2, 1:-1:2]]
        return np.stack([a.ravel() for a in q], axis=1)

    planes = [
        (1, g[:, :, 0]), (2, g[:, :, last]),      # x = 0, x = 1  ([k, j])
        (3, g[:, 0, :]), (4, g[:, last, :]),      # y = 0, y = 1  ([k, i])
        (5, g[0, :, :]), (6, g[last, :, :]),      # z = 0, z = 1  ([j, i])
    ]
    for bid, f in planes:
        q = quads(f)
        out.append(q)
        ids.append(np.full(len(q), bid, np.int32))
    return np.concatenate(out).astype(np.int32), np.concatenate(ids)


# elements.def 827 reference coordinates of the 27 nodes
_H27 = np.array([
    [-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
    [0, -1, -1], [1, 0, -1], [0, 1, -1], [-1, 0, -1], [-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0],
    [0, -1, 1], [1, 0, 1], [0, 1, 1], [-1, 0, 1], [0, -1, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0],
    [0, 0, -1], [0, 0, 1], [0, 0, 0]], dtype=np.int64)


def hex27_cube(n: int) -> Mesh:
    """n^3 Hex27 elements (quadratic Lagrange) with Quad9 boundary faces."""
    m = 2 * n + 1
    g = (np.arange(m ** 3, dtype=np.int64).reshape(m, m, m) + 1).astype(np.int32)
    c = _coords3(m, 1.0 / (2 * n))
    ez, ey, ex = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    cols = []
    for (u, v, w) in _H27:
        cols.append(g[2 * ez + w + 1, 2 * ey + v + 1, 2 * ex + u + 1])
    el = np.stack(cols, axis=1).astype(np.int32)
    faces, ids = _quad_faces(g, 2)
    return Mesh(3, c, 827, el, 409, faces, ids, {"n": n, "kind": "hex27"})


def quad4_square(n: int) -> Mesh:
    """n^2 Quad4 elements on the unit square with Line2 boundary edges (ids 1..4)."""
    m = n + 1
    g = (np.arange(m * m, dtype=np.int64).reshape(m, m) + 1).astype(np.int32)
    j, i = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    c = np.zeros((m * m, 3))
    c[:, 0] = i.ravel() / n
    c[:, 1] = j.ravel() / n
    el = np.stack([g[:-1, :-1].ravel(), g[:-1, 1:].ravel(), g[1:, 1:].ravel(), g[1:, :-1].ravel()], axis=1)
    segs, ids = [], []
    for bid, line in ((1, g[:, 0]), (2, g[:, -1]), (3, g[0, :]), (4, g[-1, :])):
        s = np.stack([line[:-1], line[1:]], axis=1)
        segs.append(s)
        ids.append(np.full(len(s), bid, np.int32))
    return Mesh(2, c, 404, el.astype(np.int32), 202, np.concatenate(segs).astype(np.int32),
                np.concatenate(ids), {"n": n, "kind": "quad4"})


# Kuhn split of the unit cube along its main diagonal: the 6 axis orders
_KUHN = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def tet4_cube(n: int, jitter: float = 0.2, seed: int = DEFAULT_SEED) -> Mesh:
    """n^3 hexes, each split into 6 Tet4 (conforming Kuhn split), with the
    interior nodes jittered uniformly by +/- jitter*h per coordinate.

    Tetrahedra are ordered so that each has positive signed volume; a
    jitter that would invert one is rejected with ValueError.
    """
    g = _grid3(n)
    h = 1.0 / n
    c = _coords3(n + 1, h)
    if jitter > 0:
        m = n + 1
        k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
        interior = ((i > 0) & (i < n) & (j > 0) & (j < n) & (k > 0) & (k < n)).ravel()
        ids = np.arange(1, m ** 3 + 1, dtype=np.uint64)
        for d in range(3):
            u1, _ = uniforms(seed, ids, stream=1 + d)
            off = (2.0 * (u1 - 0.5)) * jitter * h
            c[interior, d] += off[interior]
    ez, ey, ex = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    tets = []
    for order in _KUHN:
        v = np.zeros((len(ex), 3), np.int64)
        pts = [v.copy()]
        for ax in order:
            v[:, ax] += 1
            pts.append(v.copy())
        nodes = [g[ez + p[:, 2], ey + p[:, 1], ex + p[:, 0]] for p in pts]
        tets.append(np.stack(nodes, axis=1))
    # hex-major order: all 6 tets of hex 0, then hex 1, ...
    el = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    x = c[el - 1]
    vol = np.einsum("ij,ij->i", np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]), x[:, 3] - x[:, 0])
    neg = vol < 0
    el[neg, 2], el[neg, 3] = el[neg, 3].copy(), el[neg, 2].copy()
    x = c[el - 1]
    vol = np.einsum("ij,ij->i", np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]), x[:, 3] - x[:, 0])
    if np.any(vol <= 0):
        raise ValueError("jitter inverted a tetrahedron")
    # boundary: triangles of the boundary quads (two per quad)
    quads, qids = _quad_faces(g, 1)
    tri = np.concatenate([quads[:, [0, 1, 2]], quads[:, [0, 2, 3]]])
    tid = np.concatenate([qids, qids])
    return Mesh(3, c, 504, el, 303, tri.astype(np.int32), tid.astype(np.int32),
                {"n": n, "kind": "tet4", "jitter": jitter, "seed": seed})


def tet10_cube(n: int, jitter: float = 0.2, seed: int = DEFAULT_SEED) -> Mesh:
    """tet4_cube with a node at the middle of every edge: Tet10 elements in
    elements.def 510 order (vertices, then mid-edges (1,2) (2,3) (3,1) (1,4)
    (2,4) (3,4)); boundary Tri6 (306: vertices, then (1,2) (2,3) (3,1)) on the
    faces that belong to one tetrahedron, with the cube-face ids of tet4_cube."""
    t4 = tet4_cube(n, jitter=jitter, seed=seed)
    el4 = t4.elements.astype(np.int64)
    pairs = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    e = np.concatenate([np.sort(el4[:, [a, b]], axis=1) for a, b in pairs])
    uniq, inv = np.unique(e, axis=0, return_inverse=True)
    mid = t4.nnodes + 1 + inv.reshape(len(pairs), -1).T          # (nelem, 6) mid-edge node ids
    coords = np.vstack([t4.coords, 0.5 * (t4.coords[uniq[:, 0] - 1] + t4.coords[uniq[:, 1] - 1])])
    el = np.hstack([el4, mid]).astype(np.int32)
    edge_id = {(int(a), int(b)): t4.nnodes + 1 + k for k, (a, b) in enumerate(uniq)}
    faces = {}
    for t in el4:
        for f in ((t[0], t[1], t[2]), (t[0], t[1], t[3]), (t[0], t[2], t[3]), (t[1], t[2], t[3])):
            key = tuple(sorted(int(v) for v in f))
            faces[key] = faces.get(key, 0) + 1
    bnd, ids = [], []
    for (a, b, c), cnt in faces.items():
        if cnt != 1:
            continue
        tri = [a, b, c]
        m = [edge_id[tuple(sorted((tri[i], tri[(i + 1) % 3])))] for i in range(3)]
        bnd.append(tri + m)
        x = t4.coords[np.array(tri) - 1]
        for d, (lo, hi) in enumerate(((1, 2), (3, 4), (5, 6))):
            if np.all(x[:, d] == 0.0):
                ids.append(lo)
                break
            if np.all(x[:, d] == 1.0):
                ids.append(hi)
                break
        else:
            raise ValueError("boundary face off the cube faces")
    return Mesh(3, coords, 510, el, 306, np.array(bnd, np.int32), np.array(ids, np.int32),
                {"n": n, "kind": "tet10", "jitter": jitter, "seed": seed})
