"""SELL-64 with slice-shared column offsets (EHIP_SPMV_SELL_SHARED).

The layout is built, and made the default of both precisions, on stencil-
like patterns (meshes of regular topology in natural node order).  Bars:
FP64 SpMV bit-exact against the oracle's CRS loop (CRS_MatrixVectorProd,
CRSMatrix.src:3565-3573), also when x holds infinities at columns that are
in a slice's offset list but not in a row's pattern (the per-row mask keeps
them out); FP32 SpMV and whole FP64 / mixed solves give the bits of the
SELL-64 layout; patterns whose slices use more than 32 offsets have no such
layout.
"""
import numpy as np
import pytest

import oracle_lib as O
from elmerfem_amd import _lib, mesh as M, workloads as W
from elmerfem_amd.system import System

pytestmark = pytest.mark.gpu

SHARED, SELL, CSR = _lib.SPMV_SELL_SHARED, _lib.SPMV_SELL, _lib.SPMV_CSR

WORKLOADS = {"poisson_hex8_12": lambda: W.poisson_hex8(12, sigma=1.0),
             "heat_tet4_9": lambda: W.heat_tet4(9),
             "convdiff_quad4_70": lambda: W.convdiff_quad4(70)}


def oracle_spmv(rows, cols, vals, u):
    out = np.zeros(len(rows) - 1)
    O.lib().or_spmv(len(rows) - 1, O._p(rows), O._p(cols), O._p(vals), O._p(np.ascontiguousarray(u)), O._p(out))
    return out


def both(s, fmt):
    s.spmv_set_format(fmt, 0)
    s.spmv_set_format(fmt, 1)


@pytest.mark.parametrize("name", sorted(WORKLOADS))
@pytest.mark.parametrize("optimize", [False, True])
def test_default_on_regular_meshes_and_bit_exact(ehip, name, optimize):
    """Natural order, and the bandwidth optimisation (rejected on these
    meshes): the shared-offset layout is the default, and its FP64 products
    are the reference loop's bits before and after new values."""
    w = WORKLOADS[name]()
    s, _ = W.build_system(w, optimize_bandwidth=optimize)
    for prec in (0, 1):
        cfg = s.spmv_config(prec)
        assert cfg["format"] == "sell_shared" and cfg["index_bits"] == 0, cfg
    _, rows, cols, _ = s.pattern()
    vals, rhs = s.values()
    rng = np.random.default_rng(3)
    u = rng.standard_normal(s.n)
    ref = oracle_spmv(rows, cols, vals, u)
    for fmt in (SHARED, SELL, CSR):
        s.spmv_set_format(fmt)
        assert np.array_equal(s.spmv(u), ref), fmt
    s.spmv_set_format(SHARED)
    vals2 = vals * (1.0 + 0.25 * rng.random(len(vals)))
    s.set_values(vals2, rhs)
    assert np.array_equal(s.spmv(u), oracle_spmv(rows, cols, vals2, u))
    s.free()


def _five_point(nx, ny):
    """5-point Laplacian pattern of an nx x ny grid in natural order (1-based CRS)."""
    rows, cols = [1], []
    for j in range(ny):
        for i in range(nx):
            r = j * nx + i
            c = [r - nx] * (j > 0) + [r - 1] * (i > 0) + [r] + [r + 1] * (i < nx - 1) + [r + nx] * (j < ny - 1)
            cols.extend(k + 1 for k in c)
            rows.append(len(cols) + 1)
    rows = np.array(rows, np.int32)
    cols = np.array(cols, np.int32)
    n = nx * ny
    diag = np.array([rows[i] + np.nonzero(cols[rows[i] - 1:rows[i + 1] - 1] == i + 1)[0][0] for i in range(n)],
                    np.int32)
    return rows, cols, diag


def test_fp64_adds_only_the_rows_own_entries(ehip):
    """Offset +1 is in every slice's list, but the last node of a grid line
    has no right neighbour: with x = inf at the next line's first node, that
    row's FP64 product stays finite and equal to the reference loop's, and
    every other row matches too (infinities where the reference has them)."""
    nx, ny = 40, 30
    rows, cols, diag = _five_point(nx, ny)
    n = nx * ny
    rng = np.random.default_rng(11)
    vals = -rng.random(len(cols))
    vals[diag - 1] = 4.5
    s = System.from_csr(rows, cols, diag)
    s.set_values(vals, np.ones(n))
    assert s.spmv_config(0)["format"] == "sell_shared"
    u = rng.standard_normal(n)
    for line in (0, 7, 28):
        u2 = u.copy()
        u2[(line + 1) * nx] = np.inf  # first node of line + 1
        ref = oracle_spmv(rows, cols, vals, u2)
        got = s.spmv(u2)
        r = line * nx + nx - 1  # last node of the line: not a neighbour of the inf node
        assert np.isfinite(got[r]) and got[r] == ref[r]
        assert np.array_equal(got, ref, equal_nan=True)
    s.free()


@pytest.mark.parametrize("name", ["poisson_hex8_12", "heat_tet4_9"])
@pytest.mark.parametrize("precision,mode", [(_lib.FP64, _lib.FAITHFUL), (_lib.FP64, _lib.FAST),
                                            (_lib.MIXED_IR, _lib.FAST)])
def test_solves_identical_to_sell(ehip, name, precision, mode):
    """Same row sums, same workgroups: a whole solve gives the same
    iterations, replacements and solution bits in both layouts, to the
    unscaled residual bar recomputed on the host."""
    w = WORKLOADS[name]()
    s, _ = W.build_system(w)
    _, rows, cols, _ = s.pattern()
    vals, rhs = s.values()
    out = []
    for fmt in (SHARED, SELL, SHARED):
        s.set_values(vals, rhs)
        both(s, fmt)
        x, rep = s.solve(w["kw"], precision=precision, mode=mode, true_residual_stop=True)
        assert rep["info"] == 1 and rep["res_true"] <= 1e-10, rep
        ax = oracle_spmv(rows, cols, vals, x)
        assert np.linalg.norm(rhs - ax) / np.linalg.norm(rhs) <= 1e-10
        out.append((x, rep["iters"], rep["outer_iters"]))
    for x, it, up in out[1:]:
        assert (it, up) == out[0][1:] and np.array_equal(x, out[0][0])
    s.free()


def test_fp32_products_identical_to_sell(ehip):
    """The FP32 copy (made by a mixed solve) gives the same product bits in
    both layouts: absent entries add +0 to a finite sum."""
    w = WORKLOADS["heat_tet4_9"]()
    s, _ = W.build_system(w)
    s.solve(w["kw"], precision=_lib.MIXED_IR, mode=_lib.FAST)
    u = np.random.default_rng(2).standard_normal(s.n)
    got = {}
    for fmt in (SHARED, SELL, CSR):
        s.spmv_set_format(fmt, 1)
        got[fmt] = s.spmv(u, precision=1)
    assert np.array_equal(got[SHARED], got[SELL]) and np.array_equal(got[SHARED], got[CSR])
    s.free()


@pytest.mark.parametrize("kind", ["scrambled", "hex27_elastic"])
def test_absent_without_a_stencil(ehip, kind):
    """Randomly relabelled nodes (a slice's rows then use far more than 32
    offsets) and Hex27 elasticity (up to 125 neighbour nodes per row;
    node-blocked SELL-64 instead) have no shared-offset layout."""
    if kind == "scrambled":
        m = M.hex8_cube(6)
        relabel = (np.random.default_rng(9).permutation(m.nnodes) + 1).astype(np.int32)
        m.elements, m.bnd_elements = relabel[m.elements - 1], relabel[m.bnd_elements - 1]
        s = System.from_mesh(m, dofs=1, optimize_bandwidth=False)
    else:
        s, _ = W.build_system(W.elasticity_hex(3, quadratic=True))
    assert s.spmv_config(0)["format"] in ("sell", "csr")
    with pytest.raises(_lib.EhipError):
        s.spmv_set_format(SHARED)
    s.free()


@pytest.mark.parametrize("precision,mode,method", [(_lib.FP64, _lib.FAITHFUL, _lib.BICGSTAB),
                                                   (_lib.FP64, _lib.FAST, _lib.BICGSTAB),
                                                   (_lib.MIXED_IR, _lib.FAST, _lib.BICGSTAB),
                                                   (_lib.MIXED_IR, _lib.FAST, _lib.CG)])
def test_batch_engine_identical_to_crs_stream(ehip, precision, mode, method):
    """The multi-system engine sums a slot's rows on its shared-offset layout
    one thread per row, the rows and partial sums the staged CRS-stream path
    gives each thread: whole batched solves -- one slot kept on the CRS
    stream in the first run, every slot in the second -- are identical."""
    from elmerfem_amd.batch import solve_batch
    if method == _lib.BICGSTAB:
        ws = [W.convdiff_quad4(10 + 9 * i, index=i + 1) for i in range(4)]
    else:
        ws = [W.heat_tet4(7 + 2 * i, index=i + 1) for i in range(4)]
    systems = [W.build_system(w)[0] for w in ws]
    assert all(s.spmv_config(0)["format"] == "sell_shared" for s in systems)
    vals = [s.values() for s in systems]
    kw = ws[0]["kw"]
    kw.method, kw.tol, kw.maxit = method, 1e-10, 20000
    out = []
    for run in range(2):
        for i, (s, (v, b)) in enumerate(zip(systems, vals)):
            s.set_values(v, b)
            both(s, CSR if run == 1 or i == 2 else SHARED)
        xs, reps, _ = solve_batch(systems, kw, precision=precision, mode=mode, true_residual_stop=True)
        out.append((xs, [(r["iters"], r["info"], r["outer_iters"]) for r in reps]))
    assert out[0][1] == out[1][1]
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    for s in systems:
        s.free()


def test_node_blocks_default_and_bit_exact(ehip):
    """3-DOF Hex8 elasticity in natural node order: the shared-offset layout
    with one node offset and one mask bit per 3 x 3 block is the default,
    and its FP64 products are the reference loop's bits, as are those of the
    node-blocked SELL-64 and the CSR stream."""
    w = W.elasticity_hex(16)
    s, _ = W.build_system(w)
    assert s.spmv_config(0)["format"] == "sell_shared"
    assert s.spmv_shared_layout()["node_block"] == 3
    _, rows, cols, _ = s.pattern()
    vals, rhs = s.values()
    u = np.random.default_rng(5).standard_normal(s.n)
    ref = oracle_spmv(rows, cols, vals, u)
    for fmt in (SHARED, SELL, CSR):
        s.spmv_set_format(fmt)
        assert np.array_equal(s.spmv(u), ref), fmt
    assert s.spmv_sell_layout()["node_block"] == 3
    s.free()


@pytest.mark.parametrize("precision,mode", [(_lib.FP64, _lib.FAST), (_lib.MIXED_IR, _lib.FAST)])
def test_node_block_solves_identical_to_sell(ehip, precision, mode):
    """Same row sums, same workgroups as the node-blocked SELL-64: the
    elasticity solve gives the same iterations and solution bits."""
    w = W.elasticity_hex(16)
    s, _ = W.build_system(w)
    _, rows, cols, _ = s.pattern()
    vals, rhs = s.values()
    out = []
    for fmt in (SHARED, SELL):
        s.set_values(vals, rhs)
        both(s, fmt)
        x, rep = s.solve(w["kw"], precision=precision, mode=mode, true_residual_stop=True)
        assert rep["info"] == 1 and rep["res_true"] <= 1e-10, rep
        ax = oracle_spmv(rows, cols, vals, x)
        assert np.linalg.norm(rhs - ax) / np.linalg.norm(rhs) <= 1e-10
        out.append((x, rep["iters"], rep["outer_iters"]))
    assert out[0][1:] == out[1][1:] and np.array_equal(out[0][0], out[1][0])
    s.free()


def _kron_five_point(nx, ny, B):
    """The 5-point Laplacian of an nx x ny grid (natural order) times a D x D
    SPD block B, as 1-based CRS with node-interleaved unknowns: SPD."""
    D = B.shape[0]
    rows, cols, vals = [1], [], []
    for j in range(ny):
        for i in range(nx):
            node = j * nx + i
            nb = [(node - nx, -1.0)] * (j > 0) + [(node - 1, -1.0)] * (i > 0) + [(node, 4.2)] + \
                 [(node + 1, -1.0)] * (i < nx - 1) + [(node + nx, -1.0)] * (j < ny - 1)
            for d in range(D):
                for m, lv in nb:
                    for e in range(D):
                        cols.append(D * m + e + 1)
                        vals.append(lv * B[d, e])
                rows.append(len(cols) + 1)
    rows = np.array(rows, np.int32)
    cols = np.array(cols, np.int32)
    n = nx * ny * D
    diag = np.array([rows[r] + np.nonzero(cols[rows[r] - 1:rows[r + 1] - 1] == r + 1)[0][0] for r in range(n)],
                    np.int32)
    return rows, cols, diag, np.array(vals)


def test_two_dof_node_blocks(ehip):
    """2-DOF node blocks (D = 2): the shared-offset layout per block gives the
    reference loop's FP64 bits, and a mixed CG solve is bit-identical to the
    one on the node-blocked SELL-64."""
    B = np.array([[2.0, 0.5], [0.5, 1.0]])
    rows, cols, diag, vals = _kron_five_point(60, 45, B)
    n = len(rows) - 1
    s = System.from_csr(rows, cols, diag)
    s.set_values(vals, np.ones(n))
    assert s.spmv_config(0)["format"] == "sell_shared"
    assert s.spmv_shared_layout()["node_block"] == 2
    u = np.random.default_rng(8).standard_normal(n)
    ref = oracle_spmv(rows, cols, vals, u)
    for fmt in (SHARED, SELL, CSR):
        s.spmv_set_format(fmt)
        assert np.array_equal(s.spmv(u), ref), fmt
    assert s.spmv_sell_layout()["node_block"] == 2
    kw = _lib.Params(0, 1, 1e-10, 5000, 1, _lib.MIXED_IR, _lib.FAST, 0, 0, 0, 0, 0, 0, 1)
    out = []
    for fmt in (SHARED, SELL):
        s.set_values(vals, np.ones(n))
        both(s, fmt)
        x, rep = s.solve(kw)
        assert rep["info"] == 1 and rep["res_true"] <= 1e-10, rep
        out.append((x, rep["iters"], rep["outer_iters"]))
    assert out[0][1:] == out[1][1:] and np.array_equal(out[0][0], out[1][0])
    s.free()


@pytest.mark.parametrize("precision,mode", [(_lib.FP64, _lib.FAST), (_lib.MIXED_IR, _lib.FAST)])
def test_cached_and_streamed_value_loads_same_bits(ehip, monkeypatch, precision, mode):
    """A values copy larger than the Infinity Cache is read with cached loads
    up to a budget of slices and streamed after it.  The budget changes only
    the kind of load, never a sum's order: forced here on a 48^3 system
    (24 MB of FP64 values) to none, a third and all of its slices
    (EHIP_MALL_CACHED_MB 0, 8, unset), products and whole solves (fresh
    systems, so no graph recorded under another budget is replayed) give
    the same bits, the FP64 products the reference loop's."""
    w = W.poisson_hex8(48, sigma=1.0)
    u = np.random.default_rng(5).standard_normal(w["mesh"].nnodes)
    got = []
    for mb in (None, "8", "0"):
        if mb is None:
            monkeypatch.delenv("EHIP_MALL_CACHED_MB", raising=False)
        else:
            monkeypatch.setenv("EHIP_MALL_CACHED_MB", mb)
        s, _ = W.build_system(w)
        assert s.spmv_config(0)["format"] == "sell_shared"
        _, rows, cols, _ = s.pattern()
        vals, rhs = s.values()
        y64 = s.spmv(u)
        assert np.array_equal(y64, oracle_spmv(rows, cols, vals, u[:s.n]))
        x, rep = s.solve(w["kw"], precision=precision, mode=mode, true_residual_stop=True)
        assert rep["info"] == 1 and rep["res_true"] <= 1e-10, rep
        y32 = s.spmv(u, precision=1) if precision == _lib.MIXED_IR else None
        got.append((y64, y32, x, rep["iters"], rep["outer_iters"]))
        s.free()
    for y64, y32, x, it, up in got[1:]:
        assert np.array_equal(y64, got[0][0]) and np.array_equal(x, got[0][2])
        assert (it, up) == got[0][3:]
        if y32 is not None:
            assert np.array_equal(y32, got[0][1])
