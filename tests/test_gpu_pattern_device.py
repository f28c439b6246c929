"""The device build of the CRS pattern and the row-owner assembly maps
(the default) against the host build (EHIP_PATTERN_BUILD=host) and the
oracle: Perm, Rows, Cols, Diag and the bandwidth record bit-exact, and the
assembled Values and RHS -- which go through the row maps -- bit-identical
between the two builds."""
import os

import numpy as np
import pytest

import oracle_lib as O
from elmerfem_amd import mesh as M, workloads as W
from elmerfem_amd.system import System

pytestmark = pytest.mark.gpu


def build(w, opt, host):
    if host:
        os.environ["EHIP_PATTERN_BUILD"] = "host"
    try:
        s = System.from_mesh(w["mesh"], dofs=w["dofs"], optimize_bandwidth=opt)
    finally:
        os.environ.pop("EHIP_PATTERN_BUILD", None)
    m = w["mesh"]
    s.assemble(w["physics"], m.coords, elem_coef=w["coef"], elem_conv=w.get("conv"), elem_velo=w.get("velo"),
               elem_load=w.get("load"), poisson=w.get("poisson", 0.3))
    return s


def scrambled(w, seed=3):
    m = w["mesh"]
    relabel = (np.random.default_rng(seed).permutation(m.nnodes) + 1).astype(np.int32)
    inv = np.empty_like(relabel)
    inv[relabel - 1] = np.arange(m.nnodes)
    m.elements = relabel[m.elements - 1]
    m.bnd_elements = relabel[m.bnd_elements - 1]
    m.coords = m.coords[inv]
    return w


CASES = {
    "hex8_12": lambda: W.poisson_hex8(12, sigma=1.0),
    "hex8_10_scrambled": lambda: scrambled(W.poisson_hex8(10, sigma=1.0)),
    "tet4_8": lambda: W.heat_tet4(8),
    "quad4_40_conv": lambda: W.convdiff_quad4(40),
    "elastic_hex8_5": lambda: W.elasticity_hex(5),
    "elastic_hex27_2": lambda: W.elasticity_hex(2, quadratic=True),
    "tet10_3": lambda: W.heat_tet10(3),
    "hex27_3": lambda: W.poisson_hex27(3),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("opt", [False, True])
def test_device_pattern_equals_host(ehip, case, opt):
    w = CASES[case]()
    print("pattern %s optimize=%d" % (case, opt), flush=True)
    sd = build(w, opt, host=False)
    sh = build(w, opt, host=True)
    for a, b in zip(sd.pattern(), sh.pattern()):
        assert np.array_equal(a, b)
    assert sd.bandwidth == sh.bandwidth
    vd, bd = sd.values()
    vh, bh = sh.values()
    assert np.array_equal(vd, vh) and np.array_equal(bd, bh)
    m = w["mesh"]
    pat = O.Pattern(m.elements, m.nnodes, w["dofs"], optimize=opt, extra_elements=[m.bnd_elements])
    perm, rows, cols, diag = sd.pattern()
    assert np.array_equal(perm, pat.perm) and np.array_equal(rows, pat.rows) and np.array_equal(cols, pat.cols)
    assert np.array_equal(diag, pat.diag)
    assert sd.bandwidth == (pat.bw_before, pat.bw_after, pat.accepted)
    sd.free()
    sh.free()


def test_device_pattern_rejects_bad_meshes(ehip):
    from types import SimpleNamespace
    from elmerfem_amd import _lib
    el = np.array([[1, 2, 3, 4]], np.int32)
    m = SimpleNamespace(elements=el, bnd_elements=np.zeros((0, 4), np.int32), nnodes=5)   # node 5 in no element
    with pytest.raises(_lib.EhipError, match="no element"):
        System.from_mesh(m, dofs=1, optimize_bandwidth=False)
    m = SimpleNamespace(elements=np.array([[1, 2, 3, 9]], np.int32), bnd_elements=np.zeros((0, 4), np.int32),
                        nnodes=4)
    with pytest.raises(_lib.EhipError, match="out of range"):
        System.from_mesh(m, dofs=1, optimize_bandwidth=False)
