"""The C ABI without a GPU: libehip.so loads and exports every function
include/ehip.h declares, the ctypes structs match the C layout, and the
Fortran interface module declares every entry point it binds with the C
name the header gives.  No compute call is made here."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from elmerfem_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ehip.h")
FORTRAN = os.path.join(ROOT, "elmerfem_amd", "fortran", "ehip.F90")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ehip_[a-z0-9_]+)\s*\(", text)))


def test_header_parses():
    names = declared_functions()
    assert "ehip_matvec" in names and "ehip_solve" in names and "ehip_server_step" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)
    _lib.load(require_all=True)


def test_struct_layout_matches_c():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "ehip.h"
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  printf("ehip_params %zu\nehip_report %zu\n", sizeof(ehip_params), sizeof(ehip_report));
  P(ehip_params, tol) P(ehip_params, inner_tol) P(ehip_params, divtol) P(ehip_params, profile_every)
  P(ehip_report, res_scaled) P(ehip_report, spmv_us) P(ehip_report, spmv_precision)
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-I" + os.path.join(ROOT, "include"), "-o", exe, c])
        out = subprocess.check_output([exe], text=True).split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    assert int(got["ehip_params"]) == C.sizeof(_lib.Params)
    assert int(got["ehip_report"]) == C.sizeof(_lib.Report)
    for key, val in got.items():
        if "." in key:
            struct, field = key.split(".")
            cls = _lib.Params if struct == "ehip_params" else _lib.Report
            assert getattr(cls, field).offset == int(val), key


def test_fortran_interface_names():
    text = open(FORTRAN).read().lower()
    bound = set(re.findall(r'bind\(c,\s*name="(ehip_[a-z0-9_]+)"\)', text))
    assert bound, "no bind(C) interfaces found"
    assert bound <= set(declared_functions())
    for must in ("ehip_matvec", "ehip_solve_csr", "ehip_solve", "ehip_system_create", "ehip_assemble"):
        assert must in bound


@pytest.mark.skipif(not os.path.exists("/opt/rocm/lib/llvm/bin/flang"), reason="flang absent")
def test_fortran_struct_layout():
    """The bind(C) derived types have the C sizes (flang lays them out)."""
    prog = """
program l
  use ehip
  type(ehip_params) :: p
  type(ehip_report) :: r
  print '(i0,1x,i0)', c_sizeof(p), c_sizeof(r)
end program l
"""
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "l.F90")
        exe = os.path.join(d, "l")
        open(f, "w").write(prog)
        libdir = os.path.dirname(_lib.LIB_PATH)
        subprocess.check_call(["/opt/rocm/lib/llvm/bin/flang", "-J", d, "-o", exe, FORTRAN, f, "-L" + libdir,
                               "-lehip", "-Wl,-rpath," + libdir], cwd=d)
        a, b = map(int, subprocess.check_output([exe], text=True).split())
    assert a == C.sizeof(_lib.Params) and b == C.sizeof(_lib.Report)


def test_schedule_table_host_logic():
    """The multi-system schedule table is host state: size classes by
    nonzeros, defaults, per-form keys, and argument checks."""
    L = _lib.load()
    p = _lib.Params()
    p.method, p.precision, p.mode = _lib.BICGSTAB, _lib.FP64, _lib.FAST
    nb, grid = C.c_int(), C.c_int()

    def get(nnz, q=p):
        assert L.ehip_batch_get_schedule(C.byref(q), nnz, C.byref(nb), C.byref(grid)) == 0
        return nb.value, grid.value

    assert get(100) == (2048, 2048)
    assert L.ehip_batch_set_schedule(C.byref(p), (1 << 17) + 5, 4096, 1024) == 0
    # same class: [2^17, 2^19); neighbours keep the default chunk size, the grid is per form
    assert get(1 << 17) == (4096, 1024) and get((1 << 19) - 1) == (4096, 1024)
    assert get((1 << 17) - 1) == (2048, 1024) and get(1 << 19) == (2048, 1024)
    assert get(1 << 40) == (2048, 1024)
    q = _lib.Params()
    q.method, q.precision, q.mode = _lib.BICGSTAB, _lib.FP64, _lib.FAITHFUL   # another form
    assert get(1 << 17, q) == (2048, 2048)
    assert L.ehip_batch_set_schedule(C.byref(p), 1 << 17, -1, -1) == 0      # unchanged
    assert get(1 << 17) == (4096, 1024)
    for bad_nb, bad_grid in ((3000, 0), (0, 12), (0, -2), (8192, 0), (0, 1 << 20)):
        assert L.ehip_batch_set_schedule(C.byref(p), 1, bad_nb, bad_grid) == _lib.EARG
    assert L.ehip_batch_set_schedule(C.byref(p), 1 << 17, 0, 0) == 0        # back to defaults
    assert get(1 << 17) == (2048, 2048)
    p.method = 7
    assert L.ehip_batch_get_schedule(C.byref(p), 1, C.byref(nb), C.byref(grid)) == _lib.EARG


@pytest.mark.skipif(not os.path.exists("/opt/rocm/lib/llvm/bin/flang"), reason="flang absent")
def test_before_linsolve_shim_compiles():
    """The Before-Linsolve shim builds against the ISO_C_BINDING module and
    the stand-in for Elmer's DefUtils (it runs in the GPU tests)."""
    import fortran_shim
    with tempfile.TemporaryDirectory() as d:
        exe = fortran_shim.build(d)
        assert os.path.exists(exe)
