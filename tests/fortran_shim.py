"""Builds the Before-Linsolve shim (elmerfem_amd/fortran/ehip_linsolve.F90)
with the test stand-in for Elmer's DefUtils and the test driver; shared by
the CPU compile test and the GPU run."""
import os
import subprocess

from elmerfem_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


def build(outdir):
    libdir = os.path.dirname(_lib.LIB_PATH)
    exe = os.path.join(outdir, "linsolve_driver")
    srcs = [os.path.join(ROOT, "tests", "fortran", "elmer_defutils_mock.F90"),
            os.path.join(ROOT, "elmerfem_amd", "fortran", "ehip.F90"),
            os.path.join(ROOT, "elmerfem_amd", "fortran", "ehip_linsolve.F90"),
            os.path.join(ROOT, "tests", "fortran", "linsolve_driver.F90")]
    subprocess.check_call([FLANG, "-O2", "-J", outdir, "-o", exe] + srcs +
                          ["-L" + libdir, "-lehip", "-Wl,-rpath," + libdir], cwd=outdir)
    return exe


def laplacian(m):
    """The driver's matrix: 1-based CRS of the m x m 5-point Laplacian."""
    import numpy as np
    rows, cols, vals, diag = [1], [], [], []
    for j in range(1, m + 1):
        for i in range(1, m + 1):
            r = (j - 1) * m + i
            ent = []
            if j > 1: ent.append((r - m, -1.0))
            if i > 1: ent.append((r - 1, -1.0))
            ent.append((r, 4.0))
            if i < m: ent.append((r + 1, -1.0))
            if j < m: ent.append((r + m, -1.0))
            for c, v in ent:
                cols.append(c)
                vals.append(v)
                if c == r:
                    diag.append(len(cols))
            rows.append(len(cols) + 1)
    return (np.array(rows, np.int32), np.array(cols, np.int32), np.array(diag, np.int32), np.array(vals))
