"""Writes the HUTI golden vectors in this directory.

Runs the reference's own HUTI CG/BiCGStab and reference-BLAS ddot/dnrm2,
compiled from the reference sources into oracle/_ref (make -C oracle ref),
on the scaled systems IterSolver would see for small workloads, and stores
iteration counts, HUTI_INFO, per-iteration residuals and solutions.

    python tests/golden/make_golden.py [case ...]     # all cases, or the named ones
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as O  # noqa: E402
from golden_cases import CASES, scaled_case  # noqa: E402


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libhuti_ref.so is not built (make -C oracle ref)")
    names = sys.argv[1:] or list(CASES)
    for name in names:
        spec = CASES[name]
        pat, v, b, x0 = scaled_case(spec)
        info, k, x, hist = O.ref_huti_solve(pat.rows, pat.cols, pat.diag, v, b, x0, method=spec["method"],
                                            precond=spec["precond"], tol=spec["tol"], maxit=spec["maxit"])
        keep_x = x if pat.n <= 4000 else x[::97]
        np.savez_compressed(os.path.join(HERE, "huti_%s.npz" % name), info=info, iters=k, hist=hist, x=keep_x,
                            full_x=pat.n <= 4000, n=pat.n, nnz=pat.nnz,
                            bsum=np.sum(b), vsum=np.sum(v))
        print(name, "n", pat.n, "info", info, "iters", k, "final", hist[-1])


if __name__ == "__main__":
    main()
