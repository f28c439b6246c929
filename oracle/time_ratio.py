"""Times the oracle's HUTI CG (the C restatement bench.py uses as the CPU
baseline) against the reference's own HUTI CG and BLAS compiled into
oracle/_ref, on the same scaled system and thread count (the row SpMV is
row-parallel in both).  Test infrastructure: run where the reference tree
and oracle/_ref exist.

    python oracle/time_ratio.py [--n 64] [--threads 8]
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle_lib as O  # noqa: E402
from elmerfem_amd import workloads as W  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    if not O.ref_available():
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    os.environ["OMP_NUM_THREADS"] = str(a.threads)
    O.lib().or_set_threads(a.threads)
    w = W.poisson_hex8(a.n, sigma=1.0)
    m = w["mesh"]
    pat = O.Pattern(m.elements, m.nnodes, 1, optimize=True, extra_elements=[m.bnd_elements])
    vals, rhs = O.assemble(0, m, pat, w["coef"], load=w["load"])
    O.dirichlet(pat, vals, rhs, pat.perm[w["bc_nodes"] - 1])
    v, b, x0 = O.scaled_system(pat, vals, rhs)
    best = {}
    for name, fn in (("port", O.huti_solve), ("reference", O.ref_huti_solve)):
        ts = []
        for _ in range(a.reps):
            t = time.perf_counter()
            info, k, x, hist = fn(pat.rows, pat.cols, pat.diag, v, b, x0, method=0, precond=1, tol=1e-10,
                                  maxit=20000)
            ts.append(time.perf_counter() - t)
        best[name] = (min(ts), k)
        print("%-9s %d iterations, best of %d: %.3f s" % (name, k, a.reps, min(ts)))
    print("port / reference time ratio: %.3f (%d^3 Hex8 lognormal CG + Jacobi to 1e-10, %d threads)"
          % (best["port"][0] / best["reference"][0], a.n, a.threads))


if __name__ == "__main__":
    main()
