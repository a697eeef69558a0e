"""Writes tests/golden/lanczos_ref.json: the reference's Lanczos recurrence on
the baseline workloads, driven by the reference's own mltply.

oracle/_ref/ref_harness (this project's harness linked against the
reference's hot-path sources compiled in place) runs the recurrence of
src/Lanczos_EigenValue.c:102-177 in its `lanczos` mode from the normalised
SplitMix64 vector of the workload's seed, and prints alpha_m and beta_m of
every step.  The tridiagonal eigenvalues are taken here from those numbers
(scipy's eigvalsh_tridiagonal stands in for DSEVvalue -> LAPACK dsyev_, which
the image lacks).  Stored per workload: dim, seed, steps, alpha[1..n],
beta[1..n], and the lowest Ritz values of T_m at every `every`-th step m.

    chain16  (cfg1)  seed 0, 80 steps  (dim 12,870: E0 converged to 1e-15)
    hubbard  (cfg2)  seed 1, steps as given (Hubbard 4x4, dim 165,636,900)
    j1j2     (cfg3)  seed 2, steps as given (J1-J2 8x4, dim 601,080,390)

usage: python tests/golden/gen_lanczos.py <cfg> <steps> [threads]
       python tests/golden/gen_lanczos.py <cfg> --from <ab.txt> [threads]   (record a run's steps so far)
The harness writes ab.txt as it goes; a run cut short keeps the steps done.
Where the lowest two filtered values have stopped moving (1e-13 over the last 5
steps) they are stored as "converged".
"""

from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np
from scipy.linalg import eigvalsh_tridiagonal

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from hphi_amd import lattice as L  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "lanczos_ref.json")
SEEDS = {1: 0, 2: 1, 3: 2}


def ritz(alpha, beta, m, n=4):
    """lowest n eigenvalues of T_m (alpha_1..m on the diagonal, beta_1..m-1 beside it)"""
    if m == 1:
        return [float(alpha[0])]
    w = eigvalsh_tridiagonal(np.asarray(alpha[:m]), np.asarray(beta[:m - 1]))
    return [float(x) for x in w[:n]]


def parse(path):
    a, b, t = [], [], []
    for ln in open(path):
        p = ln.split()
        if len(p) == 4:
            a.append(float(p[1]))
            b.append(float(p[2]))
            t.append(float(p[3]))
    return a, b, t


def main():
    cfg = int(sys.argv[1])
    spec = L.baseline_config(cfg)
    seed = SEEDS[cfg]
    if sys.argv[2] == "--from":
        threads = int(sys.argv[4]) if len(sys.argv) > 4 else None
        a, b, t = parse(sys.argv[3])
    else:
        steps = int(sys.argv[2])
        threads = int(sys.argv[3]) if len(sys.argv) > 3 else None
        work = tempfile.mkdtemp(prefix=f"lanczos_cfg{cfg}_", dir=os.environ.get("LANCZOS_WORK"))
        d = os.path.join(work, "def")
        L.write_expert(spec, d)
        env = dict(os.environ)
        if threads:
            env["OMP_NUM_THREADS"] = str(threads)
        print(f"cfg{cfg}: dim {spec.dim}, {steps} steps, work {work}", flush=True)
        r = subprocess.run([O.REF_HARNESS, d, "lanczos", os.path.join(work, "out"), str(seed), str(steps)], env=env)
        a, b, t = parse(os.path.join(work, "out", "ab.txt"))
        if not a:
            raise SystemExit(f"no steps done (rc {r.returncode})")
    every = 1 if len(a) <= 100 else 5
    # the lowest distinct values once they stop moving (Cullum-Willoughby
    # filtered, as the device's distinct mode reports them): converged when the
    # filtered values of the last step and of 5 steps before agree to 1e-13
    conv = None
    if len(a) > 10:
        from hphi_amd import capi
        import ctypes as C
        def good(m):
            out = np.zeros(16)
            aa, bb = np.array(a[:m]), np.array(b[:m])
            k = capi.lib().hphi_lanczos_good_eigenvalues(aa.ctypes.data_as(C.POINTER(C.c_double)),
                                                         bb.ctypes.data_as(C.POINTER(C.c_double)), m, 4,
                                                         out.ctypes.data_as(C.POINTER(C.c_double)))
            return out[:k]
        g1, g0 = good(len(a)), good(len(a) - 5)
        if len(g1) >= 2 and len(g0) >= 2 and np.max(np.abs(g1[:2] - g0[:2]) / np.abs(g1[:2])) < 1e-13:
            conv = [float(x) for x in g1]
    rec = {"dim": spec.dim, "seed": seed, "steps": len(a), "alpha": a, "beta": b, "converged": conv,
           "hv_seconds_mean": float(np.mean(t)), "threads": threads or os.cpu_count(),
           "ritz": {str(m): ritz(a, b, m) for m in range(1, len(a) + 1) if m % every == 0 or m == len(a)}}
    gold = json.load(open(OUT)) if os.path.exists(OUT) else {}
    gold[f"cfg{cfg}"] = rec
    json.dump(gold, open(OUT, "w"), indent=1)
    print(f"cfg{cfg}: {len(a)} steps, lowest Ritz values {rec['ritz'][str(len(a))]}", flush=True)


if __name__ == "__main__":
    main()
