"""Writes tests/golden/serve_stream.json: reference eigenvalues of a sample of
the served stream itself (generator lattice.stream_request, seed 3), so that
the service's results on the stream's own requests, up to N = 24 spins
(dim 2,704,156), are checked against an independent solver.

Selection, in stream order: the first request of every model name (chain
12..24, ladder 12..24, square J1-J2 on every torus, Hubbard chain 8..12)
among the first 2,000, then a second request of every N >= 22 name (other
couplings, other arrival).  For each, the lowest distinct eigenvalues of the
sector from the CPU oracle's H·v: dense eigvalsh up to dim 1,000, else scipy
eigsh (ARPACK, tol 1e-14), distinct where they differ by more than 1e-8
relative (a Sz = 0 sector holds one state per spin multiplet, so repeats
come only from lattice symmetry).

usage: python tests/golden/gen_serve_stream.py [threads]
"""

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

from hphi_amd import lattice as L  # noqa: E402
from oracle import oracle as O  # noqa: E402

STREAM_SEED = 3
SCAN = 2000


def pick():
    rng = np.random.default_rng(STREAM_SEED)
    reqs = [L.stream_request(rng) for _ in range(SCAN)]
    first, second = {}, {}
    for i, r in enumerate(reqs):
        if r.name not in first:
            first[r.name] = i
        elif r.name not in second and r.nsite >= 22:
            second[r.name] = i
    idx = sorted(set(first.values()) | set(second.values()))
    return [(i, reqs[i]) for i in idx]


def distinct_lowest(spec, n=6):
    import scipy.sparse.linalg as sla
    orc = O.Oracle(spec, L.processed_terms(spec)).prepare()
    if orc.dim <= 1000:
        H = np.array([orc.hv(e)[0] for e in np.eye(orc.dim, dtype=np.complex128)]).T
        w = np.linalg.eigvalsh(H)
    else:
        A = sla.LinearOperator((orc.dim, orc.dim), dtype=np.complex128,
                               matvec=lambda x: orc.hv(np.asarray(x, np.complex128).ravel())[0])
        k = min(10, orc.dim - 2)
        w = np.sort(sla.eigsh(A, k=k, which="SA", tol=1e-14, ncv=min(orc.dim - 1, 4 * k))[0])
    out = []
    for x in w:
        if not out or abs(x - out[-1]) > 1e-8 * max(1.0, abs(x)):
            out.append(float(x))
    return out[:n]


def main():
    out_path = os.path.join(HERE, "serve_stream.json")
    gold = json.load(open(out_path)) if os.path.exists(out_path) else {"stream_seed": STREAM_SEED, "picks": []}
    done = {p["index"] for p in gold["picks"]}
    for i, spec in pick():
        if i in done:
            continue
        t = time.time()
        ev = distinct_lowest(spec)
        gold["picks"].append({"index": i, "name": spec.name, "dim": int(spec.dim), "eigenvalues": ev})
        gold["picks"].sort(key=lambda p: p["index"])
        json.dump(gold, open(out_path, "w"), indent=1)  # kept as it goes
        print(f"{i} {spec.name} dim {spec.dim}: {ev[:4]} ({time.time() - t:.1f} s)", flush=True)


if __name__ == "__main__":
    main()
