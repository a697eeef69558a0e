"""Writes tests/golden/serve_eigen.json: the lowest distinct eigenvalues of the
service's test models from the CPU oracle (dense eigvalsh for small sectors,
scipy eigsh otherwise), so GPU tests need no CPU eigen-solve.

usage: python tests/golden/gen_serve_eigen.py
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

from hphi_amd import lattice as L  # noqa: E402
from oracle import oracle as O  # noqa: E402


def serve_models():
    return {
        "chain12_j2_0.3": L.heisenberg_chain(12, 1.0, 0.3),
        "chain13_j2_0.6_sz1": L.heisenberg_chain(13, 1.0, 0.6, two_sz=1),
        "ladder12_j2_0.4": L.j1j2_ladder(6, 1.0, 0.4),
        "square4x3_j2_0.55": L.j1j2_square(4, 3, 1.0, 0.55),
        "hubbard_chain8_u3.3": L.hubbard_chain(8, 1.0, 3.3),
        "hubbard_chain7_u5.1": L.hubbard_chain(7, 1.0, 5.1),
        "chain14": L.heisenberg_chain(14, 1.0, 0.0),
        "ladder10_j2_0.9": L.j1j2_ladder(5, 1.0, 0.9),
    }


def distinct_lowest(spec, n=6):
    import scipy.sparse.linalg as sla
    orc = O.Oracle(spec, L.processed_terms(spec)).prepare()
    if orc.dim <= 1000:
        H = np.array([orc.hv(e)[0] for e in np.eye(orc.dim, dtype=np.complex128)]).T
        w = np.linalg.eigvalsh(H)
    else:
        A = sla.LinearOperator((orc.dim, orc.dim), dtype=np.complex128,
                               matvec=lambda x: orc.hv(np.asarray(x, np.complex128).ravel())[0])
        w = np.sort(sla.eigsh(A, k=min(20, orc.dim - 2), which="SA", tol=1e-14)[0])
    out = []
    for x in w:
        if not out or abs(x - out[-1]) > 1e-8 * max(1.0, abs(x)):
            out.append(float(x))
    return out[:n]


if __name__ == "__main__":
    res = {name: distinct_lowest(spec) for name, spec in serve_models().items()}
    json.dump(res, open(os.path.join(HERE, "serve_eigen.json"), "w"), indent=1)
    print(res)
