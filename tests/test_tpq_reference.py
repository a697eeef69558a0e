"""The TPQ restatements (tests/tpq_restated.py) against the reference's own
numbers, and the canonical one against the mathematics.

The fixtures under tests/golden/files/{tpq,ctpq}_* were written by the
reference's CalcByTPQ (src/CalcByTPQ.c:52-294) and CalcByCanonicalTPQ
(src/CalcByCanonicalTPQ.c:49-300) compiled from its sources, with each
sample's start vector as the reference's MakeIniVec made it in that run
(ini_rand<i>.bin: global_1st_norm, then the normalised v1[1..dim]).  The
restatements, fed the same start vector and the dense textbook Hamiltonian
(tests/brute.py), must reproduce every printed number of SS_rand, Norm_rand
and Flct_rand (printed with 16 decimals; the reference's OpenMP reductions
round differently from numpy, hence a tolerance of 1e-12 relative).
"""

import json
import os

import numpy as np
import pytest

from hphi_amd import lattice as L
from tests import brute
from tests.golden.cases import ctpq_file_cases, tpq_file_cases
from tests.tpq_restated import ctpq_restated, tpq_restated

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = os.path.join(HERE, "golden", "files")
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))["files"]
CASES = {**tpq_file_cases(), **ctpq_file_cases()}


def load_ini(name, i):
    raw = np.fromfile(os.path.join(FILES, name, f"ini_rand{i}.bin"), np.float64)
    return raw[0], raw[1:].view(np.complex128)


def _table(path):
    return np.array([list(map(float, ln.split())) for ln in open(path) if not ln.startswith(" #")])


def _obs(spec, H):
    def f(v):
        p = brute.phys(spec, v, H)
        return (p["doublon"], p["doublon2"], p["num"], p["num2"], p["Sz"], p["Sz2"])
    return f


@pytest.mark.parametrize("name", sorted(CASES))
def test_tpq_restatement_reproduces_reference_files(name):
    spec, rec = CASES[name], GOLD[name]
    H = brute.dense(spec)
    assert H.shape[0] == rec["dim"]
    hv = lambda v: H @ v  # noqa: E731
    steps = int(spec.params["Lanczos_max"])
    lv = float(spec.params["LargeValue"])
    canonical = name.startswith("ctpq_")
    for i in range(int(spec.params["NumAve"])):
        norm1, v1 = load_ini(name, i)
        assert abs(np.linalg.norm(v1) - 1.0) < 1e-13
        v_init = v1 * norm1
        obs = _obs(spec, H)
        if canonical:
            rows = ctpq_restated(hv, v_init, steps, 1.0 / lv, rec["expand_coef"], obs)
            moments = rows[:, 4:10]
        else:
            base = tpq_restated(hv, v_init, steps, lv, spec.nsite)
            # the same vectors again for the moments: rerun the recurrence
            vs, v = [], v_init / np.linalg.norm(v_init)
            for s in range(steps):
                vs.append(v)
                w = lv * v - hv(v) / spec.nsite
                v = w / np.linalg.norm(w)
            rows = base
            moments = np.array([obs(u) for u in vs])
        d = os.path.join(FILES, name)
        ss = _table(os.path.join(d, f"SS_rand{i}.dat"))
        nm = _table(os.path.join(d, f"Norm_rand{i}.dat"))
        fl = _table(os.path.join(d, f"Flct_rand{i}.dat"))
        tol = dict(rtol=1e-12, atol=1e-13)
        assert len(ss) == steps and np.all(ss[:, -1] == np.arange(steps))
        np.testing.assert_allclose(ss[:, 0], rows[:, 0], **tol)  # inv_temp
        np.testing.assert_allclose(ss[:, 1], rows[:, 1], **tol)  # energy
        np.testing.assert_allclose(ss[:, 2], rows[:, 2], **tol)  # <H^2>
        np.testing.assert_allclose(ss[:, 3], moments[:, 0], **tol)  # doublon
        np.testing.assert_allclose(ss[:, 4], moments[:, 2], **tol)  # num
        np.testing.assert_allclose(nm[:, 1], rows[:, 3], **tol)  # global_norm (step 0: 1st norm)
        np.testing.assert_allclose(nm[:, 2], norm1, **tol)
        first = 0 if canonical else 1  # Flct from step 0 (canonical) or 1
        assert np.all(fl[:, -1] == np.arange(first, steps))
        # Flct columns: inv_temp, N, N^2, D, D^2, Sz, Sz^2
        want = moments[first:][:, [2, 3, 0, 1, 4, 5]]
        np.testing.assert_allclose(fl[:, 1:7], want, **tol)


def test_ctpq_restatement_pinned_by_mathematics():
    """Step s of the canonical TPQ is the normalised P(H)^s psi_0 with
    P(x) = sum_{m<=coef} (-dtau x/2)^m / m!, evaluated here in H's eigenbasis
    (no H·v recurrence); with coef large enough P(H) is exp(-dtau H/2), so
    the energy approaches the canonical ensemble's as the samples grow."""
    spec = L.heisenberg_chain(8)
    H = brute.dense(spec)
    lam, U = np.linalg.eigh(H)
    rng = np.random.default_rng(7)
    psi0 = rng.standard_normal(H.shape[0]) + 1j * rng.standard_normal(H.shape[0])
    dtau, steps = 0.25, 12
    for coef in (2, 4, 10):
        rows = ctpq_restated(lambda v: H @ v, psi0, steps, dtau, coef)
        x = -0.5 * dtau * lam
        P = sum(x ** m / np.prod(np.arange(1, m + 1)) for m in range(coef + 1))
        c0 = U.conj().T @ psi0
        for s in range(steps):
            c = P ** s * c0
            c = c / np.linalg.norm(c)
            E = np.sum(np.abs(c) ** 2 * lam)
            E2 = np.sum(np.abs(c) ** 2 * lam ** 2)
            assert abs(rows[s, 1] - E) <= 1e-12 * max(1.0, abs(E))
            assert abs(rows[s, 2] - E2) <= 1e-12 * max(1.0, E2)
            if s > 0:  # global_norm = |P(H) psi_{s-1}| for the normalised psi_{s-1}
                cp = P ** (s - 1) * c0
                cp = cp / np.linalg.norm(cp)
                assert abs(rows[s, 3] - np.linalg.norm(P * cp)) <= 1e-12
    # coef 10 is exp(-dtau H / 2) up to |x|^11 / 11! < 5e-12 per step here
    # (|x| = |dtau H / 2| < 0.46), so step s sits at beta = s dtau
    rows = ctpq_restated(lambda v: H @ v, psi0, steps, dtau, 10)
    c0 = U.conj().T @ psi0
    for s in range(steps):
        c = np.exp(-0.5 * s * dtau * lam) * c0
        c = c / np.linalg.norm(c)
        assert abs(rows[s, 1] - np.sum(np.abs(c) ** 2 * lam)) <= 1e-10
