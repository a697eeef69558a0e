"""Eigenvector and expectation values: expec_energy_flct, the one- and
two-body Green's functions (expec_cisajs, expec_cisajscktaltdc) and the Lanczos
eigenvector pass.

Parity rests on fixtures written by the reference itself (oracle/_ref in
`green` mode: tests/golden/gen_golden.py --green) for a seeded normalised
vector, and on the operator definitions (tests/brute.py).  The reference prints
Green's functions with 10 decimals, so fixtures are compared to 2e-10; the
definitions pin the full precision.  The CPU tests pin the fixtures to the
definitions and the reader rules; the GPU tests put the device code against both."""

import json
import os

import numpy as np
import pytest

from hphi_amd import lattice as L
from tests import brute
from tests.golden.cases import SEED, green_cases

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PRINT_TOL = 2e-10  # %.10lf in the reference's output files
CASES = sorted(green_cases())


def _rec(name):
    return json.load(open(os.path.join(GOLD, "golden.json")))["cases_green"][name]


def _vec(oracle_lib, dim):
    v = oracle_lib.fill_vector(SEED, dim)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("name", CASES)
def test_reference_green_fixture_matches_definitions(oracle_lib, name):
    spec = green_cases()[name]
    rec = _rec(name)
    v = _vec(oracle_lib, rec["dim"])
    g1, g2 = L.processed_green(spec)
    assert [tuple(r[:4]) for r in rec["green1"]] == g1   # the reader's drops
    assert [tuple(r[:8]) for r in rec["green2"]] == g2
    for r in rec["green1"]:
        want = brute.green1(spec, v, r[:4])
        assert abs(complex(r[4], r[5]) - want) <= PRINT_TOL, (r, want)
    for r in rec["green2"]:
        want = brute.green2(spec, v, r[:8])
        assert abs(complex(r[8], r[9]) - want) <= PRINT_TOL, (r, want)
    ph = brute.phys(spec, v)
    for k, want in ph.items():
        assert abs(rec["phys"][k] - want) <= 1e-12 * max(1.0, abs(want)), (k, rec["phys"][k], want)


def test_processed_green_rejects_bad_spin_two_body():
    s = green_cases()["green_spin_interall"]
    s.green2 = [(0, 0, 1, 0, 2, 0, 2, 0)]
    with pytest.raises(ValueError):
        L.processed_green(s)


@pytest.fixture(scope="module")
def A():
    from hphi_amd import capi
    capi.lib()
    return capi


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_green_and_flct_vs_reference(A, oracle_lib, name):
    print(f"[gpu] green {name}", flush=True)
    spec = green_cases()[name]
    rec = _rec(name)
    eng = A.Engine.from_spec(spec)
    v = _vec(oracle_lib, eng.dim)
    dv = A.DeviceArray(eng.dim).upload(v)
    g1, g2 = L.processed_green(spec)
    got1 = eng.cisajs(dv, g1)
    got2 = eng.cisajscktalt(dv, g2)
    for r, g in zip(rec["green1"], got1):
        assert abs(complex(r[4], r[5]) - g) <= PRINT_TOL, (r, g)
        assert abs(brute.green1(spec, v, r[:4]) - g) <= 1e-13, (r, g)
    for r, g in zip(rec["green2"], got2):
        assert abs(complex(r[8], r[9]) - g) <= PRINT_TOL, (r, g)
        assert abs(brute.green2(spec, v, r[:8]) - g) <= 1e-13, (r, g)
    ph = eng.expec_energy_flct(dv)
    names = {"sz": "Sz", "sz2": "Sz2"}
    for k, val in ph.items():
        want = rec["phys"][names.get(k, k)]
        assert abs(val - want) <= 1e-12 * max(1.0, abs(want)), (k, val, want)
    dv.free()
    eng.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES + ["chain13_sz1", "hubbard_chain5", "tri_4x2_gc"])
def test_lanczos_eigenvector_is_the_ground_state(A, name):
    """The eigenvector pass returns the dense ground state: E = w0 to 1e-10,
    <H^2> - E^2 <= 1e-8 |E|^2, and |<v|psi0>| = 1 to 1e-8 when w0 is simple."""
    from tests.golden.cases import golden_cases
    print(f"[gpu] eigenvector {name}", flush=True)
    spec = green_cases()[name] if name in green_cases() else golden_cases()[name]
    H = brute.dense(spec)
    w, U = np.linalg.eigh(H)
    eng = A.Engine.from_spec(spec)
    dv, r = eng.lanczos_eigenvector(k=0, n_eigen=1, max_iter=500, eps=1e-14, target=0, seed=3)
    v = dv.download()
    assert abs(np.linalg.norm(v) - 1.0) <= 1e-12
    ph = eng.expec_energy_flct(dv)
    E = ph["energy"]
    assert abs(E - w[0]) <= 1e-10 * max(1.0, abs(w[0])), (E, w[0])
    assert ph["var"] - E * E <= 1e-8 * max(1.0, E * E)
    if w[1] - w[0] > 1e-8:
        assert abs(abs(np.vdot(U[:, 0], v)) - 1.0) <= 1e-8
    dv.free()
    eng.close()


NBODY = json.load(open(os.path.join(GOLD, "golden.json")))["cases_green"]


def _nbody_cases():
    from tests.golden.cases import nbody_green_cases
    return nbody_green_cases()


@pytest.mark.parametrize("name", sorted(_nbody_cases()))
def test_reference_nbody_green_fixture_matches_definitions(oracle_lib, name):
    """The reference's SpinGC three/four/six-body lines (expec_Threebody/
    Fourbody/Sixbody_SpinGCHalf, written by the compiled reference) against the
    operator definitions; entries it does not evaluate print eight indices."""
    spec = _nbody_cases()[name]
    rec = NBODY[name]
    v = _vec(oracle_lib, rec["dim"])
    for nb in (3, 4, 6):
        ents = getattr(spec, f"green{nb}")
        lines = rec[f"green{nb}"]
        assert len(lines) == len(ents)
        for e, r in zip(ents, lines):
            want = brute.green_nbody(spec, v, e)
            if want is None:
                assert r[:-2] == list(e[:8]) and r[-2:] == [0.0, 0.0]
                continue
            assert r[:-2] == list(e)
            assert abs(complex(r[-2], r[-1]) - want) <= PRINT_TOL, (nb, e, r[-2:], want)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(_nbody_cases()))
def test_device_nbody_green_vs_reference(name, tmp_path, oracle_lib):
    """hphi_expec_spin_nbody against the reference's lines (2e-10, its 10
    printed decimals) and the definitions (1e-13); hphi_write_nbody_green
    re-emits the reference's parsed values byte for byte."""
    from hphi_amd import capi
    spec = _nbody_cases()[name]
    rec = NBODY[name]
    print(f"[gpu] n-body Green's functions {name}", flush=True)
    eng = capi.Engine.from_spec(spec)
    v = _vec(oracle_lib, rec["dim"])
    dv = capi.DeviceArray(eng.dim).upload(v)
    for nb, fname in ((3, "zvo_ThreeBody.dat"), (4, "zvo_FourBody.dat"), (6, "zvo_SixBody.dat")):
        ents = getattr(spec, f"green{nb}")
        got, ok = eng.expec_spin_nbody(dv, nb, ents)
        for e, r, g, k in zip(ents, rec[f"green{nb}"], got, ok):
            want = brute.green_nbody(spec, v, e)
            assert k == (want is not None), e
            if want is None:
                continue
            assert abs(g - want) <= 1e-13, (nb, e, g, want)
            assert abs(g - complex(r[-2], r[-1])) <= PRINT_TOL, (nb, e, g, r[-2:])
        ref_vals = [complex(r[-2], r[-1]) for r in rec[f"green{nb}"]]
        out = tmp_path / fname
        capi.write_nbody_green(str(out), nb, ents, ref_vals, ok)
        assert open(out, "rb").read() == open(os.path.join(GOLD, "files", name, fname), "rb").read()
    dv.free()
    eng.close()
