"""Two-body terms: Hubbard InterAll, PairHop and Exchange; spin and SpinGC
InterAll.

The C restatement in oracle/ does not cover these terms; their parity rests
on fixtures written by the reference itself (tests/golden/gen_golden.py
--two-body, from oracle/_ref) and on a dense Hamiltonian built from the
operator definitions (tests/brute.py).  The CPU test pins the fixtures to the
definitions; the GPU tests put the device engine against the fixtures."""

import hashlib
import json
import os

import numpy as np
import pytest

from hphi_amd import lattice as L
from tests import brute
from tests.golden.cases import SEED, two_body_cases

GOLD = os.path.join(os.path.dirname(__file__), "golden")
HV_TOL = 1e-12


CASES = sorted(two_body_cases())


def _rec(name):
    g = json.load(open(os.path.join(GOLD, "golden.json")))
    return g["cases_two_body"][name]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", CASES)
def test_reference_two_body_fixture_matches_definitions(oracle_lib, name):
    """The reference's H·v, prdct and diagonal for each two-body model equal the
    dense Hamiltonian from the textbook operator definitions."""
    spec = two_body_cases()[name]
    rec = _rec(name)
    arr = np.load(os.path.join(GOLD, rec["arrays"]))
    H = brute.dense(spec)
    v = oracle_lib.fill_vector(SEED, rec["dim"])
    hv = H @ v
    assert np.max(np.abs(hv - arr["hv"])) <= HV_TOL * np.max(np.abs(hv))
    assert np.allclose(np.real(np.diag(H)), arr["list_Diagonal"][1:], atol=1e-13, rtol=0)
    assert abs(np.vdot(v, hv) - complex(*rec["prdct"])) <= HV_TOL * np.linalg.norm(v) * np.linalg.norm(hv)
    if "list_1" in arr:  # the basis the reference enumerated is the ascending canonical one
        assert np.array_equal(arr["list_1"][1:], np.array(brute.basis(spec), dtype=np.uint64))


@pytest.fixture(scope="module")
def A():
    from hphi_amd import capi
    capi.lib()
    return capi


@pytest.mark.gpu
@pytest.mark.parametrize("source", ["spec", "def"])
@pytest.mark.parametrize("name", CASES)
def test_two_body_tables_diag_hv_vs_reference(A, oracle_lib, tmp_path, name, source):
    print(f"[gpu] two-body {name} from {source}", flush=True)
    spec = two_body_cases()[name]
    rec = _rec(name)
    arr = np.load(os.path.join(GOLD, rec["arrays"]))
    eng = A.Engine.from_spec(spec) if source == "spec" else A.Engine.from_def(L.write_expert(spec, str(tmp_path)))
    if "sha_list_1" in rec:
        l1, l21, l22 = eng.sz()
        assert sha(l1) == rec["sha_list_1"] and sha(l21) == rec["sha_list_2_1"] and sha(l22) == rec["sha_list_2_2"]
    else:
        eng.sz()
    diag = eng.diagonalcalc()
    assert sha(diag) == rec["sha_list_Diagonal"]  # bit-exact, including the InterAll diagonal
    v = oracle_lib.fill_vector(SEED, eng.dim)
    hv, pr = eng.hv(v)
    ref = arr["hv"]
    assert np.max(np.abs(hv - ref)) <= HV_TOL * np.max(np.abs(ref))
    assert abs(pr - complex(*rec["prdct"])) <= HV_TOL * np.linalg.norm(v) * np.linalg.norm(ref)
    eng.close()


@pytest.mark.gpu
@pytest.mark.parametrize("tuning", [dict(use_c2=0), dict(use_c2=1, lds21=1, tpb=256, wgs_per_cu=4),
                                    dict(use_c2=1, lds21=0, tpb=512, wgs_per_cu=2),
                                    dict(use_c2=1, lds21=0, c2_unroll=16, tpb=256, wgs_per_cu=4),
                                    dict(use_c2=1, c2_crank=1, tpb=256, wgs_per_cu=8)],
                         ids=lambda t: "-".join(f"{k}{v}" for k, v in t.items()))
@pytest.mark.parametrize("name", CASES)
def test_two_body_every_kernel(A, oracle_lib, tuning, name):
    """Single-vector kernels under each tuning, and the block kernel for k = 4
    and 16, against the reference's H·v."""
    spec = two_body_cases()[name]
    rec = _rec(name)
    print(f"[gpu] two-body kernels {name} {tuning}", flush=True)
    if tuning.get("c2_crank") and spec.model not in (L.SPIN, L.HUBBARD):
        pytest.skip("rank tables are for canonical sectors")
    eng = A.Engine.from_spec(spec)
    if spec.model not in (L.SPIN_GC, L.HUBBARD_GC):  # the canonical kernels' knobs
        eng.tune(**tuning)
    v = oracle_lib.fill_vector(SEED, eng.dim)
    hv, _ = eng.hv(v)
    ref = np.load(os.path.join(GOLD, rec["arrays"]))["hv"]
    assert np.max(np.abs(hv - ref)) <= HV_TOL * np.max(np.abs(ref))
    H = brute.dense(spec)
    for k in (4, 16):
        V1 = A.DeviceArray(eng.dim * k)
        eng.fill_random(V1, k, 30)
        V0 = A.DeviceArray(eng.dim * k).zero()
        eng.mltply_block_dev(k, V0, V1)
        out = V0.download().reshape(eng.dim, k)
        for c in range(k):
            want = H @ oracle_lib.fill_vector(30 + c, eng.dim)
            assert np.max(np.abs(out[:, c] - want)) <= HV_TOL * np.max(np.abs(want)), (k, c)
        V0.free()
        V1.free()
    eng.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_two_body_lanczos_and_service(A, name):
    print(f"[gpu] two-body Lanczos {name}", flush=True)
    spec = two_body_cases()[name]
    w = np.linalg.eigvalsh(brute.dense(spec))
    eng = A.Engine.from_spec(spec)
    r = eng.lanczos(n_eigen=1, max_iter=400, eps=1e-14, target=0, seed=3)
    assert r.converged and abs(r.eigenvalues[0] - w[0]) <= 1e-10
    eng.close()
    if spec.model in (L.SPIN_GC, L.HUBBARD_GC):
        return  # the service runs canonical sectors
    srv = A.Server(hbm_budget=1 << 28, max_jobs=8, n_eigen=2, eps=1e-12, check_every=5)
    srv.submit(spec, 4, 0.0)
    res = []
    for _ in range(10000):
        a, wq = srv.tick()
        res += srv.results()
        if a == 0 and wq == 0:
            break
    srv.close()
    distinct = [w[0]] + [x for x in w[1:] if x - w[0] > 1e-9][:1]
    assert res and res[0].converged
    assert np.allclose(list(res[0].eigenvalues[:2]), distinct, atol=1e-10)


@pytest.mark.gpu
def test_canonical_spin_interall_must_conserve_sz(A):
    """In a canonical Spin sector an off-diagonal InterAll entry must be two
    opposite flips on distinct sites; the engine refuses the others by name."""
    s = L.heisenberg_chain(6)
    s.interall = [(0, 1, 0, 0, 2, 1, 2, 0, 0.2, 0.0), (2, 0, 2, 1, 0, 0, 0, 1, 0.2, 0.0)]  # two raising flips
    with pytest.raises(A.HphiError, match="conserve Sz"):
        A.Engine.from_spec(s)


@pytest.mark.gpu
def test_hubbard_gc_tpq_matches_restated_step(A, oracle_lib):
    """mTPQ on the HubbardGC ring: per-step rows equal the reference's step
    restated in numpy on the dense Hamiltonian, including <doublon>,
    <doublon^2> and the electron number <N> (expec_energy_flct.c:186-300)."""
    from tests.tpq_restated import tpq_restated
    print("[gpu] tpq hubbard_gc", flush=True)
    spec = two_body_cases()["hubbard_gc"]
    arr = np.load(os.path.join(GOLD, _rec("hubbard_gc")["arrays"]))
    H = brute.dense(spec)
    eng = A.Engine.from_spec(spec)
    k, steps = 2, 20
    out, res = eng.tpq(k, steps, seed=3, init_kind=0, layout=1, use_graph=True)
    lv = 3.0 * np.max(np.abs(arr["list_Diagonal"][1:])) / spec.nsite
    assert abs(res.large_value - lv) <= 1e-15 * lv
    xs = np.arange(eng.dim)
    doub = np.array([bin(int(x) & (int(x) >> 1) & 0x5555555555555555).count("1") for x in xs], float)
    num = np.array([bin(int(x)).count("1") for x in xs], float)
    for c in range(k):
        ref = tpq_restated(lambda v: H @ v, oracle_lib.fill_vector(3 + c, eng.dim), steps, lv, spec.nsite, doub, num)
        got = out[:, c, :7]
        assert np.allclose(got, ref, rtol=1e-10, atol=1e-12), (got[:3], ref[:3])
    eng.close()


@pytest.mark.gpu
def test_hubbard_gc_canonical_tpq_matches_restated_step(A, oracle_lib):
    """Canonical TPQ on the HubbardGC ring: per-step rows (inv_temp, E, <H^2>,
    norm, doublon and electron-number moments, Sz moments) equal the reference's
    step (CalcByCanonicalTPQ, MultiplyForCanonicalTPQ) restated in numpy on the
    dense Hamiltonian, Taylor order 6."""
    from tests.tpq_restated import ctpq_restated
    print("[gpu] ctpq hubbard_gc", flush=True)
    spec = two_body_cases()["hubbard_gc"]
    H = brute.dense(spec)
    eng = A.Engine.from_spec(spec)
    k, steps, lv = 2, 12, 2.5
    out, res = eng.ctpq(k, steps, seed=3, init_kind=0, expand_coef=6, large_value=lv)
    assert res.large_value == lv
    E = 0x5555555555555555
    xs = [int(x) for x in range(eng.dim)]
    D = np.array([bin(x & (x >> 1) & E).count("1") for x in xs], float)
    up = np.array([bin(x & E).count("1") for x in xs], float)
    dn = np.array([bin(x & (E << 1)).count("1") for x in xs], float)
    N, S = up + dn, up - dn

    def obs(v):
        w = np.abs(v) ** 2
        return (w @ D, w @ D ** 2, w @ N, w @ N ** 2, 0.5 * (w @ S), 0.25 * (w @ S ** 2))
    for c in range(k):
        ref = ctpq_restated(lambda v: H @ v, oracle_lib.fill_vector(3 + c, eng.dim), steps, 1.0 / lv, 6, obs)
        got = out[:, c, :10]
        assert np.allclose(got, ref, rtol=1e-10, atol=1e-12), (got[:3], ref[:3])
    eng.close()
