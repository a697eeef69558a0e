"""GPU parity: the HIP path through the C ABI against the oracle and the
reference's golden fixtures.

Bars: basis tables and list_Diagonal bit-exact; H·v within 1e-12 relative
infinity-norm of the oracle (the gather form sums in another order than the
reference's scatter, so it is not bit-identical); prdct within 1e-12
relative; eigenvalues within 1e-10 of an independent solver."""

import hashlib
import json
import os

import numpy as np
import pytest

from hphi_amd import lattice as L
from tests.golden.cases import SEED, golden_cases

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
HV_TOL = 1e-12


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def A():
    from hphi_amd import capi
    capi.lib()
    return capi


@pytest.fixture(scope="module")
def gold():
    return json.load(open(os.path.join(GOLD, "golden.json")))


def rel_inf(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def prdct_ok(pr, opr, v, hv):
    """prdct = <v|Hv> is a dot product: its rounding error scales with
    |v| |Hv| (Cauchy-Schwarz), not with the possibly cancelled result."""
    return abs(pr - opr) <= HV_TOL * np.linalg.norm(v) * np.linalg.norm(hv)


@pytest.mark.parametrize("name", sorted(golden_cases()))
def test_tables_diag_hv_vs_reference_and_oracle(A, oracle_lib, gold, name):
    print(f"[gpu] case {name}", flush=True)
    spec = golden_cases()[name]
    rec = gold["cases"][name]
    eng = A.Engine.from_spec(spec)
    assert eng.dim == rec["dim"]
    l1, l21, l22 = eng.sz()
    if spec.model in (L.SPIN, L.HUBBARD):
        assert sha(l1) == rec["sha_list_1"]
        assert sha(l21) == rec["sha_list_2_1"]
        assert sha(l22) == rec["sha_list_2_2"]
    diag = eng.diagonalcalc()
    assert sha(diag) == rec["sha_list_Diagonal"]
    v = oracle_lib.fill_vector(SEED, eng.dim)
    hv, pr = eng.hv(v)
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    ohv, opr = orc.hv(v)
    assert rel_inf(hv, ohv) <= HV_TOL
    assert prdct_ok(pr, opr, v, ohv)
    assert prdct_ok(pr, complex(*rec["prdct"]), v, ohv)
    if "arrays" in rec:
        ref = np.load(os.path.join(GOLD, rec["arrays"]))["hv"]
        assert rel_inf(hv, ref) <= HV_TOL
    eng.close()


def test_device_generator_matches_host(A, oracle_lib):
    eng = A.Engine.from_spec(L.heisenberg_chain(10))
    d = A.DeviceArray(eng.dim)
    eng.fill_random(d, 1, 3)
    assert np.array_equal(d.download(), oracle_lib.fill_vector(3, eng.dim))
    k = 4
    b = A.DeviceArray(eng.dim * k)
    eng.fill_random(b, k, 11)
    B = b.download().reshape(eng.dim, k)
    for c in range(k):
        assert np.array_equal(B[:, c], oracle_lib.fill_vector(11 + c, eng.dim))


@pytest.mark.parametrize("name", ["chain13_sz1", "hubbard_4x2", "tri_4x3_gc"])
def test_mltply_reference_shaped_host_arrays(A, oracle_lib, name):
    """hphi_mltply on the reference's 1-indexed host arrays: accumulates into v0,
    leaves element 0 alone, returns prdct."""
    spec = golden_cases()[name]
    eng = A.Engine.from_spec(spec)
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    dim = eng.dim
    v1 = np.zeros(dim + 1, np.complex128)
    v1[1:] = oracle_lib.fill_vector(5, dim)
    v0 = np.zeros(dim + 1, np.complex128)
    v0[1:] = oracle_lib.fill_vector(6, dim)
    v0[0] = 123.0 + 4.0j
    w0, w1, v0_in = v0.copy(), v1.copy(), v0.copy()
    pr = eng.mltply(v0, v1)
    opr = orc.mltply(w0, w1)
    assert v0[0] == 123.0 + 4.0j
    assert rel_inf(v0[1:], w0[1:]) <= HV_TOL
    assert prdct_ok(pr, opr, v1[1:], w0[1:] - v0_in[1:])
    eng.close()


@pytest.mark.parametrize("k", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("name", ["j1j2_4x4", "hubbard_4x2", "tri_4x3_gc"])
def test_block_hv_equals_columns(A, oracle_lib, name, k):
    print(f"[gpu] block {name} k={k}", flush=True)
    spec = golden_cases()[name]
    eng = A.Engine.from_spec(spec)
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    dim = eng.dim
    V1 = A.DeviceArray(dim * k)
    eng.fill_random(V1, k, 40)
    V0 = A.DeviceArray(dim * k).zero()
    pr = eng.mltply_block_dev(k, V0, V1)
    out = V0.download().reshape(dim, k)
    for c in range(k):
        vc = oracle_lib.fill_vector(40 + c, dim)
        ohv, opr = orc.hv(vc)
        assert rel_inf(out[:, c], ohv) <= HV_TOL
        assert prdct_ok(pr[c], opr, vc, ohv)
    eng.close()


BLOCK_SHAPES = [(2, 1), (2, 2), (4, 1), (4, 4), (8, 1), (8, 4), (8, 8), (16, 1), (16, 4), (16, 16)]


@pytest.mark.parametrize("tpb", [256, 1024])
@pytest.mark.parametrize("k,cpl", BLOCK_SHAPES + [(4, 0), (16, 0)], ids=lambda v: str(v))
@pytest.mark.parametrize("name", ["j1j2_4x4", "hubbard_3x2_n5", "hubbard_chain6_complex", "tri_4x3_gc",
                                  "chain10_field"])
def test_block_shapes_match_oracle(A, oracle_lib, name, k, cpl, tpb):
    """Every block-kernel shape (k columns, cpl columns per lane, workgroup
    size; cpl 0 = the per-column kernel) against the oracle, column by column."""
    print(f"[gpu] block shape {name} k={k} cpl={cpl} tpb={tpb}", flush=True)
    spec = golden_cases()[name]
    eng = A.Engine.from_spec(spec)
    if cpl == 0:
        eng.tune(use_blk=0)
    else:
        eng.tune(**{f"blk_cpl{k}": cpl, "blk_tpb": tpb})
        assert eng.get_tune(f"blk_cpl{k}") == cpl
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    dim = eng.dim
    V1 = A.DeviceArray(dim * k)
    eng.fill_random(V1, k, 60)
    V0 = A.DeviceArray(dim * k)
    eng.fill_random(V0, k, 90)  # v0 += H v1: the old content must be kept
    v0_in = V0.download().reshape(dim, k)
    pr = eng.mltply_block_dev(k, V0, V1)
    out = V0.download().reshape(dim, k) - v0_in
    for c in range(k):
        vc = oracle_lib.fill_vector(60 + c, dim)
        ohv, opr = orc.hv(vc)
        assert rel_inf(out[:, c], ohv) <= HV_TOL, c
        assert prdct_ok(pr[c], opr, vc, ohv), c
    V0.free()
    V1.free()
    eng.close()


def test_from_def_equals_from_terms(A, tmp_path, oracle_lib):
    spec = golden_cases()["hubbard_chain6_complex"]
    nl = L.write_expert(spec, str(tmp_path))
    e1 = A.Engine.from_def(nl)
    e2 = A.Engine.from_spec(spec)
    v = oracle_lib.fill_vector(9, e1.dim)
    h1, p1 = e1.hv(v)
    h2, p2 = e2.hv(v)
    assert np.array_equal(h1, h2) and p1 == p2
    assert np.array_equal(e1.diagonalcalc(), e2.diagonalcalc())


def test_hv_is_deterministic(A, oracle_lib):
    spec = golden_cases()["j1j2_4x4"]
    eng = A.Engine.from_spec(spec)
    v = oracle_lib.fill_vector(1, eng.dim)
    a, pa = eng.hv(v)
    b, pb = eng.hv(v)
    assert np.array_equal(a, b) and pa == pb


LANCZOS_KNOWN = [
    # doc/tutorial/en/source/zero_temperature/spin_chain.rst:30,39 (6 decimals)
    ("chain16", L.heisenberg_chain(16), [-7.142296], 5e-7),
    # test/lobcg_hubbard_square.sh reference.dat
    ("hubbard_4x2", L.hubbard_square(4, 2, ncond=8), [-10.2529529552637033], 1e-10),
]


@pytest.mark.parametrize("name,spec,ev,tol", LANCZOS_KNOWN, ids=[k[0] for k in LANCZOS_KNOWN])
def test_lanczos_known_answers(A, oracle_lib, name, spec, ev, tol):
    import scipy.sparse.linalg as sla
    print(f"[gpu] lanczos {name}", flush=True)
    eng = A.Engine.from_spec(spec)
    r = eng.lanczos(n_eigen=1, max_iter=2000, eps=1e-14, target=2, seed=0)
    assert r.converged
    assert abs(r.eigenvalues[0] - ev[0]) <= tol
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    Aop = sla.LinearOperator((orc.dim, orc.dim), dtype=np.complex128,
                             matvec=lambda x: orc.hv(np.asarray(x, np.complex128).ravel())[0])
    w = np.sort(sla.eigsh(Aop, k=3, which="SA", tol=1e-14)[0])
    assert abs(r.eigenvalues[0] - w[0]) <= 1e-10
    eng.close()


TUNINGS = [dict(use_c2=0), dict(use_c2=1, lds21=0, tpb=256, wgs_per_cu=4),
           dict(use_c2=1, lds21=1, tpb=1024, wgs_per_cu=1, xcd_map=1),
           dict(use_c2=1, lds21=1, tpb=512, wgs_per_cu=1, xcd_map=0),
           dict(use_c2=1, lds21=1, c2_unroll=16, tpb=1024, wgs_per_cu=1),
           dict(use_c2=1, lds21=0, c2_unroll=16, tpb=256, wgs_per_cu=4),
           dict(use_c2=1, c2_crank=1, tpb=256, wgs_per_cu=8),           # spin only: rank tables
           dict(use_c2=1, c2_crank=1, c2_unroll=16, tpb=512, wgs_per_cu=2, xcd_map=1),
           # canonical Spin: the high-half terms in their own pass over a tile-major copy of v1
           dict(use_c2=1, c2_crank=1, hi_split=1), dict(use_c2=1, c2_crank=1, hi_split=1, hi_fuse_copy=0),
           dict(use_c2=1, lds21=1, tpb=1024, wgs_per_cu=1, hi_split=1, hi_wgs_per_cu=2)]


@pytest.mark.parametrize("tuning", TUNINGS, ids=lambda t: "-".join(f"{k}{v}" for k, v in t.items()))
@pytest.mark.parametrize("name", ["chain16", "j1j2_4x3_sz2", "hubbard_4x2", "hubbard_3x2_n5",
                                  "hubbard_chain6_complex", "chain10_field"])
def test_every_tuning_matches_oracle(A, oracle_lib, name, tuning):
    print(f"[gpu] tuning {name} {tuning}", flush=True)
    spec = golden_cases()[name]
    eng = A.Engine.from_spec(spec)
    try:
        eng.tune(**tuning)
    except A.HphiError:
        pytest.skip("tuning not applicable to this model")
    v = oracle_lib.fill_vector(SEED, eng.dim)
    hv, pr = eng.hv(v)
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    ohv, opr = orc.hv(v)
    assert rel_inf(hv, ohv) <= HV_TOL
    assert prdct_ok(pr, opr, v, ohv)
    eng.close()


@pytest.mark.parametrize("spec", [L.j1j2_square(4, 4, 1.0, 0.5), L.j1j2_ladder(9, 1.0, 0.3),
                                  L.heisenberg_chain(18, 1.0, 0.4, two_sz=2), L.heisenberg_chain(17, 1.0, 0.2, two_sz=1),
                                  L.hubbard_square(4, 2, 1.0, 4.0), L.hubbard_chain(10, 1.0, 2.0, ncond=9),
                                  L.hubbard_square(3, 3, 1.0, 8.0, two_sz=1)],
                         ids=lambda s: s.name)
def test_high_half_pass_equals_single_pass(A, spec):
    """The two-pass H·v (high-half exchange terms or hops over the tile-major
    copy of v1) against the single pass on the same vector: 1e-14 relative
    (only the order of the sums differs), prdct likewise, in accumulate and in
    overwrite form."""
    print(f"[gpu] high-half pass {spec.name}", flush=True)
    eng = A.Engine.from_spec(spec)
    rng = np.random.default_rng(11)
    v = rng.standard_normal(eng.dim) + 1j * rng.standard_normal(eng.dim)
    w = rng.standard_normal(eng.dim) + 1j * rng.standard_normal(eng.dim)
    out = {}
    for hs in (0, 1):
        eng.tune(hi_split=hs)
        assert eng.get_tune("hi_split") == hs
        a, b = A.DeviceArray(eng.dim).upload(v), A.DeviceArray(eng.dim).upload(w)
        p1 = eng.mltply_dev(b, a)
        p2 = eng.mltply_dev(b, a)  # w + 2 H v
        acc = b.download()
        eng.mltply_dev_scaled(b, a, 0.0, 2.0, overwrite=True)  # 2 H v
        out[hs] = (acc, b.download(), p1, p2)
        a.free(); b.free()
    (y0, z0, p0, q0), (y1, z1, p1, q1) = out[0], out[1]
    assert rel_inf(y1, y0) <= 1e-14 and rel_inf(z1, z0) <= 1e-14
    assert rel_inf(y1 - w, z1) <= 1e-13  # accumulate and overwrite forms agree
    assert abs(p1 - p0) <= 1e-13 * abs(p0) and abs(q1 - q0) <= 1e-13 * abs(q0)
    eng.close()


def test_high_half_pass_is_the_default_at_full_size(A):
    """The 8x4 J1-J2 sector takes the two-pass H·v by default (its full-size
    checksums against the reference run through it); a small sector and the
    4x4 Hubbard sector (slower in two passes) do not."""
    big = A.Engine.from_spec(L.baseline_config(3))
    small = A.Engine.from_spec(L.j1j2_square(4, 4, 1.0, 0.5))
    hub = A.Engine.from_spec(L.baseline_config(2))
    assert big.get_tune("hi_split") == 1 and small.get_tune("hi_split") == 0 and hub.get_tune("hi_split") == 0
    for e in (big, small, hub):
        e.close()


def _extreme_eigenvalues(orc):
    """lowest and highest eigenvalue of the oracle's H (ARPACK, both ends)"""
    from scipy.sparse.linalg import LinearOperator, eigsh
    op = LinearOperator((orc.dim, orc.dim), matvec=lambda v: orc.hv(np.asarray(v, np.complex128).ravel())[0],
                        dtype=np.complex128)
    lo = eigsh(op, k=1, which="SA", tol=1e-10, return_eigenvectors=False)[0]
    hi = eigsh(op, k=1, which="LA", tol=1e-10, return_eigenvectors=False)[0]
    return lo, hi


def check_default_large_value(spec, orc, lv):
    """The default l (no LargeValue given) is at least 3 max|diag| / Nsite, as the
    cfg4 workload defines it, and l Nsite bounds |E| so that l - H/Nsite is
    positive definite (the mTPQ recurrence converges to the lowest state)."""
    m = np.max(np.abs(orc.diag[1:]))
    assert lv >= 3.0 * m / spec.nsite * (1 - 1e-15)
    lo, hi = _extreme_eigenvalues(orc)
    assert lv * spec.nsite > max(abs(lo), abs(hi))


@pytest.mark.parametrize("make", [
    lambda: L.spin_model("xy_chain10", 10, [(i, j, 1.0) for i, j in L.chain_bonds(10)], two_sz=0, jz_scale=0.0),
    lambda: L.hubbard_chain(6, U=0.0),
    lambda: L.spin_model("tfield_gc8", 8, [(i, j, 1.0) for i, j in L.chain_bonds(8)], two_sz=None, jz_scale=0.0,
                         hx=0.7),
], ids=["xy_chain10", "hubbard_chain6_U0", "xy_field_gc8"])
def test_tpq_default_large_value_bounds_zero_diagonal_models(A, oracle_lib, make):
    """Hamiltonians without a diagonal part: 3 max|diag| = 0 does not bound the
    spectrum, the Gershgorin rule does; mTPQ then lowers the energy."""
    spec = make()
    print(f"[gpu] tpq default l {spec.name}", flush=True)
    eng = A.Engine.from_spec(spec)
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    assert np.max(np.abs(orc.diag[1:])) == 0.0
    out, res = eng.tpq(1, 40, seed=5, init_kind=0)
    check_default_large_value(spec, orc, res.large_value)
    assert out[-1, 0, 1] < out[0, 0, 1]
    eng.close()


@pytest.mark.parametrize("layout,graph", [(1, True), (0, True), (1, False)])
@pytest.mark.parametrize("name,k", [("tri_4x3_gc", 4), ("hubbard_4x2", 2), ("j1j2_4x4", 2)])
def test_tpq_matches_restated_reference_step(A, oracle_lib, name, k, layout, graph):
    """Per-step TPQ output (inv_temp, E, <H^2>, norm, Sz moments) equals the
    reference's step restated in numpy on the oracle's H·v, same initial vectors."""
    from tests.tpq_restated import tpq_restated
    print(f"[gpu] tpq {name} k={k} layout={layout} graph={graph}", flush=True)
    spec = golden_cases()[name]
    eng = A.Engine.from_spec(spec)
    steps = 25
    out, res = eng.tpq(k, steps, seed=3, init_kind=0, layout=layout, use_graph=graph)
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    lv = res.large_value
    check_default_large_value(spec, orc, lv)
    sz = None  # per-state observable: SpinGC 2 Sz(x), Hubbard doublon count
    if spec.model == L.SPIN_GC:
        sz = np.array([2 * bin(x).count("1") - spec.nsite for x in range(orc.dim)], float)
    elif spec.model == L.HUBBARD:
        sz = np.array([bin(int(x) & (int(x) >> 1) & 0x5555555555555555).count("1")
                       for x in orc.list_1[1:]], float)
    hv = lambda v: orc.hv(v)[0]
    for c in range(k):
        ref = tpq_restated(hv, oracle_lib.fill_vector(3 + c, orc.dim), steps, lv, spec.nsite, sz)
        got = out[:, c]
        assert np.allclose(got[:, :4], ref[:, :4], rtol=1e-10, atol=1e-12), (got[:3], ref[:3])
        if spec.model == L.SPIN_GC:  # Sz = <2Sz>/2, Sz^2 = <(2Sz)^2>/4
            assert np.allclose(got[:, 8], 0.5 * ref[:, 4], rtol=1e-10, atol=1e-12)
            assert np.allclose(got[:, 9], 0.25 * ref[:, 5], rtol=1e-10, atol=1e-12)
            assert np.all(got[:, 6] == spec.nsite) and np.all(got[:, 4] == 0)
        elif spec.model == L.HUBBARD:  # doublon moments; fixed electron number
            assert np.allclose(got[:, 4:6], ref[:, 4:6], rtol=1e-10, atol=1e-12)
            assert np.allclose(got[:, 6], spec.ncond, rtol=1e-12)
            assert np.allclose(got[:, 8], 0.5 * spec.two_sz, rtol=1e-12, atol=1e-12)
        else:  # canonical Spin: constants of the sector
            assert np.all(got[:, 6] == spec.nsite) and np.all(got[:, 8] == 0.5 * spec.two_sz)
        assert np.all(np.abs(got[:, 10]) <= 1e-10 * np.abs(got[:, 1]) + 1e-12)  # <H> real
    eng.close()


def _state_moments(spec, orc):
    """Per-state doublon count D, electron number N and 2 Sz of the oracle's basis."""
    E = 0x5555555555555555
    xs = [int(x) for x in orc.list_1[1:]] if spec.model in (L.SPIN, L.HUBBARD) else list(range(orc.dim))
    if spec.model in (L.SPIN, L.SPIN_GC):
        D = np.zeros(len(xs))
        N = np.full(len(xs), float(spec.nsite))
        S = np.array([2 * bin(x).count("1") - spec.nsite for x in xs], float)
    else:
        D = np.array([bin(x & (x >> 1) & E).count("1") for x in xs], float)
        up = np.array([bin(x & E).count("1") for x in xs], float)
        dn = np.array([bin(x & (E << 1)).count("1") for x in xs], float)
        N, S = up + dn, up - dn
    return D, N, S


@pytest.mark.parametrize("name,k,coef", [("tri_4x3_gc", 4, 0), ("hubbard_4x2", 2, 0), ("j1j2_4x4", 1, 4),
                                         ("chain16", 16, 10), ("hubbard_chain6_complex", 2, 6)])
def test_ctpq_matches_restated_reference_step(A, oracle_lib, name, k, coef):
    """Canonical TPQ rows (inv_temp, E, <H^2>, norm, moments) equal the
    reference's step (CalcByCanonicalTPQ, MultiplyForCanonicalTPQ) restated in
    numpy on the oracle's H·v, from the same start vectors."""
    from tests.tpq_restated import ctpq_restated
    print(f"[gpu] ctpq {name} k={k} coef={coef}", flush=True)
    spec = golden_cases()[name]
    eng = A.Engine.from_spec(spec)
    steps = 15
    out, res = eng.ctpq(k, steps, seed=3, init_kind=0, expand_coef=coef)
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    lv = res.large_value
    check_default_large_value(spec, orc, lv)
    D, N, S = _state_moments(spec, orc)

    def obs(v):
        w = np.abs(v) ** 2
        return (w @ D, w @ D ** 2, w @ N, w @ N ** 2, 0.5 * (w @ S), 0.25 * (w @ S ** 2))
    hv = lambda v: orc.hv(v)[0]  # noqa: E731
    for c in range(k):
        ref = ctpq_restated(hv, oracle_lib.fill_vector(3 + c, orc.dim), steps, 1.0 / lv, coef or 10, obs)
        got = out[:, c]
        assert np.allclose(got[:, :4], ref[:, :4], rtol=1e-10, atol=1e-12), (got[:3, :4], ref[:3, :4])
        assert np.allclose(got[:, 4:10], ref[:, 4:10], rtol=1e-10, atol=1e-12), (got[:3, 4:10], ref[:3, 4:10])
        assert np.all(np.abs(got[:, 10]) <= 1e-10 * np.abs(got[:, 1]) + 1e-12)
    # the energy falls as beta grows (imaginary-time evolution from beta = 0)
    assert np.all(out[-1, :, 1] < out[0, :, 1])
    eng.close()


def test_tpq_gaussian_start_is_normalised(A):
    spec = golden_cases()["tri_4x3_gc"]
    eng = A.Engine.from_spec(spec)
    out, _ = eng.tpq(4, 3, seed=0, init_kind=-1, layout=1)
    # |w_0|^2 of a complex Gaussian vector ~ 2 dim
    assert np.all(np.abs(out[0, :, 3] ** 2 / (2 * eng.dim) - 1.0) < 0.1)
    eng.close()


@pytest.mark.parametrize("name", ["chain16", "hubbard_4x2", "tri_4x3_gc", "hubbard_chain6_complex"])
def test_scaled_hv_matches_oracle(A, oracle_lib, name):
    """hphi_mltply_dev_scaled: v0 = c0 v0 + c1 H v1 (accumulate) and c1 H v1
    (overwrite), with prdct = <v1|H|v1>, against the oracle's H·v."""
    print(f"[gpu] scaled hv {name}", flush=True)
    spec = golden_cases()[name]
    eng = A.Engine.from_spec(spec)
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    v1h = oracle_lib.fill_vector(SEED, eng.dim)
    v0h = oracle_lib.fill_vector(SEED + 1, eng.dim)
    ohv, opr = orc.hv(v1h)
    c0, c1 = -0.7, 1.3
    v1 = A.DeviceArray(eng.dim).upload(v1h)
    for overwrite in (False, True):
        v0 = A.DeviceArray(eng.dim).upload(v0h)
        pr = eng.mltply_dev_scaled(v0, v1, c0, c1, overwrite)
        want = c1 * ohv + (0.0 if overwrite else c0 * v0h)
        assert rel_inf(v0.download(), want) <= HV_TOL
        assert prdct_ok(pr, opr, v1h, ohv)
        v0.free()
    eng.close()


def test_lanczos_residual_pass(A):
    """hphi_vec_lanczos_resid: r <- r - f w, |r|^2, and vg += cg w."""
    spec = golden_cases()["chain16"]
    eng = A.Engine.from_spec(spec)
    rng = np.random.default_rng(3)
    n = eng.dim
    rh, wh, gh = (rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(3))
    r, w, g = (A.DeviceArray(n).upload(x) for x in (rh, wh, gh))
    f, cg = 0.37, -1.9
    r2 = eng.lanczos_resid(r, w, g, f, cg)
    want = rh - f * wh
    assert rel_inf(r.download(), want) <= 1e-15
    assert rel_inf(g.download(), gh + cg * wh) <= 1e-15
    assert abs(r2 - np.vdot(want, want).real) <= 1e-13 * r2
    r3 = eng.lanczos_resid(w, r, None, 0.0)  # without the eigenvector sum, f = 0: |w|^2
    assert abs(r3 - np.vdot(wh, wh).real) <= 1e-13 * r3
    for x in (r, w, g):
        x.free()
    eng.close()


@pytest.mark.parametrize("which", ["torus_4x2_gc", "chain10_doubled"])
def test_repeated_terms_merge_to_the_same_hamiltonian(A, oracle_lib, which):
    """Identical compiled terms are merged into one with the summed coupling;
    the oracle applies every listed entry separately.  The 4x2 triangular torus
    lists its a2 bonds twice (24 entries, 20 pairs), as the 13x2 one does; the
    chain lists every exchange and Ising bond twice."""
    print(f"[gpu] merged terms {which}", flush=True)
    if which == "torus_4x2_gc":
        spec = L.triangular_xxz_gc(4, 2, 1.0, 0.8, 0.1)
    else:
        spec = L.heisenberg_chain(10)
        spec.exchange = spec.exchange + spec.exchange
        spec.ising = spec.ising + spec.ising
    eng = A.Engine.from_spec(spec)
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    v = oracle_lib.fill_vector(SEED, eng.dim)
    hv, pr = eng.hv(v)
    ohv, opr = orc.hv(v)
    assert rel_inf(hv, ohv) <= HV_TOL
    assert prdct_ok(pr, opr, v, ohv)
    eng.close()


@pytest.mark.parametrize("which", ["hubbard17_n2", "chain40_one_magnon"])
def test_states_wider_than_32_bits(A, oracle_lib, which):
    """Sectors of 34- and 40-bit states run the 64-bit kernels: H·v against the
    oracle, and for the 40-site chain with one flipped spin the closed-form
    one-magnon spectrum N J/4 - J (1 - cos 2 pi m / N) of the dense matrix the
    engine's H·v builds column by column."""
    print(f"[gpu] wide states {which}", flush=True)
    if which == "hubbard17_n2":
        spec = L.hubbard_chain(17, 1.0, 3.0, ncond=2, two_sz=0)
    else:
        spec = L.heisenberg_chain(40, 1.0, 0.0, two_sz=38)
    assert spec.nbits > 32
    eng = A.Engine.from_spec(spec)
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    v = oracle_lib.fill_vector(SEED, eng.dim)
    hv, pr = eng.hv(v)
    ohv, opr = orc.hv(v)
    assert rel_inf(hv, ohv) <= HV_TOL
    assert prdct_ok(pr, opr, v, ohv)
    if which == "chain40_one_magnon":
        n = eng.dim
        H = np.zeros((n, n), complex)
        for c in range(n):
            e = np.zeros(n, complex)
            e[c] = 1.0
            H[:, c] = eng.hv(e)[0]
        assert np.max(np.abs(H - H.conj().T)) <= 1e-14
        want = np.sort([40 / 4 - (1 - np.cos(2 * np.pi * m / 40)) for m in range(40)])
        assert np.max(np.abs(np.linalg.eigvalsh(H) - want)) <= 1e-12
    eng.close()


def test_one_state_and_diagonal_only_sectors(A, oracle_lib):
    """Degenerate inputs: the fully polarised 8-site chain has one state (H·v =
    diag·v, and Lanczos returns N J/4 = 2 at once); an Ising-only chain has no
    off-diagonal term at all (every term group empty)."""
    spec = L.heisenberg_chain(8, 1.0, 0.0, two_sz=8)
    eng = A.Engine.from_spec(spec)
    assert eng.dim == 1
    hv, pr = eng.hv(np.array([0.6 - 0.8j]))
    assert abs(hv[0] - 2.0 * (0.6 - 0.8j)) <= 1e-15 and abs(pr - 2.0) <= 1e-15
    r = eng.lanczos(n_eigen=1, max_iter=50, eps=1e-12, seed=1)
    assert r.iterations == 1 and abs(r.eigenvalues[0] - 2.0) <= 1e-15
    eng.close()
    ising = L.heisenberg_chain(10, 1.0, 0.0, two_sz=0)
    ising.exchange = []
    eng = A.Engine.from_spec(ising)
    orc = oracle_lib.Oracle(ising, L.processed_terms(ising)).prepare()
    v = oracle_lib.fill_vector(SEED, eng.dim)
    hv, pr = eng.hv(v)
    ohv, opr = orc.hv(v)
    assert rel_inf(hv, ohv) <= HV_TOL and prdct_ok(pr, opr, v, ohv)
    assert np.allclose(hv, orc.diag[1:] * v, rtol=1e-15, atol=1e-15)
    eng.close()


@pytest.mark.parametrize("name,k", [("chain16", 1), ("hubbard_4x2", 4), ("tri_4x3_gc", 8)])
def test_autotuner_applies_a_correct_configuration(A, oracle_lib, name, k):
    """hphi_autotune times every candidate shape (kernel, workgroup size,
    workgroups per CU, columns per lane, XCD map) and leaves the fastest
    applied: the report lists the timed candidates, the winner is no slower
    than the default shape, and the tuned engine's H·v still equals the oracle's."""
    print(f"[gpu] autotune {name} k={k}", flush=True)
    spec = golden_cases()[name]
    eng = A.Engine.from_spec(spec)
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    dim = eng.dim
    V1 = A.DeviceArray(dim * k)
    eng.fill_random(V1, k, 40)
    V0 = A.DeviceArray(dim * k).zero()
    t_default = eng.time_mltply(k, V0, V1, 1, 3)
    report = eng.autotune(k, V0, V1, reps=2)
    times = [float(ln.rsplit(":", 1)[1].split()[0]) for ln in report.strip().split("\n")]
    assert len(times) >= 8, report
    assert min(times) <= 1.05 * t_default + 0.01, (min(times), t_default)
    V0.zero()
    pr = eng.mltply_block_dev(k, V0, V1) if k > 1 else [eng.mltply_dev(V0, V1)]
    out = V0.download().reshape(dim, k)
    for c in range(k):
        vc = oracle_lib.fill_vector(40 + c, dim)
        ohv, opr = orc.hv(vc)
        assert rel_inf(out[:, c], ohv) <= HV_TOL
        assert prdct_ok(pr[c], opr, vc, ohv)
    eng.close()


def test_host_pointer_mltply_equals_device_mltply_across_chunks(A):
    """hphi_mltply on the reference's 1-indexed host arrays (pinned bounce
    buffers, up to 16 workers with 16 MiB chunks, v0 added on the host) equals
    the device-resident call bit for bit: J1-J2 7x4 Sz=0 (dim 40,116,600, 642 MB
    per vector, several chunks per worker), v0 non-zero so the accumulation
    counts, and index 0 untouched."""
    spec = L.j1j2_square(7, 4, 1.0, 0.35)
    print(f"[gpu] host-pointer mltply {spec.name} dim {spec.dim}", flush=True)
    eng = A.Engine.from_spec(spec)
    dim = eng.dim
    d1 = A.DeviceArray(dim)
    d0 = A.DeviceArray(dim)
    eng.fill_random(d1, 1, 5)
    eng.fill_random(d0, 1, 6)
    h1 = np.zeros(dim + 1, np.complex128)
    h0 = np.zeros(dim + 1, np.complex128)
    h1[1:] = d1.download()
    h0[1:] = d0.download()
    h0[0] = 7.0 + 3.0j  # the reference's arrays are 1-indexed: element 0 is not part of the vector
    pd = eng.mltply_dev(d0, d1)
    ph = eng.mltply(h0, h1)
    ref = d0.download()
    assert h0[0] == 7.0 + 3.0j
    assert np.array_equal(h0[1:], ref)
    assert ph == pd
    d0.free()
    d1.free()
    eng.close()
