"""C ABI on the CPU: the library loads, exports every declared symbol, and its
host-only parts (expert-mode reader, tridiagonal eigensolver) are correct.
No device work here."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from hphi_amd import lattice as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def capi():
    from hphi_amd import capi as A
    if not os.path.exists(A.LIB_PATH):
        from hphi_amd import build
        build.build()
    return A


def _declared():
    txt = open(os.path.join(ROOT, "include", "hphi_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hphi_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(capi):
    lib = capi.lib()
    names = _declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(capi.EXPORTED) <= set(names)


def _read(capi, path):
    lib = capi.lib()
    lib.hphi_read_def.argtypes = [C.c_char_p, C.POINTER(C.POINTER(capi.ModelDesc)), C.POINTER(capi.ModPara)]
    lib.hphi_model_desc_free.argtypes = [C.POINTER(capi.ModelDesc)]
    pd = C.POINTER(capi.ModelDesc)()
    mp = capi.ModPara()
    rc = lib.hphi_read_def(path.encode(), C.byref(pd), C.byref(mp))
    return rc, pd, mp


def _arr(ptr, n, w=1):
    return np.array([ptr[i] for i in range(n * w)]).reshape(n, w) if n else np.zeros((0, w))


def _two_body():
    from tests.golden.cases import two_body_cases
    return list(two_body_cases().values())


@pytest.mark.parametrize("spec", [
    *_two_body(),
    L.baseline_config(1), L.hubbard_chain(5), L.hubbard_square(4, 2, ncond=8),
    L.triangular_xxz_gc(4, 2), L.j1j2_square(4, 3, 1.0, 0.31, two_sz=2),
], ids=lambda s: s.name)
def test_def_reader_matches_python_mirror(capi, tmp_path, spec):
    """The C reader reproduces readdef's term processing (merge, Hermitian
    pairing, Ising split), cross-checked against lattice.processed_terms."""
    nl = L.write_expert(spec, str(tmp_path))
    rc, pd, mp = _read(capi, nl)
    assert rc == 0, capi.lib().hphi_last_error()
    d = pd.contents
    t = L.processed_terms(spec)
    assert (d.model, d.nsite, d.nup, d.ndown) == (spec.model, spec.nsite, spec.nup, spec.ndown)
    assert np.array_equal(_arr(d.exchange_ij, d.n_exchange, 2), t["exchange_ij"].reshape(-1, 2))
    assert np.array_equal(_arr(d.exchange_J, d.n_exchange)[:, 0], t["exchange_J"])
    assert np.array_equal(_arr(d.transfer_idx, d.n_transfer, 4), t["trans_idx"].reshape(-1, 4))
    assert np.array_equal(_arr(d.transfer_t, d.n_transfer, 2), t["trans_t"].reshape(-1, 2))
    assert np.array_equal(_arr(d.chemi_mu, d.n_chemi)[:, 0] if d.n_chemi else np.zeros(0), t["chemi_mu"])
    assert np.array_equal(_arr(d.intra_U, d.n_intra)[:, 0] if d.n_intra else np.zeros(0), t["intra_U"])
    assert np.array_equal(_arr(d.inter_ij, d.n_inter, 2), t["inter_ij"].reshape(-1, 2))
    assert np.array_equal(_arr(d.inter_V, d.n_inter)[:, 0] if d.n_inter else np.zeros(0), t["inter_V"])
    assert np.array_equal(_arr(d.hund_ij, d.n_hund, 2), t["hund_ij"].reshape(-1, 2))
    assert np.array_equal(_arr(d.hund_J, d.n_hund)[:, 0] if d.n_hund else np.zeros(0), t["hund_J"])
    assert np.array_equal(_arr(d.interall_idx, d.n_interall, 8), t["interall_idx"].reshape(-1, 8))
    assert np.array_equal(_arr(d.interall_V, d.n_interall, 2), t["interall_V"].reshape(-1, 2))
    assert np.array_equal(_arr(d.interall_diag_idx, d.n_interall_diag, 4), t["interall_diag_idx"].reshape(-1, 4))
    assert np.array_equal(_arr(d.interall_diag_V, d.n_interall_diag)[:, 0] if d.n_interall_diag else np.zeros(0),
                          t["interall_diag_V"])
    assert np.array_equal(_arr(d.pairhop_ij, d.n_pairhop, 2), t["pairhop_ij"].reshape(-1, 2))
    assert np.array_equal(_arr(d.pairhop_J, d.n_pairhop)[:, 0] if d.n_pairhop else np.zeros(0), t["pairhop_J"])
    assert np.array_equal(_arr(d.chemi_site, d.n_chemi)[:, 0] if d.n_chemi else np.zeros(0), t["chemi_site"])
    assert mp.exct == int(spec.params["exct"]) and mp.lanczos_max == int(spec.params["Lanczos_max"])
    capi.lib().hphi_model_desc_free(pd)


def test_def_reader_merges_and_pairs_transfers(capi, tmp_path):
    spec = L.hubbard_chain(4)
    spec.trans = spec.trans + [(0, 0, 1, 0, 0.25, 0.0), (1, 0, 0, 0, 0.25, 0.0), (2, 1, 2, 1, 0.3, 0.0)]
    nl = L.write_expert(spec, str(tmp_path))
    rc, pd, _ = _read(capi, nl)
    assert rc == 0
    d = pd.contents
    t = L.processed_terms(spec)
    assert d.n_transfer == len(t["trans_t"]) == 16
    assert np.array_equal(_arr(d.transfer_t, d.n_transfer, 2), t["trans_t"])
    assert d.n_chemi == 1 and d.chemi_site[0] == 2 and d.chemi_spin[0] == 1 and d.chemi_mu[0] == 0.3
    capi.lib().hphi_model_desc_free(pd)


@pytest.mark.parametrize("mutate,msg", [
    (lambda s: s.trans.append((0, 0, 1, 0, 1.0, 0.0)), "Hermitian"),
    (lambda s: setattr(s, "two_sz", None), "2Sz"),
])
def test_def_reader_errors(capi, tmp_path, mutate, msg):
    spec = L.hubbard_chain(4) if msg == "Hermitian" else L.heisenberg_chain(6)
    mutate(spec)
    if msg == "2Sz":
        nl = L.write_expert(L.heisenberg_chain(6), str(tmp_path))
        txt = open(os.path.join(str(tmp_path), "modpara.def")).read().replace("2Sz 0\n", "")
        open(os.path.join(str(tmp_path), "modpara.def"), "w").write(txt)
    else:
        nl = L.write_expert(spec, str(tmp_path))
    rc, _, _ = _read(capi, nl)
    assert rc == -1
    assert msg in capi.lib().hphi_last_error().decode()


def test_def_reader_rejects_offsite_spin_transfer(capi, tmp_path):
    """A canonical Spin Trans entry must act on one site (readdef.c:1106-1110)."""
    spec = L.heisenberg_chain(6)
    spec.trans = [(0, 0, 1, 0, 0.1, 0.0), (1, 0, 0, 0, 0.1, 0.0)]
    nl = L.write_expert(spec, str(tmp_path))
    rc, _, _ = _read(capi, nl)
    assert rc == -1 and "one site" in capi.lib().hphi_last_error().decode()


def _calcmod(tmp_path, text):
    nl = L.write_expert(L.heisenberg_chain(6), str(tmp_path))
    p = os.path.join(str(tmp_path), "calcmod.def")
    open(p, "w").write(text)
    return nl


def test_def_reader_passes_driver_flags(capi, tmp_path):
    nl = _calcmod(tmp_path, "CalcType 1\nCalcModel 1\nReStart 2\nInitialVecType -1\nCalcEigenVec 1\n"
                            "OutputEigenVec 1\nInputEigenVec 0\nOutputMode 1\nNGPU 2\n")
    rc, pd, mp = _read(capi, nl)
    assert rc == 0, capi.lib().hphi_last_error()
    assert (mp.calc_type, mp.restart, mp.initial_vec_type, mp.calc_eigenvec, mp.output_eigenvec,
            mp.input_eigenvec) == (1, 2, -1, 1, 1, 0)
    capi.lib().hphi_model_desc_free(pd)


@pytest.mark.parametrize("line,msg", [
    ("Bogus 1", "unknown keyword Bogus"), ("InputHam 1", "InputHam 1 is not supported"),
    ("CalcSpec 1", "CalcSpec 1 is not supported"), ("ReStart 4", "ReStart must be"),
])
def test_def_reader_calcmod_errors(capi, tmp_path, line, msg):
    nl = _calcmod(tmp_path, f"CalcType 0\nCalcModel 1\n{line}\n")
    rc, _, _ = _read(capi, nl)
    assert rc == -1 and msg in capi.lib().hphi_last_error().decode()


def test_def_reader_rejects_unsupported_keyword(capi, tmp_path):
    nl = L.write_expert(L.heisenberg_chain(6), str(tmp_path))
    with open(nl, "a") as f:
        f.write("Boost boost.def\n")
    rc, _, _ = _read(capi, nl)
    assert rc == -1 and "Boost" in capi.lib().hphi_last_error().decode()


@pytest.mark.parametrize("what", ["spin_format", "pairhop_spin", "no_partner"])
def test_def_reader_rejects_bad_two_body_input(capi, tmp_path, what):
    """Spin InterAll entries must read (i s1 i s2 k s3 k s4)
    (CheckFormatForSpinInt, readdef.c:2637-2650); PairHop needs a Hubbard
    model; an off-diagonal entry needs its Hermitian partner."""
    s = L.heisenberg_chain(6)
    if what == "spin_format":
        s.interall = [(0, 0, 1, 0, 2, 1, 2, 1, 0.5, 0.0)]
        msg = "i == j"
    elif what == "pairhop_spin":
        s.pairhop = [(0, 1, 0.5)]
        msg = "Hubbard"
    else:
        s.interall = [(0, 0, 0, 1, 2, 1, 2, 0, 0.5, 0.1)]
        msg = "Hermitian"
    nl = L.write_expert(s, str(tmp_path))
    rc, _, _ = _read(capi, nl)
    assert rc == -1 and msg in capi.lib().hphi_last_error().decode()
    with pytest.raises(ValueError):
        L.processed_terms(s)


def test_tridiagonal_eigensolver(capi):
    rng = np.random.default_rng(5)
    for n in (1, 2, 3, 10, 200):
        a = rng.normal(size=n)
        b = np.abs(rng.normal(size=n - 1)) + 1e-3
        T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
        k = min(n, 4)
        ev = capi.tridiag_lowest(a, b, k)
        assert np.allclose(ev, np.linalg.eigvalsh(T)[:k], atol=1e-13, rtol=0)
    # degenerate pair
    a = np.array([1.0, 1.0, 3.0])
    b = np.array([0.0, 0.0])
    assert np.allclose(capi.tridiag_lowest(a, b, 3), [1.0, 1.0, 3.0])


def test_good_eigenvalues_survive_many_ghost_copies(capi, oracle_lib):
    """Plain Lanczos without reorthogonalisation keeps adding ghost copies of
    converged eigenvalues.  After hundreds of steps the lowest eigenvalues of T_m
    are mostly copies, and the Cullum-Willoughby filter must look past them:
    the lowest 4 distinct good values equal the dense spectrum's at every
    check, from step 60 to step 600 of a 12-site chain (dim 924)."""
    from tests import brute
    spec = L.heisenberg_chain(12, 1.0, 0.3)
    w = np.linalg.eigvalsh(brute.dense(spec))
    distinct = [w[0]]
    for x in w[1:]:
        if x - distinct[-1] > 1e-8:
            distinct.append(x)
    orc = oracle_lib.Oracle(spec, L.processed_terms(spec)).prepare()
    n = orc.dim
    v = oracle_lib.fill_vector(5, n)
    v = v / np.linalg.norm(v)
    vprev = np.zeros(n, complex)
    alpha, beta, b = [], [], 0.0
    for _ in range(600):
        x = orc.hv(v)[0] - b * vprev
        a = np.vdot(v, x).real
        r = x - a * v
        b = np.linalg.norm(r)
        alpha.append(a)
        beta.append(b)
        vprev, v = v, r / b
    A = np.array(alpha)
    B = np.array(beta)
    dp = C.POINTER(C.c_double)
    fn = capi.lib().hphi_lanczos_good_eigenvalues
    fn.argtypes = [dp, dp, C.c_int, C.c_int, dp]
    for m in range(60, 601, 10):
        out = np.zeros(16)
        k = fn(A.ctypes.data_as(dp), B.ctypes.data_as(dp), m, 4, out.ctypes.data_as(dp))
        assert k == 4, (m, k, out[:k])
        assert np.allclose(out[:4], distinct[:4], atol=1e-8, rtol=0), (m, out[:4], distinct[:4])


def test_ritz_residual_equals_the_ritz_vectors_residual(capi):
    """hphi_lanczos_ritz_residual (beta_m |y_m| by inverse iteration) against
    |H z - theta z| of the explicit Ritz vector z = V_m y of a plain Lanczos
    run on a dense random Hermitian matrix, before and after convergence; and
    hphi_lanczos_residuals_small follows it."""
    lib = capi.lib()
    lib.hphi_lanczos_ritz_residual.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_double,
                                               C.c_double, C.POINTER(C.c_double)]
    lib.hphi_lanczos_residuals_small.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_double,
                                                 C.POINTER(C.c_double), C.c_int, C.c_double]
    rng = np.random.default_rng(7)
    n = 300
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = (X + X.conj().T) / 2
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    v /= np.linalg.norm(v)
    V, a, b = [v], [], []
    w = H @ v
    for m in range(1, 61):
        al = np.vdot(V[-1], w).real
        w = w - al * V[-1] - (b[-1] * V[-2] if b else 0)
        a.append(al)
        b.append(np.linalg.norm(w))
        V.append(w / b[-1])
        w = H @ V[-1]
        if m in (10, 30, 60):
            T = np.diag(a) + np.diag(b[:-1], 1) + np.diag(b[:-1], -1)
            th, Y = np.linalg.eigh(T)
            Vm = np.array(V[:m]).T
            aa, bb = np.array(a), np.array(b[:-1] if m > 1 else [0.0])
            for i in (0, 1, m // 2):
                z = Vm @ Y[:, i]
                true = np.linalg.norm(H @ z - th[i] * z)
                got = C.c_double()
                assert lib.hphi_lanczos_ritz_residual(aa.ctypes.data_as(C.POINTER(C.c_double)),
                                                      bb.ctypes.data_as(C.POINTER(C.c_double)), m, b[-1], th[i],
                                                      C.byref(got)) == 0
                assert abs(got.value - true) <= 1e-8 * max(true, 1.0) + 1e-10, (m, i, got.value, true)
                assert abs(got.value - b[-1] * abs(Y[-1, i])) <= 1e-9 * max(got.value, 1e-3)
            cur = th[:2].copy()
            small = lib.hphi_lanczos_residuals_small(aa.ctypes.data_as(C.POINTER(C.c_double)),
                                                     bb.ctypes.data_as(C.POINTER(C.c_double)), m, b[-1],
                                                     cur.ctypes.data_as(C.POINTER(C.c_double)), 2, 1e-12)
            r = [b[-1] * abs(Y[-1, i]) for i in (0, 1)]
            assert small == int(all(x <= 1e-7 * max(1.0, abs(t)) for x, t in zip(r, cur))), (m, r)
