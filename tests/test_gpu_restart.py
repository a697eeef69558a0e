"""GPU tests of the solvers' start vectors, results and restart files:

* TPQ and canonical TPQ on the device from the reference's own start vectors
  (tests/golden/files/*/ini_rand<i>.bin, written by the reference's MakeIniVec
  in the run that wrote the fixtures), written with the project's file
  writers and compared number by number with the SS_rand, Norm_rand and
  Flct_rand files the reference's CalcByTPQ / CalcByCanonicalTPQ wrote;
* a run cut at step n and resumed from its restart file (tmpvec_set<i>: step
  field n, the normalised vector of step n - 1, CalcByTPQ.c:86-117,272-287)
  equals the straight run, rows and files;
* LOBPCG eigenvectors (CalcByLOBPCG.c:716-724): orthonormal, eigenvectors of H
  to the solver's residual, their energies (expec_energy_flct) the returned
  eigenvalues; written as restart files and read back they start a converged
  run (Initialize_wave, :172-213); the block solvers' energy file.
"""

import json
import os

import numpy as np
import pytest

from hphi_amd import lattice as L
from tests.golden.cases import ctpq_file_cases, tpq_file_cases

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = os.path.join(HERE, "golden", "files")
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))["files"]
CASES = {**tpq_file_cases(), **ctpq_file_cases()}


@pytest.fixture(scope="module")
def A():
    from hphi_amd import capi
    capi.lib()
    return capi


def _table(path):
    return np.array([list(map(float, ln.split())) for ln in open(path) if not ln.startswith(" #")])


def _start_block(A, name, dim, n):
    """the reference's start vectors w_0 = global_1st_norm * v1 of samples 0..n-1"""
    cols = []
    for i in range(n):
        raw = np.fromfile(os.path.join(FILES, name, f"ini_rand{i}.bin"), np.float64)
        norm1, v1 = raw[0], raw[1:].view(np.complex128)
        assert v1.size == dim
        cols.append(v1 * norm1)
    return A.DeviceArray(dim * n).upload(np.concatenate(cols))


def _run(A, eng, name, spec, W0, steps, step0=0, W_end=None):
    if name.startswith("ctpq_"):
        return eng.ctpq_from(W0, int(spec.params["NumAve"]), steps, step0, W_end,
                             expand_coef=GOLD[name]["expand_coef"], large_value=float(spec.params["LargeValue"]))
    return eng.tpq_from(W0, int(spec.params["NumAve"]), steps, step0, W_end,
                        large_value=float(spec.params["LargeValue"]))


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_tpq_from_reference_start_vectors_reproduces_reference_files(A, name, tmp_path):
    spec = CASES[name]
    print(f"[gpu] tpq from reference start vectors {name}", flush=True)
    eng = A.Engine.from_spec(spec)
    n, steps = int(spec.params["NumAve"]), int(spec.params["Lanczos_max"])
    W0 = _start_block(A, name, eng.dim, n)
    out, _ = _run(A, eng, name, spec, W0, steps)
    write = A.write_ctpq_files if name.startswith("ctpq_") else A.write_tpq_files
    for i in range(n):
        write(str(tmp_path), i, out[:, i])
        for kind in ("SS", "Norm", "Flct"):
            got = _table(tmp_path / f"{kind}_rand{i}.dat")
            ref = _table(os.path.join(FILES, name, f"{kind}_rand{i}.dat"))
            assert got.shape == ref.shape, (kind, i)
            assert np.array_equal(got[:, -1], ref[:, -1]), (kind, i)
            assert np.allclose(got, ref, rtol=1e-10, atol=1e-12), (kind, i, np.max(np.abs(got - ref)))
    W0.free()
    eng.close()


@pytest.mark.parametrize("name", ["tpq_chain8", "tpq_hubbard_gc", "ctpq_hubbard_chain4", "ctpq_tri_3x2_gc"])
def test_resumed_tpq_equals_straight_run(A, name, tmp_path):
    spec = CASES[name]
    print(f"[gpu] tpq resume {name}", flush=True)
    eng = A.Engine.from_spec(spec)
    n, steps, cut = int(spec.params["NumAve"]), int(spec.params["Lanczos_max"]), 5
    dim = eng.dim
    W0 = _start_block(A, name, dim, n)
    full, _ = _run(A, eng, name, spec, W0, steps)
    W_end = A.DeviceArray(dim * n)
    first, _ = _run(A, eng, name, spec, W0, cut, W_end=W_end)
    np.testing.assert_array_equal(first, full[:cut])
    # the restart files as the reference writes them: step field = steps done
    ve = W_end.download().reshape(n, dim)
    for i in range(n):
        assert abs(np.linalg.norm(ve[i]) - 1.0) < 1e-13
        A.write_vec_file(str(tmp_path / f"tmpvec_set{i}_rank_0.dat"), cut, ve[i])
    back, st = [], []
    for i in range(n):
        v, s = A.read_vec_file(str(tmp_path / f"tmpvec_set{i}_rank_0.dat"), dim)
        back.append(v)
        st.append(s)
    assert st == [cut] * n
    W1 = A.DeviceArray(dim * n).upload(np.concatenate(back))
    rest, _ = _run(A, eng, name, spec, W1, steps, step0=cut)
    assert rest.shape[0] == steps - cut
    # the resumed run re-normalises the stored vector: equal to rounding
    np.testing.assert_allclose(rest, full[cut:], rtol=1e-11, atol=1e-12)
    # appended to the first run's files they read as the straight run's
    write = A.write_ctpq_files if name.startswith("ctpq_") else A.write_tpq_files
    for i in range(n):
        d_full, d_cut = tmp_path / "full", tmp_path / "cut"
        d_full.mkdir(exist_ok=True)
        d_cut.mkdir(exist_ok=True)
        write(str(d_full), i, full[:, i])
        write(str(d_cut), i, first[:, i])
        A.append_tpq_files(str(d_cut), i, cut, rest[:, i], norm1=float(full[0, i, 3]))
        for kind in ("SS", "Norm", "Flct"):
            a, b = _table(d_full / f"{kind}_rand{i}.dat"), _table(d_cut / f"{kind}_rand{i}.dat")
            assert a.shape == b.shape and np.array_equal(a[:, -1], b[:, -1]), (kind, i)
            np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-12)
    for arr in (W0, W_end, W1):
        arr.free()
    eng.close()


@pytest.mark.parametrize("spec,k", [(L.heisenberg_chain(12, 1.0, 0.3), 4), (L.hubbard_chain(6, 1.0, 4.0), 2),
                                    (L.triangular_xxz_gc(3, 3, 1.0, 0.8, 0.1), 8)],
                         ids=["chain12", "hubbard6", "tri3x3gc"])
def test_lobpcg_eigenvectors_and_restart_files(A, spec, k, tmp_path):
    print(f"[gpu] lobpcg eigenvectors {spec.name} k={k}", flush=True)
    eng = A.Engine.from_spec(spec)
    dim = eng.dim
    X = A.DeviceArray(dim * k)
    r, hist = eng.lobpcg_ex(k, Xout=X, max_iter=400, eps=1e-7, seed=3)
    assert r.converged
    ev = np.array(r.eigenvalues[:k])
    xs = X.download().reshape(k, dim)
    # orthonormal columns; eigenvectors of H to the residual threshold
    G = xs.conj() @ xs.T
    assert np.max(np.abs(G - np.eye(k))) < 1e-8
    for c in range(k):
        hv, pr = eng.hv(xs[c])
        res = np.linalg.norm(hv - ev[c] * xs[c])
        assert res <= 2.0 * max(1e-7, 1e-7 * abs(ev[c])), (c, res)
        assert abs(pr.real - ev[c]) < 1e-10 * max(1.0, abs(ev[c]))
    # phys of each eigenvector (phys.c): energy = the eigenvalue
    phys = []
    for c in range(k):
        v = A.DeviceArray(dim).upload(xs[c])
        p = eng.expec_energy_flct(v)
        v.free()
        assert abs(p["energy"] - ev[c]) < 1e-10 * max(1.0, abs(ev[c]))
        phys.append(p)
    A.write_energy_states(str(tmp_path / "zvo_energy.dat"), phys)
    txt = open(tmp_path / "zvo_energy.dat").read().split("\n")
    assert txt[0] == "State 0" and txt[1].startswith("  Energy  ") and txt[4] == ""
    assert abs(float(txt[1].split()[1]) - ev[0]) < 1e-12 * max(1.0, abs(ev[0]))
    # restart files tmpvec_set<c>_rank_0.dat (Output_restart): step field, dim, v[0..dim]
    for c in range(k):
        A.write_vec_file(str(tmp_path / f"tmpvec_set{c}_rank_0.dat"), r.iterations, xs[c])
    cols = []
    for c in range(k):
        v, st = A.read_vec_file(str(tmp_path / f"tmpvec_set{c}_rank_0.dat"), dim)
        assert st == r.iterations
        cols.append(v)
    X0 = A.DeviceArray(dim * k).upload(np.concatenate(cols))
    r2, hist2 = eng.lobpcg_ex(k, X0=X0, max_iter=400, eps=1e-7, seed=99)
    assert r2.converged and r2.iterations <= 2, r2.iterations
    assert np.max(np.abs(np.array(r2.eigenvalues[:k]) - ev)) < 1e-9 * max(1.0, np.max(np.abs(ev)))
    for arr in (X, X0):
        arr.free()
    eng.close()
