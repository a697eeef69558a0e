"""The reference's output and restart files (host-side C writers and readers).

The Green's-function files and the TPQ files (SS_rand, Norm_rand, Flct_rand)
in tests/golden/files/ were written by the reference itself (oracle/_ref:
expec_cisajs, expec_cisajscktaltdc, CalcByTPQ; tests/golden/gen_golden.py
--files).  The writers re-emit the values parsed from those files and must
reproduce them byte for byte: headers, column order, spacing, line ends, the
first-step norm columns and the step from which Flct lines start.  Binary
vector files and TMComponents are checked by layout and round trip."""

import json
import os
import struct

import numpy as np
import pytest

from hphi_amd import capi
from hphi_amd import lattice as L

FILES = os.path.join(os.path.dirname(__file__), "golden", "files")
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))["files"]


def _read(p):
    return open(p, "rb").read()


@pytest.mark.parametrize("name", sorted(k for k in GOLD if k.startswith("green_")))
def test_green_files_byte_identical(tmp_path, name):
    rec = GOLD[name]
    for fname, width in (("zvo_cisajs.dat", 4), ("zvo_cisajscktalt.dat", 8)):
        if fname not in rec["files"]:
            continue
        src = os.path.join(FILES, name, fname)
        rows = [ln.split() for ln in open(src) if ln.strip()]
        idx = [[int(x) for x in r[:width]] for r in rows]
        vals = [complex(float(r[width]), float(r[width + 1])) for r in rows]
        out = str(tmp_path / fname)
        if width == 4:
            capi.write_cisajs(out, idx, vals)
        else:
            capi.write_cisajscktalt(out, rec["model"], idx, vals)
        assert _read(out) == _read(src), fname


def _tpq_rows(d, i):
    ss = [list(map(float, ln.split())) for ln in open(os.path.join(d, f"SS_rand{i}.dat")) if not ln.startswith(" #")]
    nm = [list(map(float, ln.split())) for ln in open(os.path.join(d, f"Norm_rand{i}.dat")) if not ln.startswith(" #")]
    fl = {int(r[-1]): r for r in ([list(map(float, ln.split())) for ln in open(os.path.join(d, f"Flct_rand{i}.dat"))
                                   if not ln.startswith(" #")])}
    rows = np.zeros((len(ss), 12))
    for s, (a, b) in enumerate(zip(ss, nm)):
        rows[s, 0:3] = a[0:3]
        rows[s, 3] = b[1]
        rows[s, 4] = a[3]
        rows[s, 6] = a[4]
        if s in fl:
            f = fl[s]
            rows[s, 7], rows[s, 5], rows[s, 8], rows[s, 9] = f[2], f[4], f[5], f[6]
    return rows


@pytest.mark.parametrize("name", sorted(k for k in GOLD if k.startswith(("tpq_", "ctpq_"))))
def test_tpq_files_byte_identical(tmp_path, name):
    """SS_rand / Norm_rand / Flct_rand as the reference's CalcByTPQ (tpq_*) and
    CalcByCanonicalTPQ (ctpq_*, with a Flct line at step 0) wrote them."""
    d = os.path.join(FILES, name)
    nsamp = sum(1 for f in GOLD[name]["files"] if f.startswith("SS_rand"))
    assert nsamp >= 1
    write = capi.write_ctpq_files if name.startswith("ctpq_") else capi.write_tpq_files
    for i in range(nsamp):
        write(str(tmp_path), i, _tpq_rows(d, i))
        for k in ("SS", "Norm", "Flct"):
            f = f"{k}_rand{i}.dat"
            assert _read(str(tmp_path / f)) == _read(os.path.join(d, f)), f


def test_vec_file_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    v = rng.standard_normal(37) + 1j * rng.standard_normal(37)
    p = str(tmp_path / "zvo_eigenvec_0_rank_0.dat")
    capi.write_vec_file(p, 42, v)
    raw = _read(p)
    assert len(raw) == 4 + 8 + 16 * 38  # int step, unsigned long dim, dim+1 complex128
    step, dim = struct.unpack("<iQ", raw[:12])
    assert (step, dim) == (42, 37)
    arr = np.frombuffer(raw[12:], np.complex128)
    assert arr[0] == 0 and np.array_equal(arr[1:], v)
    back, st = capi.read_vec_file(p, 37)
    assert st == 42 and np.array_equal(back, v)
    with pytest.raises(capi.HphiError, match="incorrect"):
        capi.read_vec_file(p, 36)


def test_restart_vecs_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    v0 = rng.standard_normal(20) + 1j * rng.standard_normal(20)
    v1 = rng.standard_normal(20) + 1j * rng.standard_normal(20)
    p = str(tmp_path / "zvo_recalcvec_rank_0.dat")
    capi.write_restart_vecs(p, 17, v0, v1)
    raw = _read(p)
    assert len(raw) == 8 + 8 + 2 * 16 * 21 and struct.unpack("<QQ", raw[:16]) == (17, 20)
    a, b, st = capi.read_restart_vecs(p, 20)
    assert st == 17 and np.array_equal(a, v0) and np.array_equal(b, v1)


def test_tmcomponents_format_and_round_trip(tmp_path):
    alpha = np.array([-1.25, 0.5, 2.0 / 3.0])
    beta = np.array([0.75, 1.0 / 7.0, 0.0])
    p = str(tmp_path / "zvo_TMComponents.dat")
    capi.write_tmcomponents(p, 3.5, alpha, beta)
    text = open(p).read().split("\n")
    assert text[0] == "3 " and text[1] == "3.5000000000 " and text[2] == "-1.2500000000 0.7500000000 "
    nrm, a, b = capi.read_tmcomponents(p)
    assert nrm == 3.5 and np.allclose(a, alpha, atol=1e-10) and np.allclose(b, beta, atol=1e-10)


def test_energy_file(tmp_path):
    p = str(tmp_path / "zvo_energy.dat")
    capi.write_energy(p, {"energy": -7.125, "doublon": 0.25, "sz": -0.5})
    assert open(p).read() == "Energy  -7.1250000000000000 \nDoublon  0.2500000000000000 \nSz  -0.5000000000000000 \n"


@pytest.fixture(scope="module")
def A():
    capi.lib()
    return capi


@pytest.mark.gpu
def test_device_outputs_to_reference_files(A, tmp_path, oracle_lib):
    """Device results written in the reference's formats: TPQ rows of the
    Hubbard chain (k = 2) as SS_rand / Norm_rand / Flct_rand, the Lanczos
    eigenvector as an eigenvec file, its energy file and Green's functions;
    everything read back equals what the device produced (to print precision)."""
    from tests import brute
    print("[gpu] device outputs to files", flush=True)
    spec = L.hubbard_chain(4)
    eng = A.Engine.from_spec(spec)
    out, _ = eng.tpq(2, 12, seed=5, init_kind=0, layout=1)
    for c in range(2):
        capi.write_tpq_files(str(tmp_path), c, out[:, c])
        ss = np.loadtxt(str(tmp_path / f"SS_rand{c}.dat"))
        fl = np.loadtxt(str(tmp_path / f"Flct_rand{c}.dat"))
        assert np.allclose(ss[:, [0, 1, 2, 3, 4]], out[:, c, [0, 1, 2, 4, 6]], rtol=0, atol=1e-15)
        assert np.allclose(fl[:, 1:7], out[1:, c][:, [6, 7, 4, 5, 8, 9]], rtol=0, atol=1e-15)
        assert np.allclose(fl[:, 1], spec.ncond, atol=1e-12)  # N fixed in the canonical sector
    dv, r = eng.lanczos_eigenvector(k=0, max_iter=200, eps=1e-14, target=0, seed=1)
    v = dv.download()
    p = str(tmp_path / "zvo_eigenvec_0_rank_0.dat")
    capi.write_vec_file(p, r.iterations, v)
    back, st = capi.read_vec_file(p, eng.dim)
    assert st == r.iterations and np.array_equal(back, v)
    ph = eng.expec_energy_flct(dv)
    capi.write_energy(str(tmp_path / "zvo_energy.dat"), ph)
    e_line = open(str(tmp_path / "zvo_energy.dat")).readline().split()
    assert e_line[0] == "Energy" and abs(float(e_line[1]) - ph["energy"]) <= 1e-15
    g1 = [(i, s, j, s) for i in range(4) for j in range(4) for s in (0, 1)]
    vals = eng.cisajs(dv, g1)
    capi.write_cisajs(str(tmp_path / "zvo_cisajs.dat"), g1, vals)
    rows = np.loadtxt(str(tmp_path / "zvo_cisajs.dat"))
    assert np.allclose(rows[:, 4] + 1j * rows[:, 5], vals, atol=1e-10)
    assert abs(sum(vals[k] for k, e in enumerate(g1) if e[0] == e[2]) - spec.ncond) <= 1e-10  # sum of n
    want = [brute.green1(spec, v, e) for e in g1[:6]]
    assert np.allclose(vals[:6], want, atol=1e-13)
    dv.free()
    eng.close()


@pytest.mark.gpu
def test_lanczos_restart_continues_the_recurrence(A, tmp_path):
    """30 steps written to TMComponents + recalcvec, continued for 40 more,
    equal a straight 70-step run: the first 30 components identically, the rest
    and the eigenvalues to the precision of the 10-decimal alpha, beta the
    continued run starts from (as the reference's restart)."""
    print("[gpu] lanczos restart", flush=True)
    spec = L.heisenberg_chain(20, 1.0, 0.3)  # dim 184,756: 70 steps stop short of machine-precision convergence
    eng = A.Engine.from_spec(spec)
    f = lambda n: (str(tmp_path / f"tm{n}.dat"), str(tmp_path / f"vec{n}.dat"))
    kw = dict(n_eigen=1, eps=1e-300, target=0, seed=4)
    eng.lanczos_restart(restart_out=f(1), max_iter=30, **kw)
    rb = eng.lanczos_restart(restart_in=f(1), restart_out=f(2), max_iter=40, **kw)
    rc = eng.lanczos_restart(restart_out=f(3), max_iter=70, **kw)
    _, a2, b2 = capi.read_tmcomponents(f(2)[0])
    _, a3, b3 = capi.read_tmcomponents(f(3)[0])
    assert len(a2) == len(a3) == 70 and rb.iterations == rc.iterations == 70
    assert np.array_equal(a2[:30], a3[:30]) and np.array_equal(b2[:30], b3[:30])
    assert np.allclose(a2, a3, rtol=0, atol=1e-8) and np.allclose(b2, b3, rtol=0, atol=1e-8)
    assert abs(rb.eigenvalues[0] - rc.eigenvalues[0]) <= 1e-9
    _, _, st = capi.read_restart_vecs(f(2)[1], eng.dim)
    assert st == 70
    with pytest.raises(A.HphiError, match="holds step"):  # mismatched pair
        eng.lanczos_restart(restart_in=(f(1)[0], f(2)[1]), max_iter=5, **kw)
    eng.close()
