"""The drop-in binding (integration/mltply_hphi_amd.c) inside the reference.

oracle/_ref/ref_dropin is the reference's start-up sequence and callers
compiled from its sources, linked with the binding in place of src/mltply.c
and with libhphi_amd (oracle/Makefile target `dropin`; built here, where the
reference tree is, and carried to the GPU with the other built files).

* CPU: the binding's descriptor of the reference's X->Def (after
  ReadDefFileIdxPara and HPhiTrans) equals hphi_read_def's of the same files,
  field by field, for every fixture model (two-body terms included).
* GPU: the reference's own code paths run unchanged on the GPU engine:
  one mltply against the reference's H·v fixtures; CalcByTPQ and
  CalcByCanonicalTPQ against the files the reference wrote on the CPU (same
  start vectors: MakeIniVec with the 8 threads the fixtures were made with);
  the Lanczos recurrence of Lanczos_EigenValue.c against the reference's
  alpha/beta.
"""

import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from hphi_amd import lattice as L
from tests.golden.cases import SEED, ctpq_file_cases, golden_cases, tpq_file_cases, two_body_cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "oracle", "_ref", "ref_dropin")
GOLD = os.path.join(ROOT, "tests", "golden")
FILES = os.path.join(GOLD, "files")

needs_dropin = pytest.mark.skipif(not os.path.exists(DROPIN), reason="oracle/_ref/ref_dropin is not built "
                                  "(it needs the reference tree: make -C oracle dropin)")


def run(def_dir, mode, seed=0, reps=1, threads=8, timeout=300):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([DROPIN, def_dir, mode, os.path.join(def_dir, "out"), str(seed), str(reps)],
                       capture_output=True, text=True, env=env, timeout=timeout)
    return r


def _all_cases():
    return {**two_body_cases(), **golden_cases()}


@needs_dropin
@pytest.mark.parametrize("name", sorted(_all_cases()))
def test_binding_descriptor_equals_reader(name, tmp_path):
    spec = _all_cases()[name]
    L.write_expert(spec, str(tmp_path))
    r = run(str(tmp_path), "desc")
    assert r.returncode == 0 and "desc ok" in r.stdout, r.stdout[-2000:]


# ------------------------------------------------------------------ GPU ----

def _table(path):
    return np.array([list(map(float, ln.split())) for ln in open(path) if not ln.startswith(" #")])


@needs_dropin
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["chain13_sz1", "hubbard_4x2", "tri_4x3_gc", "hubbard_chain6_complex"])
def test_reference_mltply_call_on_gpu(name):
    spec = golden_cases()[name]
    print(f"[gpu] drop-in hv {name}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        L.write_expert(spec, d)
        r = run(d, "hv", seed=SEED)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        hv = np.fromfile(os.path.join(d, "out", "v0.bin"), np.complex128)
        pr = open(os.path.join(d, "out", "prdct.txt")).read().split()
    rec = json.load(open(os.path.join(GOLD, "golden.json")))["cases"][name]
    ref = np.load(os.path.join(GOLD, rec["arrays"]))["hv"]
    assert np.max(np.abs(hv - ref)) <= 1e-12 * np.max(np.abs(ref))
    p = complex(float(pr[0]), float(pr[1]))
    assert abs(p - complex(*rec["prdct"])) <= 1e-12 * np.linalg.norm(ref) * np.sqrt(len(ref))


@needs_dropin
@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted({**tpq_file_cases(), **ctpq_file_cases()}))
def test_reference_tpq_drivers_on_gpu(name):
    """CalcByTPQ / CalcByCanonicalTPQ unchanged, every mltply on the GPU: the
    SS_rand, Norm_rand and Flct_rand numbers equal the reference's CPU run."""
    canonical = name.startswith("ctpq_")
    spec = (ctpq_file_cases() if canonical else tpq_file_cases())[name]
    print(f"[gpu] drop-in {name}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        L.write_expert(spec, d, calc_type=L.CALC_CTPQ if canonical else L.CALC_TPQ)
        r = run(d, "ctpq" if canonical else "tpq")
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        for i in range(int(spec.params["NumAve"])):
            for kind in ("SS", "Norm", "Flct"):
                got = _table(os.path.join(d, "output", f"{kind}_rand{i}.dat"))
                ref = _table(os.path.join(FILES, name, f"{kind}_rand{i}.dat"))
                assert got.shape == ref.shape, (kind, i)
                assert np.array_equal(got[:, -1], ref[:, -1])  # step column
                assert np.allclose(got, ref, rtol=1e-10, atol=1e-12), (kind, i, got[:3], ref[:3])


@needs_dropin
@pytest.mark.gpu
def test_reference_lanczos_recurrence_on_gpu():
    """The recurrence of Lanczos_EigenValue.c:102-177 with every mltply on the
    GPU: alpha/beta of the first 20 steps and the lowest Ritz value after 80
    equal the reference's (chain 16, Sz = 0, seed 0)."""
    from scipy.linalg import eigvalsh_tridiagonal
    ref = json.load(open(os.path.join(GOLD, "lanczos_ref.json")))["cfg1"]
    print("[gpu] drop-in lanczos chain16", flush=True)
    with tempfile.TemporaryDirectory() as d:
        L.write_expert(L.baseline_config(1), d)
        r = run(d, "lanczos", seed=ref["seed"], reps=ref["steps"])
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        ab = np.loadtxt(os.path.join(d, "out", "ab.txt"))
    a, b = ab[:, 1], ab[:, 2]
    ra, rb = np.array(ref["alpha"]), np.array(ref["beta"])
    assert np.allclose(a[:20], ra[:20], rtol=1e-10) and np.allclose(b[:20], rb[:20], rtol=1e-10)
    m = len(a)
    e = eigvalsh_tridiagonal(a, b[:m - 1])[0]
    assert abs(e - ref["ritz"][str(m)][0]) <= 1e-10
