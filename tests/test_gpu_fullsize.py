"""Full-size parity: the H·v of the baseline workloads (Hubbard 4x4, J1-J2
8x4, triangular 13x2 SpinGC) on the synthetic vector of their seed, against
checksums and sampled entries of the reference's mltply on the same vector
(tests/golden/golden.json "large", written by gen_golden.py --large from the
reference built in oracle/_ref)."""

import json
import os

import numpy as np
import pytest

from hphi_amd import lattice as L

pytestmark = pytest.mark.gpu

TOL = 1e-12
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json"))).get("large", {})


@pytest.fixture(scope="module")
def A():
    from hphi_amd import capi
    capi.lib()
    return capi


def _checksums(h):
    """S1 = sum v0, S2 = sum |v0|^2, S3 = sum w_j v0_j with w_j = (j mod 97) + 1
    (j 1-based), and the L1 bounds that scale their rounding; in chunks."""
    s1 = 0j
    s2 = 0.0
    s3 = 0j
    l1 = 0.0
    l1w = 0.0
    step = 1 << 24
    for a in range(0, h.size, step):
        c = h[a:a + step]
        j = np.arange(a + 1, a + 1 + c.size, dtype=np.int64)
        w = (j % 97 + 1).astype(np.float64)
        ab = np.abs(c)
        s1 += c.sum()
        s2 += float(np.dot(ab, ab))
        s3 += np.dot(w, c)
        l1 += ab.sum()
        l1w += np.dot(w, ab)
    return s1, s2, s3, l1, l1w


@pytest.mark.parametrize("n,tuning", [(4, {}), (2, {}), (3, {}), (3, {"c2_crank": 1, "tpb": 256, "wgs_per_cu": 8}),
                                      (3, {"c2_crank": 0, "lds21": 1, "tpb": 1024, "wgs_per_cu": 1}),
                                      (2, {"c2_crank": 1, "tpb": 256, "wgs_per_cu": 48})],
                         ids=["cfg4", "cfg2", "cfg3", "cfg3-rank-tables", "cfg3-lds-table", "cfg2-rank-tables"])
def test_full_size_hv_matches_reference_checksums(A, n, tuning):
    rec = GOLD.get(f"cfg{n}")
    if rec is None:
        pytest.skip(f"no reference checksums for workload {n}")
    spec = L.baseline_config(n)
    print(f"[gpu] full-size H·v workload {n} dim {spec.dim} {tuning}", flush=True)
    eng = A.Engine.from_spec(spec)
    eng.tune(**tuning)
    dim = eng.dim
    assert dim == rec["dim"]
    v1 = A.DeviceArray(dim)
    eng.fill_random(v1, 1, rec["seed"])
    v0 = A.DeviceArray(dim).zero()
    pr = eng.mltply_dev(v0, v1)
    v1.free()
    h = v0.download()
    v0.free()
    eng.close()
    s1, s2, s3, l1, l1w = _checksums(h)
    assert abs(s2 - rec["S2"]) <= TOL * rec["S2"]
    assert abs(s1 - complex(*rec["S1"])) <= TOL * l1
    assert abs(s3 - complex(*rec["S3"])) <= TOL * l1w
    amax = float(np.max(np.abs(h)))
    for j, re, im in rec["samples"]:
        assert abs(h[j - 1] - complex(re, im)) <= TOL * amax, j
    # prdct = <v1|H v1>: rounding scales with |v1| |H v1| = sqrt(2 dim / 3 * S2) for this vector
    scale = np.sqrt(2.0 * dim / 3.0 * s2)
    assert abs(pr - complex(*rec["prdct"])) <= TOL * scale


def test_full_size_block16_matches_reference_checksums(A):
    """The k = 16 block H·v of the triangular 13x2 workload in its default
    grand-canonical shape (the kernel of its mTPQ step): column c is the
    synthetic vector of seed c, checked column by column against the
    reference's mltply on that vector (gen_golden.py --large-block)."""
    rec = GOLD.get("cfg4_k16")
    if rec is None:
        pytest.skip("no reference block checksums")
    k = rec["k"]
    spec = L.baseline_config(4)
    print(f"[gpu] full-size block H·v workload 4 dim {spec.dim} k={k}", flush=True)
    eng = A.Engine.from_spec(spec)
    dim = eng.dim
    assert dim == rec["dim"]
    V1 = A.DeviceArray(dim * k)
    eng.fill_random(V1, k, rec["seed"])
    V0 = A.DeviceArray(dim * k).zero()
    pr = eng.mltply_block_dev(k, V0, V1)
    V1.free()
    h = V0.download().reshape(dim, k)
    V0.free()
    eng.close()
    # per-column checksums in one pass over the block
    s1 = np.zeros(k, np.complex128)
    s2 = np.zeros(k)
    s3 = np.zeros(k, np.complex128)
    l1 = np.zeros(k)
    l1w = np.zeros(k)
    step = 1 << 22
    for a in range(0, dim, step):
        c = h[a:a + step]
        j = np.arange(a + 1, a + 1 + c.shape[0], dtype=np.int64)
        w = (j % 97 + 1).astype(np.float64)
        ab = np.abs(c)
        s1 += c.sum(axis=0)
        s2 += (ab * ab).sum(axis=0)
        s3 += w @ c
        l1 += ab.sum(axis=0)
        l1w += w @ ab
    for col, cr in enumerate(rec["columns"]):
        assert abs(s2[col] - cr["S2"]) <= TOL * cr["S2"], col
        assert abs(s1[col] - complex(*cr["S1"])) <= TOL * l1[col], col
        assert abs(s3[col] - complex(*cr["S3"])) <= TOL * l1w[col], col
        amax = float(np.max(np.abs(h[:, col])))
        for jj, re, im in cr["samples"]:
            assert abs(h[jj - 1, col] - complex(re, im)) <= TOL * amax, (col, jj)
        scale = np.sqrt(2.0 * dim / 3.0 * s2[col])
        assert abs(pr[col] - complex(*cr["prdct"])) <= TOL * scale, col


LZ = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "lanczos_ref.json")))


@pytest.mark.parametrize("n", [2, 3], ids=["cfg2", "cfg3"])
def test_full_size_lanczos_matches_reference_recurrence(A, n):
    """Lanczos at full size against the reference's own recurrence
    (Lanczos_EigenValue.c:102-177 driven by the reference's mltply on the CPU,
    tests/golden/gen_lanczos.py): from the same normalised start vector, the
    lowest two eigenvalues of T_m after the reference's m steps agree to 1e-10,
    and, where the reference ran to convergence (Hubbard 4x4), the device's
    converged lowest distinct eigenvalues (check every 5 steps, 1e-12) equal the
    reference's converged values to 1e-10."""
    from hphi_amd.capi import lib  # noqa: F401
    ref = LZ.get(f"cfg{n}")
    if ref is None:
        pytest.skip(f"no reference recurrence for workload {n}")
    m = ref["steps"]
    spec = L.baseline_config(n)
    print(f"[gpu] full-size Lanczos workload {n} dim {spec.dim}, {m} steps", flush=True)
    eng = A.Engine.from_spec(spec)
    assert eng.dim == ref["dim"]
    r = eng.lanczos(n_eigen=2, max_iter=m, eps=1e-300, target=1, seed=ref["seed"], check_every=m, mode=0)
    assert r.iterations == m
    got = np.array(r.eigenvalues[:2])
    want = np.array(ref["ritz"][str(m)][:2])
    assert np.max(np.abs(got - want)) <= 1e-10, (got, want)
    if ref.get("converged"):
        rd = eng.lanczos(n_eigen=2, max_iter=m, eps=1e-12, seed=ref["seed"], check_every=5, mode=1)
        assert rd.converged
        conv = np.array(ref["converged"][:2])
        assert np.max(np.abs(np.array(rd.eigenvalues[:2]) - conv)) <= 1e-10, (rd.eigenvalues[:2], conv)
    eng.close()
