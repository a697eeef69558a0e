"""The byte and element counts behind bench.py's roofline figures, against the
values worked out from the Hamiltonians' definitions: off-diagonal elements per
H·v of the J1-J2 8x4 Sz=0 sector (128 exchange bonds x 2 C(30,15)) and of the
Hubbard 4x4 half-filled sector (128 directed hops x C(14,7) C(16,8)), and the
algorithmic bytes per H·v (16 per element + 64 per state)."""

import math

from hphi_amd import lattice as L


def test_plain_counts_and_bytes():
    import bench
    s3, s2 = L.baseline_config(3), L.baseline_config(2)
    assert s3.dim == math.comb(32, 16) == 601_080_390
    assert s2.dim == math.comb(16, 8) ** 2 == 165_636_900
    assert bench.plain_nnz(s3) == 128 * 2 * math.comb(30, 15) == 39_710_085_120
    assert bench.plain_nnz(s2) == 128 * math.comb(14, 7) * math.comb(16, 8) == 5_653_739_520
    assert bench.b_alg(s3.dim, bench.plain_nnz(s3)) == 673_830_506_880
    assert bench.b_alg(s2.dim, bench.plain_nnz(s2)) == 101_060_593_920


def test_triangular_torus_pairs():
    """The 13x2 torus lists 78 exchange entries for 65 distinct pairs (the a2
    bonds wrap onto themselves when Ly = 2); the kernels gather once per pair."""
    s4 = L.baseline_config(4)
    assert s4.dim == 2 ** 26
    assert len(s4.exchange) == 78
    assert len({(min(i, j), max(i, j)) for i, j, _ in s4.exchange}) == 65


def test_library_nnz_count_matches_closed_forms():
    """hphi_model_nnz (host only, per-term combinatorics in the library) gives
    the closed forms of the two plain workloads."""
    from hphi_amd import capi
    s3, s2 = L.baseline_config(3), L.baseline_config(2)
    assert capi.Engine.model_nnz(s3) == (s3.dim, 128 * 2 * math.comb(30, 15))
    assert capi.Engine.model_nnz(s2) == (s2.dim, 128 * math.comb(14, 7) * math.comb(16, 8))


def test_library_nnz_count_matches_dense_operators():
    """The same count equals the number of nonzero off-diagonal entries of the
    dense Hamiltonian built column by column from the oracle's H·v, for spin
    and Hubbard sectors with odd fillings, complex hops and two-body terms."""
    import numpy as np

    from hphi_amd import capi
    from oracle import oracle as O
    from tests.golden.cases import golden_cases, two_body_cases
    cases = {**golden_cases(), **two_body_cases()}
    checked = 0
    for name, spec in cases.items():
        if spec.model not in (0, 1) or spec.dim > 400:
            continue
        if name in two_body_cases():
            continue  # the C restatement has no two-body terms
        orc = O.Oracle(spec, L.processed_terms(spec)).prepare()
        H = np.array([orc.hv(e)[0] for e in np.eye(orc.dim, dtype=np.complex128)]).T
        off = H - np.diag(np.diag(H))
        dim, nnz = capi.Engine.model_nnz(spec)
        assert dim == orc.dim
        # structural count: entries can cancel numerically only if two terms hit
        # the same pair, which the merged term table rules out for these models
        assert nnz == int(np.count_nonzero(off)), name
        checked += 1
    assert checked >= 5


def test_stratified_sequential_baseline():
    """The sequential baseline of the served stream weights each structure
    class's mean solve time by its count: with every request sampled it is the
    plain mean; with a uniform time per class it is independent of how many of
    each class were sampled."""
    import numpy as np

    import bench
    rng = np.random.default_rng(3)
    reqs = [L.stream_request(rng) for _ in range(2000)]
    by = bench.stream_classes(reqs)
    assert sum(len(v) for v in by.values()) == 2000 and len(by) == 32
    assert all(reqs[i].name == name for name, ids in by.items() for i in ids)
    t = {i: 1e-3 * (1 + reqs[i].dim ** 0.5 / 100) + 1e-6 * (i % 7) for i in range(2000)}  # any per-request time
    exact = sum(t.values()) / 2000
    full = {name: sum(t[i] for i in ids) / len(ids) for name, ids in by.items()}
    assert abs(bench.stratified_seconds_per_request(by, full) - exact) < 1e-15
    # a class-constant time is recovered exactly from one sample per class
    const = {name: 1e-3 * (k + 1) for k, name in enumerate(by)}
    want = sum(const[reqs[i].name] for i in range(2000)) / 2000
    assert abs(bench.stratified_seconds_per_request(by, const) - want) < 1e-15
