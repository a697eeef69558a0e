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
