"""Golden cases: small models whose tables, diagonal and H·v were produced by
the compiled reference (oracle/_ref) with tests/golden/gen_golden.py."""

from hphi_amd import lattice as L

SEED = 7  # synthetic-vector seed used for every golden H·v


def golden_cases():
    rng_j2 = 0.3710921
    out = {
        "chain16": L.baseline_config(1),                         # workload 1
        "chain13_sz1": L.heisenberg_chain(13, 1.0, 0.3, two_sz=1),
        "ladder12": L.j1j2_ladder(6, 1.0, 0.47),
        "j1j2_4x4": L.j1j2_square(4, 4, 1.0, rng_j2),
        "j1j2_4x3_sz2": L.j1j2_square(4, 3, 1.0, 0.5, two_sz=2),
        "hubbard_chain4": L.hubbard_chain(4),
        "hubbard_chain5": L.hubbard_chain(5),
        "hubbard_4x2": L.hubbard_square(4, 2, ncond=8),
        "hubbard_3x2_n5": L.hubbard_square(3, 2, t=1.0, U=2.5, ncond=5, two_sz=1),
        "tri_4x2_gc": L.triangular_xxz_gc(4, 2, 1.0, 0.8, 0.1),
        "tri_4x3_gc": L.triangular_xxz_gc(4, 3, 1.0, 0.8, 0.1),
    }
    # longitudinal field on a spin chain exercises EDChemi; complex hopping the
    # complex transfer path
    s = L.heisenberg_chain(10, 1.0, 0.0, two_sz=2)
    s.trans = [(i, 0, i, 0, 0.05 * (i + 1), 0.0) for i in range(10)]
    s.name = "chain10_field"
    out["chain10_field"] = s
    h = L.hubbard_chain(6, 1.0, 3.0)
    h.trans = []
    for i in range(6):
        j = (i + 1) % 6
        for sp in (0, 1):
            h.trans.append((i, sp, j, sp, 1.0, 0.2))
            h.trans.append((j, sp, i, sp, 1.0, -0.2))
    h.name = "hubbard_chain6_complex"
    out["hubbard_chain6_complex"] = h
    return out


def hubbard_two_body():
    """Hubbard ring of 6 sites at half filling with every two-body kind the
    engine implements: InterAll n n (with a duplicate to merge), density times
    hop in both operator orders, two hops of opposite spin, a spin-exchange
    product (rewritten by the reader), c+_i c_j c+_j c_i (diagonal + chemical
    potential), PairHop and Hubbard Exchange; complex couplings where the term
    allows them."""
    h = L.hubbard_chain(6, 1.0, 3.0)
    v1, v2, v3 = complex(0.35, 0.2), complex(-0.25, 0.15), complex(0.3, -0.1)
    h.interall = [
        (0, 0, 0, 0, 1, 1, 1, 1, 0.7, 0.0),             # 0.7 n_0up n_1dn
        (0, 0, 0, 0, 1, 1, 1, 1, 0.1, 0.0),             # duplicate: merged to 0.8
        (3, 1, 3, 1, 1, 0, 2, 0, v1.real, v1.imag),     # n_3dn c+_1up c_2up
        (2, 0, 1, 0, 3, 1, 3, 1, v1.real, -v1.imag),    # its Hermitian conjugate
        (0, 0, 4, 0, 2, 1, 5, 1, v2.real, v2.imag),     # c+_0up c_4up c+_2dn c_5dn
        (5, 1, 2, 1, 4, 0, 0, 0, v2.real, -v2.imag),
        (0, 0, 2, 1, 3, 1, 1, 0, v3.real, v3.imag),     # c+_0up c_2dn c+_3dn c_1up (rewritten)
        (2, 1, 3, 1, 1, 0, 0, 0, -v3.real, v3.imag),    # its conjugate, already in reader order
        (4, 1, 5, 1, 5, 1, 4, 1, 0.45, 0.0),            # c+_4dn c_5dn c+_5dn c_4dn
        (1, 0, 1, 0, 3, 1, 5, 1, -0.22, 0.0),           # n_1up c+_3dn c_5dn: high-half flip, low-half test
        (5, 1, 3, 1, 1, 0, 1, 0, -0.22, 0.0),
    ]
    h.pairhop = [(0, 2, 0.3), (1, 4, -0.2)]
    h.exchange = [(1, 2, 0.4), (3, 5, -0.15)]
    h.name = "hubbard_two_body"
    return h
