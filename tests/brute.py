"""Dense Hamiltonians from the textbook operator definitions (test helper).

Independent of the reference's code structure: states are enumerated by
brute force, and every operator is applied from its definition
(doc/en/source/filespecification/expertmode_en/*_file_en.rst):

* Exchange J:  J (S+_i S-_j + S-_i S+_j): couples states whose bits i, j differ
* Ising J:     J Sz_i Sz_j with Sz = +-1/2
* Trans t:     -t c+_{i s1} c_{j s2}; fermion order = bit index 2i+s, so
               c+_a c_b |x> = (-1)^(#occupied strictly between a, b) |x'>
* spin Trans:  on one site, c+_{i s1} c_{i s2} flips the bit s2 -> s1
* CoulombIntra U: U n_up n_down; CoulombInter V n_i n_j;
  Hund -J (n_iup n_jup + n_idn n_jdn)
* spin InterAll: (|s1><s2|)_i (|s3><s4|)_k, the right factor applied first
Used only for small sizes.
"""

from __future__ import annotations

from itertools import combinations
from typing import List

import numpy as np

from hphi_amd import lattice as L


def basis(spec) -> List[int]:
    n = spec.nbits
    if spec.model in (L.SPIN_GC, L.HUBBARD_GC):
        return list(range(1 << n))
    out = []
    for x in range(1 << n):
        if spec.model == L.SPIN:
            if bin(x).count("1") == spec.nup:
                out.append(x)
        else:
            up = sum((x >> (2 * i)) & 1 for i in range(spec.nsite))
            dn = sum((x >> (2 * i + 1)) & 1 for i in range(spec.nsite))
            if up == spec.nup and dn == spec.ndown:
                out.append(x)
    return out


def dense(spec) -> np.ndarray:
    states = basis(spec)
    index = {x: k for k, x in enumerate(states)}
    D = len(states)
    H = np.zeros((D, D), np.complex128)
    for k, x in enumerate(states):
        if spec.model in (L.SPIN, L.SPIN_GC):
            for i, j, J in spec.exchange:
                if ((x >> i) & 1) != ((x >> j) & 1):
                    H[index[x ^ (1 << i) ^ (1 << j)], k] += J
            for i, j, J in spec.ising:
                si = 0.5 if (x >> i) & 1 else -0.5
                sj = 0.5 if (x >> j) & 1 else -0.5
                H[k, k] += J * si * sj
            # InterAll: c+_{i s1} c_{i s2} on a spin-1/2 site is |s1><s2| (the
            # site's state is its bit); the right factor acts first
            for i, s1, j, s2, kk, s3, l, s4, re, im in getattr(spec, "interall", []):
                y = x
                if ((y >> kk) & 1) != s4:
                    continue
                y = (y | (1 << kk)) if s3 else (y & ~(1 << kk))
                if ((y >> i) & 1) != s2:
                    continue
                y = (y | (1 << i)) if s1 else (y & ~(1 << i))
                if y in index:
                    H[index[y], k] += complex(re, im)
            for i, s1, j, s2, re, im in spec.trans:
                t = complex(re, im)
                if i == j and s1 != s2:
                    if ((x >> i) & 1) == s2:
                        H[index[x ^ (1 << i)], k] += -t
                elif i == j and s1 == s2:
                    # longitudinal field: -t n_{i s1}, n_{i s} = [bit_i == s]
                    if ((x >> i) & 1) == s1:
                        H[k, k] += -t
        else:
            for i, s1, j, s2, re, im in spec.trans:
                t = complex(re, im)
                a, b = 2 * i + s1, 2 * j + s2
                if a == b:
                    if (x >> a) & 1:
                        H[k, k] += -t
                    continue
                if not ((x >> b) & 1) or ((x >> a) & 1):
                    continue
                lo, hi = min(a, b), max(a, b)
                between = bin(x & (((1 << hi) - 1) & ~((1 << (lo + 1)) - 1))).count("1")
                y = x ^ (1 << a) ^ (1 << b)
                if y in index:
                    H[index[y], k] += -t * (-1) ** between
            for i, U in spec.coulomb_intra:
                if (x >> (2 * i)) & 1 and (x >> (2 * i + 1)) & 1:
                    H[k, k] += U
            n = [((x >> (2 * i)) & 1, (x >> (2 * i + 1)) & 1) for i in range(spec.nsite)]
            for i, j, V in spec.coulomb_inter:  # V n_i n_j
                H[k, k] += V * sum(n[i]) * sum(n[j])
            for i, j, J in spec.hund:  # -J (n_iup n_jup + n_idn n_jdn)
                H[k, k] += -J * (n[i][0] * n[j][0] + n[i][1] * n[j][1])
            # two-body terms from their definitions
            # (doc/en/source/filespecification/expertmode_en/{InterAll,PairHop,Exchange}_file_en.rst)
            terms = []
            for i, s1, j, s2, kk, s3, l, s4, re, im in getattr(spec, "interall", []):
                terms.append((complex(re, im), [(1, 2 * i + s1), (0, 2 * j + s2), (1, 2 * kk + s3), (0, 2 * l + s4)]))
            for i, j, J in getattr(spec, "pairhop", []):
                terms.append((J, [(1, 2 * i), (0, 2 * j), (1, 2 * i + 1), (0, 2 * j + 1)]))
                terms.append((J, [(1, 2 * j), (0, 2 * i), (1, 2 * j + 1), (0, 2 * i + 1)]))
            for i, j, J in spec.exchange:
                terms.append((J, [(1, 2 * i), (0, 2 * j), (1, 2 * j + 1), (0, 2 * i + 1)]))
                terms.append((J, [(1, 2 * i + 1), (0, 2 * j + 1), (1, 2 * j), (0, 2 * i)]))
            for V, ops in terms:
                y, sg = apply_ops(x, ops)
                if sg and y in index:
                    H[index[y], k] += V * sg
    return H


def apply_ops(x: int, ops):
    """Product of fermion operators (dagger, orbital), written left to right and
    applied right to left, with the sign (-1)^(occupied orbitals below) of each
    operator (orbital order 2i + sigma).  Returns (state, sign) or (None, 0)."""
    sign = 1
    for dag, a in reversed(ops):
        if ((x >> a) & 1) == dag:  # create on an occupied or annihilate an empty orbital
            return None, 0
        if bin(x & ((1 << a) - 1)).count("1") & 1:
            sign = -sign
        x ^= 1 << a
    return x, sign


def kron_xxz_gc(nsite, bonds, Jz, Jxy, hx) -> np.ndarray:
    """SpinGC XXZ + transverse field from Pauli-matrix Kronecker products."""
    sx = np.array([[0, 0.5], [0.5, 0]])
    sy = np.array([[0, -0.5j], [0.5j, 0]])
    sz = np.array([[0.5, 0], [0, -0.5]])

    def op(o, i):
        m = np.array([[1.0]])
        for s in range(nsite):
            m = np.kron(m, o if s == i else np.eye(2))
        return m

    H = np.zeros((1 << nsite, 1 << nsite), np.complex128)
    for i, j in bonds:
        H += Jz * op(sz, i) @ op(sz, j)
        H += Jxy * (op(sx, i) @ op(sx, j) + op(sy, i) @ op(sy, j))
    for i in range(nsite):
        H += -hx * op(sx, i)
    return H


def _spin_ops(x: int, ops):
    """(|a><b|)_site products on a spin-1/2 state, applied right to left."""
    for site, a, b in reversed(ops):
        if ((x >> site) & 1) != b:
            return None
        x = (x | (1 << site)) if a else (x & ~(1 << site))
    return x


def green1(spec, v, entry) -> complex:
    """<v|c+_{i s1} c_{j s2}|v> from the definitions; for spins c+_{i s1} c_{i s2}
    is |s1><s2| on site i and an entry with i != j is 0 (localized spins)."""
    states = basis(spec)
    index = {x: k for k, x in enumerate(states)}
    i, s1, j, s2 = entry
    acc = 0.0
    for k, x in enumerate(states):
        if spec.model in (L.SPIN, L.SPIN_GC):
            if i != j:
                return 0.0
            y, sg = _spin_ops(x, [(i, s1, s2)]), 1
        else:
            y, sg = apply_ops(x, [(1, 2 * i + s1), (0, 2 * j + s2)])
        if y is not None and sg and y in index:
            acc += np.conj(v[index[y]]) * sg * v[k]
    return acc


def green2(spec, v, entry) -> complex:
    """<v|c+_{i s1} c_{j s2} c+_{k s3} c_{l s4}|v> from the definitions (spins:
    i == j, k == l, (|s1><s2|)_i (|s3><s4|)_k)."""
    states = basis(spec)
    index = {x: k for k, x in enumerate(states)}
    i, s1, j, s2, kk, s3, l, s4 = entry
    acc = 0.0
    for k, x in enumerate(states):
        if spec.model in (L.SPIN, L.SPIN_GC):
            y, sg = _spin_ops(x, [(i, s1, s2), (kk, s3, s4)]), 1
        else:
            y, sg = apply_ops(x, [(1, 2 * i + s1), (0, 2 * j + s2), (1, 2 * kk + s3), (0, 2 * l + s4)])
        if y is not None and sg and y in index:
            acc += np.conj(v[index[y]]) * sg * v[k]
    return acc


def phys(spec, v, H=None) -> dict:
    """energy <v|H|v>, var <Hv|Hv>, and the doublon / electron number / Sz
    moments (Sz = (N_up - N_down)/2) of a normalised v; canonical Spin reports
    its sector's Sz."""
    H = dense(spec) if H is None else H
    hv = H @ v
    out = {"energy": float(np.vdot(v, hv).real), "var": float(np.vdot(hv, hv).real)}
    w = np.abs(v) ** 2
    states = basis(spec)
    N = spec.nsite
    if spec.model in (L.SPIN, L.SPIN_GC):
        out.update(doublon=0.0, doublon2=0.0, num=float(N), num2=float(N * N))
        if spec.model == L.SPIN:
            sz = 0.5 * spec.two_sz
            out.update(Sz=sz, Sz2=sz * sz)
        else:
            S = np.array([2 * bin(x).count("1") - N for x in states], float)
            out.update(Sz=0.5 * np.sum(w * S), Sz2=0.25 * np.sum(w * S * S))
    else:
        up = np.array([sum((x >> (2 * i)) & 1 for i in range(N)) for x in states], float)
        dn = np.array([sum((x >> (2 * i + 1)) & 1 for i in range(N)) for x in states], float)
        D = np.array([sum((x >> (2 * i)) & (x >> (2 * i + 1)) & 1 for i in range(N)) for x in states], float)
        out.update(doublon=np.sum(w * D), doublon2=np.sum(w * D * D), num=np.sum(w * (up + dn)),
                   num2=np.sum(w * (up + dn) ** 2), Sz=0.5 * np.sum(w * (up - dn)),
                   Sz2=0.25 * np.sum(w * (up - dn) ** 2))
    return out


def green_nbody(spec, v, entry):
    """<v|O_1 O_2 ... O_n|v> for a SpinGC three/four/six-body entry (n pairs
    (i, s1, j, s2)) from the definitions the reference states in
    expec_cisajscktaltdc.c: O = (|s1><s2|)_i for i == j, a later pair on two
    sites gives 0, the first two pairs as Rearray_Interactions takes them
    ((i,i)(k,k), or (i,k)(k,i) as -(|s1><s4|)_i (|s3><s2|)_k).  None when the
    reference does not evaluate the entry."""
    q = list(entry)
    n = len(q) // 4
    if q[0] == q[2] and q[4] == q[6]:
        ops, sign = [(q[0], q[1], q[3]), (q[4], q[5], q[7])], 1.0
    elif q[0] == q[6] and q[4] == q[2]:
        ops, sign = [(q[0], q[1], q[7]), (q[4], q[5], q[3])], -1.0
    else:
        return None
    for p in range(2, n):
        i, s1, j, s2 = q[4 * p:4 * p + 4]
        if i != j:
            return 0.0
        ops.append((i, s1, s2))
    acc = 0.0
    for x in range(len(v)):
        y = _spin_ops(x, ops)
        if y is not None:
            acc += np.conj(v[y]) * v[x]
    return sign * acc
