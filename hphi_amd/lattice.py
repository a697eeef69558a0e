"""Model generator and expert-mode writer for the H·v engine.

The reference builds its expert-mode input with StdFace, an un-vendored
submodule (``src/StdFace`` is empty in the reference tree), so this module
generates the same kind of input from first principles:

* lattices (chain, 2-leg ladder, square J1-J2, triangular, square Hubbard)
  as explicit bond lists;
* couplings mapped to expert-mode terms exactly as the reference's file
  specifications define them
  (``doc/en/source/filespecification/expertmode_en/{Exchange,Ising,Trans,
  CoulombIntra}_file_en.rst``):

  - Heisenberg ``J S_i·S_j``      -> Exchange ``J/2`` + Ising ``J``
  - XXZ ``Jz SzSz + Jxy(SxSx+SySy)`` -> Exchange ``Jxy/2`` + Ising ``Jz``
  - transverse field ``-h Σ S^x``  -> Trans ``(i,0,i,1,h/2)`` and ``(i,1,i,0,h/2)``
    (Trans adds ``-t c†c``; for spin-1/2 ``c†_{i1}c_{i0} = S^+`` up to the
    bit convention, and the spectrum is even in ``h``)
  - Hubbard hopping ``-t Σ c†c``   -> Trans ``(i,σ,j,σ,t)`` + ``(j,σ,i,σ,t)``
  - on-site ``U n↑n↓``             -> CoulombIntra ``U``

``write_expert`` writes ``namelist.def`` and the referenced files in the
formats the reference reader parses (``src/readdef.c:463-1005`` for counts,
``src/readdef.c:1006-2059`` for the term lines).  ``processed_terms``
restates the reader's post-processing (``readdef.c:1130-1160`` merge of
duplicate transfers, ``readdef.c:2136-2234`` Hermitian pairing into
``EDGeneralTransfer``/``EDChemi``, ``readdef.c:1300-1327`` Ising ->
Hund(-J/2) + CoulombInter(-J/4) appended after the explicit entries,
``readdef.c:2788-2792``); the C reader in the engine implements the same
rules independently and tests compare the two.
"""

from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# calcmod.def CalcModel values (reference include/DefCommon.h:22-37)
HUBBARD = 0
SPIN = 1
KONDO = 2
HUBBARD_GC = 3
SPIN_GC = 4
KONDO_GC = 5

# calcmod.def CalcType values
CALC_LANCZOS = 0
CALC_TPQ = 1
CALC_FULLDIAG = 2
CALC_CG = 3
CALC_CTPQ = 5

_MODEL_NAMES = {HUBBARD: "Hubbard", SPIN: "Spin", HUBBARD_GC: "HubbardGC", SPIN_GC: "SpinGC"}


@dataclass
class ModelSpec:
    """An expert-mode model: the raw entries as they appear in the def files."""

    name: str
    model: int
    nsite: int
    two_sz: Optional[int] = None  # canonical spin / Hubbard
    ncond: Optional[int] = None   # Hubbard electrons
    exchange: List[Tuple[int, int, float]] = field(default_factory=list)
    ising: List[Tuple[int, int, float]] = field(default_factory=list)
    pairlift: List[Tuple[int, int, float]] = field(default_factory=list)
    trans: List[Tuple[int, int, int, int, float, float]] = field(default_factory=list)
    coulomb_intra: List[Tuple[int, float]] = field(default_factory=list)
    coulomb_inter: List[Tuple[int, int, float]] = field(default_factory=list)
    hund: List[Tuple[int, int, float]] = field(default_factory=list)
    # Hubbard two-body terms: InterAll (i, s1, j, s2, k, s3, l, s4, re, im) for
    # V c+_{i s1} c_{j s2} c+_{k s3} c_{l s4}; PairHop (i, j, J); for Hubbard,
    # ``exchange`` holds the Hubbard Exchange (i, j, J)
    interall: List[Tuple[int, int, int, int, int, int, int, int, float, float]] = field(default_factory=list)
    pairhop: List[Tuple[int, int, float]] = field(default_factory=list)
    # Green's functions to measure: OneBodyG (i, s1, j, s2) for <c+_{i s1} c_{j s2}>,
    # TwoBodyG (i, s1, j, s2, k, s3, l, s4) for <c+_{i s1} c_{j s2} c+_{k s3} c_{l s4}>
    green1: List[Tuple[int, int, int, int]] = field(default_factory=list)
    green2: List[Tuple[int, int, int, int, int, int, int, int]] = field(default_factory=list)
    # SpinGC ThreeBodyG / FourBodyG / SixBodyG: 3, 4 or 6 pairs (i, s1, j, s2) each
    green3: List[Tuple[int, ...]] = field(default_factory=list)
    green4: List[Tuple[int, ...]] = field(default_factory=list)
    green6: List[Tuple[int, ...]] = field(default_factory=list)
    calc_type: int = CALC_LANCZOS
    params: Dict[str, float] = field(default_factory=dict)

    # ---- sector ---------------------------------------------------------
    @property
    def nup(self) -> int:
        """Number of 1-bits (spin) or up electrons (Hubbard), readdef.c:843-905."""
        if self.model == SPIN:
            return (self.nsite + self.two_sz) // 2
        if self.model == HUBBARD:
            return (self.ncond + self.two_sz) // 2
        return 0

    @property
    def ndown(self) -> int:
        if self.model == SPIN:
            return (self.nsite - self.two_sz) // 2
        if self.model == HUBBARD:
            return (self.ncond - self.two_sz) // 2
        return 0

    @property
    def dim(self) -> int:
        """idim_max, reference check.c:91-177."""
        if self.model == SPIN:
            return math.comb(self.nsite, self.nup)
        if self.model == HUBBARD:
            return math.comb(self.nsite, self.nup) * math.comb(self.nsite, self.ndown)
        if self.model == SPIN_GC:
            return 1 << self.nsite
        if self.model == HUBBARD_GC:
            return 1 << (2 * self.nsite)
        raise ValueError("unsupported model")

    @property
    def nbits(self) -> int:
        return 2 * self.nsite if self.model in (HUBBARD, HUBBARD_GC) else self.nsite


# ---------------------------------------------------------------------------
# lattices: bond lists (i, j) with i != j, sites 0-based
# ---------------------------------------------------------------------------

def chain_bonds(L: int, dist: int = 1) -> List[Tuple[int, int]]:
    """Periodic chain, bonds i -- i+dist."""
    return [(i, (i + dist) % L) for i in range(L)]


def square_site(x: int, y: int, Lx: int, Ly: int) -> int:
    return (x % Lx) + Lx * (y % Ly)


def square_bonds(Lx: int, Ly: int) -> Tuple[List[Tuple[int, int]], List[Tuple[int, int]]]:
    """Periodic Lx x Ly square lattice: (nearest, next-nearest) bond lists.

    Every site emits +x, +y (J1) and +x+y, +x-y (J2) bonds, so a short
    periodic direction repeats a pair; the repeated entry is kept, as a
    per-site bond generator on a small torus produces it.
    """
    nn, nnn = [], []
    for y in range(Ly):
        for x in range(Lx):
            s = square_site(x, y, Lx, Ly)
            nn.append((s, square_site(x + 1, y, Lx, Ly)))
            nn.append((s, square_site(x, y + 1, Lx, Ly)))
            nnn.append((s, square_site(x + 1, y + 1, Lx, Ly)))
            nnn.append((s, square_site(x + 1, y - 1, Lx, Ly)))
    return nn, nnn


def ladder_bonds(L: int) -> Tuple[List[Tuple[int, int]], List[Tuple[int, int]]]:
    """Periodic 2-leg ladder, site = 2*x + leg: (legs+rungs, diagonals)."""
    nn, nnn = [], []
    for x in range(L):
        a0, a1 = 2 * x, 2 * x + 1
        b0, b1 = 2 * ((x + 1) % L), 2 * ((x + 1) % L) + 1
        nn += [(a0, b0), (a1, b1), (a0, a1)]
        nnn += [(a0, b1), (a1, b0)]
    return nn, nnn


def triangular_bonds(Lx: int, Ly: int) -> List[Tuple[int, int]]:
    """Periodic triangular lattice: every site emits a1=(1,0), a2=(0,1), a1-a2=(1,-1).

    For Ly=2 the a2 bond of row 0 and of row 1 join the same pair, so that
    pair appears twice (78 entries, 65 distinct pairs for 13x2).
    """
    b = []
    for y in range(Ly):
        for x in range(Lx):
            s = square_site(x, y, Lx, Ly)
            b.append((s, square_site(x + 1, y, Lx, Ly)))
            b.append((s, square_site(x, y + 1, Lx, Ly)))
            b.append((s, square_site(x + 1, y - 1, Lx, Ly)))
    return b


# ---------------------------------------------------------------------------
# models
# ---------------------------------------------------------------------------

def _default_params(spec: ModelSpec) -> Dict[str, float]:
    # documented defaults (doc/en/source/filespecification/standardmode_en/
    # Parameters_for_the_numerical_condition.rst)
    return {"Lanczos_max": 2000, "initial_iv": -1, "exct": 1, "LanczosEps": 14,
            "LanczosTarget": 2, "LargeValue": 12, "NumAve": 5, "ExpecInterval": 20}


def spin_model(name: str, nsite: int, bonds_J: Sequence[Tuple[int, int, float]],
               two_sz: Optional[int] = 0, jz_scale: float = 1.0,
               jxy_scale: float = 1.0, hx: float = 0.0) -> ModelSpec:
    """Spin-1/2 model from (i, j, J) bonds.

    Each bond adds Exchange J*jxy_scale/2 and Ising J*jz_scale.  ``two_sz``
    None selects the grand-canonical SpinGC model.
    """
    model = SPIN if two_sz is not None else SPIN_GC
    spec = ModelSpec(name=name, model=model, nsite=nsite, two_sz=two_sz)
    for i, j, J in bonds_J:
        spec.exchange.append((i, j, 0.5 * J * jxy_scale))
        spec.ising.append((i, j, J * jz_scale))
    if hx != 0.0:
        if model != SPIN_GC:
            raise ValueError("a transverse field breaks Sz: use the SpinGC model")
        for i in range(nsite):
            spec.trans.append((i, 0, i, 1, 0.5 * hx, 0.0))
            spec.trans.append((i, 1, i, 0, 0.5 * hx, 0.0))
    spec.params = _default_params(spec)
    return spec


def hubbard_model(name: str, nsite: int, bonds_t: Sequence[Tuple[int, int, float]],
                  U: float, ncond: Optional[int], two_sz: int = 0) -> ModelSpec:
    """Hubbard model from (i, j, t) bonds; ``ncond`` None selects the
    grand-canonical HubbardGC model (every electron number and Sz)."""
    if ncond is None:
        spec = ModelSpec(name=name, model=HUBBARD_GC, nsite=nsite)
    else:
        spec = ModelSpec(name=name, model=HUBBARD, nsite=nsite, two_sz=two_sz, ncond=ncond)
    for i, j, t in bonds_t:
        for s in (0, 1):
            spec.trans.append((i, s, j, s, t, 0.0))
            spec.trans.append((j, s, i, s, t, 0.0))
    for i in range(nsite):
        spec.coulomb_intra.append((i, U))
    spec.params = _default_params(spec)
    return spec


def heisenberg_chain(L: int, J: float = 1.0, J2: float = 0.0, two_sz: int = 0) -> ModelSpec:
    bonds = [(i, j, J) for i, j in chain_bonds(L)]
    if J2 != 0.0:
        bonds += [(i, j, J2) for i, j in chain_bonds(L, 2)]
    return spin_model(f"chain{L}", L, bonds, two_sz=two_sz)


def j1j2_square(Lx: int, Ly: int, J1: float = 1.0, J2: float = 0.5, two_sz: int = 0) -> ModelSpec:
    nn, nnn = square_bonds(Lx, Ly)
    bonds = [(i, j, J1) for i, j in nn] + [(i, j, J2) for i, j in nnn]
    return spin_model(f"j1j2_{Lx}x{Ly}", Lx * Ly, bonds, two_sz=two_sz)


def j1j2_ladder(L: int, J1: float = 1.0, J2: float = 0.0, two_sz: int = 0) -> ModelSpec:
    nn, nnn = ladder_bonds(L)
    bonds = [(i, j, J1) for i, j in nn] + [(i, j, J2) for i, j in nnn]
    return spin_model(f"ladder{2 * L}", 2 * L, bonds, two_sz=two_sz)


def hubbard_square(Lx: int, Ly: int, t: float = 1.0, U: float = 4.0,
                   ncond: Optional[int] = None, two_sz: int = 0) -> ModelSpec:
    nn, _ = square_bonds(Lx, Ly)
    n = Lx * Ly
    return hubbard_model(f"hubbard_{Lx}x{Ly}", n, [(i, j, t) for i, j in nn], U,
                         n if ncond is None else ncond, two_sz)


def hubbard_chain(L: int, t: float = 1.0, U: float = 4.0, ncond: Optional[int] = None,
                  two_sz: Optional[int] = None) -> ModelSpec:
    ncond = L if ncond is None else ncond
    if two_sz is None:
        two_sz = ncond % 2
    return hubbard_model(f"hubbard_chain{L}", L, [(i, j, t) for i, j in chain_bonds(L)],
                         U, ncond, two_sz)


def hubbard_chain_gc(L: int, t: float = 1.0, U: float = 4.0, mu: float = 0.0,
                     flip: complex = 0.0) -> ModelSpec:
    """HubbardGC ring: hopping t, on-site U, chemical potential mu (a diagonal
    transfer -mu n), and an optional spin-flip hopping c+_{i up} c_{i+1 down}
    with amplitude ``flip`` (plus its Hermitian conjugate)."""
    spec = hubbard_model(f"hubbard_gc_chain{L}", L, [(i, j, t) for i, j in chain_bonds(L)], U, None)
    for i in range(L):
        for sp in (0, 1):
            if mu != 0.0:
                spec.trans.append((i, sp, i, sp, mu, 0.0))
    if flip != 0.0:
        for i, j in chain_bonds(L):
            spec.trans.append((i, 0, j, 1, flip.real, flip.imag))
            spec.trans.append((j, 1, i, 0, flip.real, -flip.imag))
    return spec


def triangular_xxz_gc(Lx: int = 13, Ly: int = 2, Jz: float = 1.0, Jxy: float = 0.8,
                      hx: float = 0.1) -> ModelSpec:
    bonds = [(i, j, 1.0) for i, j in triangular_bonds(Lx, Ly)]
    return spin_model(f"tri_xxz_{Lx}x{Ly}", Lx * Ly, bonds, two_sz=None,
                      jz_scale=Jz, jxy_scale=Jxy, hx=hx)


def baseline_config(n: int) -> ModelSpec:
    """The four fixed workloads: 1 chain16, 2 Hubbard 4x4, 3 J1-J2 8x4, 4 triangular 13x2 GC."""
    if n == 1:
        s = heisenberg_chain(16)
        s.params.update({"LanczosEps": 12})
        return s
    if n == 2:
        return hubbard_square(4, 4, t=1.0, U=4.0, ncond=16, two_sz=0)
    if n == 3:
        s = j1j2_square(8, 4, 1.0, 0.5)
        s.params.update({"exct": 2, "Lanczos_max": 200})
        return s
    if n == 4:
        s = triangular_xxz_gc(13, 2, 1.0, 0.8, 0.1)
        s.calc_type = CALC_TPQ
        return s
    raise ValueError(n)


# ---------------------------------------------------------------------------
# served-stream requests (workload 5)
# ---------------------------------------------------------------------------

SQUARE_SHAPES = {12: (4, 3), 15: (5, 3), 16: (4, 4), 18: (6, 3), 20: (5, 4), 21: (7, 3), 24: (6, 4)}


def stream_request(rng: np.random.Generator) -> ModelSpec:
    """One request of the served stream.

    Shape rules: chain N in 12..24 (J1-J2 chain, 2Sz = N mod 2); 2-leg ladder
    N even in 12..24 (legs+rungs J1, diagonals J2); square J1-J2 N drawn from
    the sizes with an Lx x Ly (Lx, Ly >= 3) torus; Hubbard chain N in 8..12,
    half filling, 2Sz = N mod 2.
    """
    kind = int(rng.integers(0, 4))
    j2 = float(rng.uniform(0.0, 1.0))
    if kind == 0:
        n = int(rng.integers(12, 25))
        return heisenberg_chain(n, 1.0, j2, two_sz=n % 2)
    if kind == 1:
        n = 2 * int(rng.integers(6, 13))
        return j1j2_ladder(n // 2, 1.0, j2)
    if kind == 2:
        n = list(SQUARE_SHAPES)[int(rng.integers(0, len(SQUARE_SHAPES)))]
        lx, ly = SQUARE_SHAPES[n]
        return j1j2_square(lx, ly, 1.0, j2)
    n = int(rng.integers(8, 13))
    return hubbard_chain(n, 1.0, float(rng.uniform(0.0, 8.0)))


# ---------------------------------------------------------------------------
# expert-mode writer
# ---------------------------------------------------------------------------

def _f(x: float) -> str:
    return repr(float(x))  # shortest round-trip repr: sscanf("%lf") reads it back exactly


def _header(tag: str, n: int, title: str) -> str:
    bar = "=" * 30
    return f"{bar}\n{tag} {n}\n{bar}\n{title}\n{bar}\n"


def write_expert(spec: ModelSpec, directory: str, calc_type: Optional[int] = None) -> str:
    """Write namelist.def and its files into ``directory``; return namelist path."""
    os.makedirs(directory, exist_ok=True)
    ct = spec.calc_type if calc_type is None else calc_type
    files = {}

    def put(kw, fname, text):
        with open(os.path.join(directory, fname), "w") as fh:
            fh.write(text)
        files[kw] = fname

    put("CalcMod", "calcmod.def",
        f"CalcType {ct}\nCalcModel {spec.model}\nReStart 0\nCalcSpec 0\n"
        f"CalcEigenVec 0\nInitialVecType 0\nInputEigenVec 0\nOutputEigenVec 0\n")
    lines = ["--------------------", "Model_Parameters   0", "--------------------",
             "HPhi_Cal_Parameters", "--------------------", "CDataFileHead  zvo",
             "CParaFileHead  zqp", "--------------------", f"Nsite {spec.nsite}"]
    if spec.model == SPIN:
        lines.append(f"2Sz {spec.two_sz}")
    elif spec.model == HUBBARD:
        lines.append(f"Ncond {spec.ncond}")
        lines.append(f"2Sz {spec.two_sz}")
    for k in ("Lanczos_max", "initial_iv", "exct", "LanczosEps", "LanczosTarget",
              "LargeValue", "NumAve", "ExpecInterval"):
        lines.append(f"{k} {_f(spec.params[k]) if k == 'LargeValue' else int(spec.params[k])}")
    if "ExpandCoef" in spec.params:  # canonical TPQ's Taylor order (0: the default 10)
        lines.append(f"ExpandCoef {int(spec.params['ExpandCoef'])}")
    put("ModPara", "modpara.def", "\n".join(lines) + "\n")

    loc = 1 if spec.model in (SPIN, SPIN_GC) else 0
    txt = _header("NlocalSpin", spec.nsite if loc else 0, "========i_0LocSpn_1IteElc ======")
    txt += "".join(f"{i} {loc}\n" for i in range(spec.nsite))
    put("LocSpin", "locspn.def", txt)

    if spec.trans:
        txt = _header("NTransfer", len(spec.trans), "========i_j_s_tijs======")
        txt += "".join(f"{i} {s1} {j} {s2} {_f(re)} {_f(im)}\n" for i, s1, j, s2, re, im in spec.trans)
        put("Trans", "trans.def", txt)
    if spec.coulomb_intra:
        txt = _header("NCoulombIntra", len(spec.coulomb_intra), "================== CoulombIntra ================")
        txt += "".join(f"{i} {_f(u)}\n" for i, u in spec.coulomb_intra)
        put("CoulombIntra", "coulombintra.def", txt)
    if spec.interall:
        txt = _header("NInterAll", len(spec.interall), "========zInterAll======")
        txt += "".join(" ".join(str(int(v)) for v in e[:8]) + f" {_f(e[8])} {_f(e[9])}\n" for e in spec.interall)
        put("InterAll", "interall.def", txt)
    for kw, fname, lst in (("CoulombInter", "coulombinter.def", spec.coulomb_inter),
                           ("Hund", "hund.def", spec.hund),
                           ("Exchange", "exchange.def", spec.exchange),
                           ("Ising", "ising.def", spec.ising),
                           ("PairLift", "pairlift.def", spec.pairlift),
                           ("PairHop", "pairhop.def", spec.pairhop)):
        if lst:
            txt = _header(f"N{kw}", len(lst), f"========{kw}======")
            txt += "".join(f"{i} {j} {_f(v)}\n" for i, j, v in lst)
            put(kw, fname, txt)
    if spec.green1:
        txt = _header("NCisAjs", len(spec.green1), "======== Green functions ======")
        txt += "".join(" ".join(str(int(v)) for v in e) + "\n" for e in spec.green1)
        put("OneBodyG", "greenone.def", txt)
    if spec.green2:
        txt = _header("NCisAjsCktAltDC", len(spec.green2), "======== Green functions for Sq AND Nq ======")
        txt += "".join(" ".join(str(int(v)) for v in e) + "\n" for e in spec.green2)
        put("TwoBodyG", "greentwo.def", txt)
    for kw, fname, lst in (("ThreeBodyG", "greenthree.def", spec.green3), ("FourBodyG", "greenfour.def", spec.green4),
                           ("SixBodyG", "greensix.def", spec.green6)):
        if lst:
            txt = _header(f"N{kw}", len(lst), f"======== {kw} ======")
            txt += "".join(" ".join(str(int(v)) for v in e) + "\n" for e in lst)
            put(kw, fname, txt)
    nl = os.path.join(directory, "namelist.def")
    with open(nl, "w") as fh:
        for kw, fname in files.items():
            fh.write(f"{kw} {fname}\n")
    return nl


# ---------------------------------------------------------------------------
# readdef post-processing, restated (test-side mirror of the C reader)
# ---------------------------------------------------------------------------

def processed_terms(spec: ModelSpec) -> Dict[str, np.ndarray]:
    """Term tables in the order the reference holds them after ReadDefFileIdxPara.

    Returns int32/float64 arrays: ``exchange_ij`` [n,2], ``exchange_J`` [n];
    ``trans_idx`` [n,4] and ``trans_t`` [n,2] (EDGeneralTransfer in Hermitian
    pairs); ``chemi_site``, ``chemi_spin``, ``chemi_mu`` (EDChemi);
    ``intra_site``, ``intra_U``; ``inter_ij``, ``inter_V`` (CoulombInter, Ising
    part appended); ``hund_ij``, ``hund_J`` (Hund, Ising part appended);
    ``pairlift_ij``, ``pairlift_J``.
    """
    # merge duplicate transfers (readdef.c:1130-1150)
    merged: List[List] = []
    for i, s1, j, s2, re, im in spec.trans:
        for m in merged:
            if m[0] == i and m[1] == s1 and m[2] == j and m[3] == s2:
                m[4] += complex(re, im)
                break
        else:
            merged.append([i, s1, j, s2, complex(re, im)])
    # Hermitian pairing (readdef.c:2136-2234)
    ed_idx, ed_t, chemi = [], [], []
    for a, (i, s1, j, s2, t) in enumerate(merged):
        found = False
        for b, (i2, t1, j2, t2, tb) in enumerate(merged):
            if i == j2 and j == i2 and s1 == t2 and s2 == t1:
                if abs(t - tb.conjugate()) > 1e-12:
                    raise ValueError("non-Hermitian transfer")
                if not found:
                    if a <= b:
                        if i != j or s1 != s2:
                            ed_idx += [(i, s1, j, s2), (i2, t1, j2, t2)]
                            ed_t += [t, tb]
                        else:
                            chemi.append((i, s1, t.real))
                    found = True
        if not found:
            raise ValueError("transfer without Hermitian partner")
    hund = list(spec.hund) + [(i, j, -J / 2.0) for i, j, J in spec.ising]
    inter = list(spec.coulomb_inter) + [(i, j, -J / 4.0) for i, j, J in spec.ising]
    # PairHop: each entry and its reverse (readdef.c:1257-1267)
    if spec.pairhop and spec.model not in (HUBBARD, HUBBARD_GC):
        raise ValueError("PairHop needs a Hubbard model")
    pairhop = []
    for i, j, J in spec.pairhop:
        pairhop += [(i, j, J), (j, i, J)]
    # InterAll: duplicates summed in first-occurrence order (InputInterAllInfo,
    # readdef.c:3054-3090), then split (GetDiagonalInterAll, readdef.c:2415-2530):
    # n_i n_k -> diagonal; c+_i c_j c+_j c_i -> diagonal -V n_i n_j plus a
    # chemical-potential entry; spin-exchange order rewritten with -V
    spin_model = spec.model not in (HUBBARD, HUBBARD_GC)
    ia_merged: List[List] = []
    for e in spec.interall:
        # CheckInterAllCondition (readdef.c:2996-3033)
        if any(not 0 <= e[u] < spec.nsite for u in (0, 2, 4, 6)):
            raise ValueError("InterAll site index out of range")
        if spin_model and (e[0] != e[2] or e[4] != e[6]):
            raise ValueError("spin InterAll needs i == j and k == l")
        for m in ia_merged:
            if m[:8] == list(e[:8]):
                m[8] += complex(e[8], e[9])
                break
        else:
            ia_merged.append(list(e[:8]) + [complex(e[8], e[9])])
    ia_off, ia_offV, ia_diag, ia_diagV = [], [], [], []
    for i, s1, j, s2, k, s3, l, s4, V in ia_merged:
        if i == j and k == l and s1 == s2 and s3 == s4:
            ia_diag.append((i, s1, k, s3))
            ia_diagV.append(V.real)
        elif i == l and j == k and s1 == s4 and s2 == s3:
            ia_diag.append((i, s1, j, s2))
            ia_diagV.append(-V.real)
            chemi.append((i, s1, -V.real))
        elif spin_model or (s1 == s2 and s3 == s4):
            ia_off.append((i, s1, j, s2, k, s3, l, s4))
            ia_offV.append(V)
        elif s1 == s4 and s2 == s3:
            ia_off.append((i, s1, l, s1, k, s2, j, s2))
            ia_offV.append(-V)
        elif spec.model == HUBBARD_GC:
            ia_off.append((i, s1, j, s2, k, s3, l, s4))
            ia_offV.append(V)
        else:
            raise ValueError("InterAll term does not conserve Sz")
    # Hermitian pairing: entry a is followed by its reverse with the conjugate
    # value (CheckInterAllHermite, readdef.c:2257-2360)
    # (spin: the partner may also list the two site operators the other way
    # round; the pair is then entry a and its reverse, readdef.c:2321-2352)
    paired, pairedV = [], []
    for a, qa in enumerate(ia_off):
        rev = (qa[6], qa[7], qa[4], qa[5], qa[2], qa[3], qa[0], qa[1])
        swp = (qa[2], qa[3], qa[0], qa[1], qa[6], qa[7], qa[4], qa[5])
        hit = next((b for b, qb in enumerate(ia_off)
                    if tuple(qb) == rev and abs(ia_offV[a] - ia_offV[b].conjugate()) < 1e-12), None)
        if hit is not None:
            if a <= hit:
                paired += [qa, ia_off[hit]]
                pairedV += [ia_offV[a], ia_offV[hit]]
            continue
        if spin_model and qa[0] != qa[4] and qa[1] != qa[5] and qa[2] != qa[6] and qa[3] != qa[7]:
            hit = next((b for b, qb in enumerate(ia_off)
                        if tuple(qb) == swp and abs(ia_offV[a] - ia_offV[b].conjugate()) < 1e-12), None)
            if hit is not None:
                if a <= hit:
                    paired += [tuple(qa), rev]
                    pairedV += [ia_offV[a], ia_offV[hit]]
                continue
        raise ValueError("non-Hermitian InterAll")
    if len(paired) != len(ia_off):
        raise ValueError("non-Hermitian InterAll")
    ia_off, ia_offV = paired, pairedV

    def arr2(lst, k):
        return np.array([x[:k] for x in lst], dtype=np.int32).reshape(-1, k)

    def arrv(lst, k):
        return np.array([x[k] for x in lst], dtype=np.float64)

    return {
        "exchange_ij": arr2(spec.exchange, 2), "exchange_J": arrv(spec.exchange, 2),
        "pairlift_ij": arr2(spec.pairlift, 2), "pairlift_J": arrv(spec.pairlift, 2),
        "trans_idx": np.array(ed_idx, dtype=np.int32).reshape(-1, 4),
        "trans_t": np.array([(t.real, t.imag) for t in ed_t], dtype=np.float64).reshape(-1, 2),
        "chemi_site": np.array([c[0] for c in chemi], dtype=np.int32),
        "chemi_spin": np.array([c[1] for c in chemi], dtype=np.int32),
        "chemi_mu": np.array([c[2] for c in chemi], dtype=np.float64),
        "intra_site": np.array([c[0] for c in spec.coulomb_intra], dtype=np.int32),
        "intra_U": np.array([c[1] for c in spec.coulomb_intra], dtype=np.float64),
        "inter_ij": arr2(inter, 2), "inter_V": arrv(inter, 2),
        "hund_ij": arr2(hund, 2), "hund_J": arrv(hund, 2),
        "interall_idx": np.array(ia_off, dtype=np.int32).reshape(-1, 8),
        "interall_V": np.array([(v.real, v.imag) for v in ia_offV], dtype=np.float64).reshape(-1, 2),
        "interall_diag_idx": np.array(ia_diag, dtype=np.int32).reshape(-1, 4),
        "interall_diag_V": np.array(ia_diagV, dtype=np.float64),
        "pairhop_ij": arr2(pairhop, 2), "pairhop_J": arrv(pairhop, 2),
    }


def processed_green(spec: ModelSpec) -> Tuple[List[Tuple[int, ...]], List[Tuple[int, ...]]]:
    """OneBodyG / TwoBodyG entries as the reference's reader keeps them
    (readdef.c:1441-1515): a canonical Spin one-body entry with i != j is
    dropped with a warning; a spin two-body entry must read (i s1 i s2 k s3 k s4)."""
    g1 = [tuple(e) for e in spec.green1 if not (spec.model == SPIN and e[0] != e[2])]
    g2 = [tuple(e) for e in spec.green2]
    if spec.model in (SPIN, SPIN_GC) and any(e[0] != e[2] or e[4] != e[6] for e in g2):
        raise ValueError("spin TwoBodyG entries need i == j and k == l")
    return g1, g2
