"""ctypes binding of the CPU oracle — TEST INFRASTRUCTURE ONLY.

Only tests, the smoke check and the benchmark's CPU-baseline leg use this
module, as the checker; the product never imports it.  See hphi_oracle.h for
what the oracle restates and how each function is pinned.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict, Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

_i32p = C.POINTER(C.c_int)
_f64p = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)


class OrModel(C.Structure):
    _fields_ = [("model", C.c_int), ("nsite", C.c_int), ("nup", C.c_int), ("ndown", C.c_int),
                ("n_ex", C.c_int), ("ex_ij", _i32p), ("ex_J", _f64p),
                ("n_pl", C.c_int), ("pl_ij", _i32p), ("pl_J", _f64p),
                ("n_tr", C.c_int), ("tr_idx", _i32p), ("tr_t", _f64p),
                ("n_ch", C.c_int), ("ch_site", _i32p), ("ch_spin", _i32p), ("ch_mu", _f64p),
                ("n_intra", C.c_int), ("intra_site", _i32p), ("intra_U", _f64p),
                ("n_inter", C.c_int), ("inter_ij", _i32p), ("inter_V", _f64p),
                ("n_hund", C.c_int), ("hund_ij", _i32p), ("hund_J", _f64p)]


def build(quiet: bool = True) -> None:
    """Compile liboracle.so, and the reference harness (plain and with the drop-in
    binding) when the reference tree is present."""
    out = subprocess.DEVNULL if quiet else None
    subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True, stdout=out)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-C", HERE, "-j8", "ref"], check=True, stdout=out)
        # the reference's callers with the drop-in binding (needs libhphi_amd.so)
        if os.path.exists(os.path.join(os.path.dirname(HERE), "hphi_amd", "libhphi_amd.so")):
            subprocess.run(["make", "-C", HERE, "dropin"], check=True, stdout=out)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_dims.argtypes = [C.POINTER(OrModel)] + [_u64p] * 6
        L.or_sz.argtypes = [C.POINTER(OrModel), _u64p, _u64p, _u64p]
        L.or_diagonalcalc.argtypes = [C.POINTER(OrModel), _u64p, C.c_uint64, _f64p]
        L.or_mltply.argtypes = [C.POINTER(OrModel), _u64p, _u64p, _u64p, _f64p, C.c_uint64,
                                _f64p, _f64p, _f64p, C.c_uint64, C.c_uint64]
        L.or_fill_vector.argtypes = [C.c_uint64, C.c_uint64, _f64p]
        L.or_num_threads.restype = C.c_int
        L.or_count_nnz.argtypes = [C.POINTER(OrModel), _u64p, C.c_uint64, C.c_uint64]
        L.or_count_nnz.restype = C.c_uint64
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None and a.size else None


class Oracle:
    """Oracle state for one model: tables, diagonal, H·v (all 1-indexed like the reference)."""

    def __init__(self, spec, terms: Dict[str, np.ndarray]):
        self.spec = spec
        self._keep = {k: np.ascontiguousarray(v) for k, v in terms.items()}
        t = self._keep
        m = OrModel()
        m.model, m.nsite, m.nup, m.ndown = spec.model, spec.nsite, spec.nup, spec.ndown
        m.n_ex, m.ex_ij, m.ex_J = len(t["exchange_J"]), _p(t["exchange_ij"], _i32p), _p(t["exchange_J"], _f64p)
        m.n_pl, m.pl_ij, m.pl_J = len(t["pairlift_J"]), _p(t["pairlift_ij"], _i32p), _p(t["pairlift_J"], _f64p)
        m.n_tr, m.tr_idx, m.tr_t = len(t["trans_t"]), _p(t["trans_idx"], _i32p), _p(t["trans_t"], _f64p)
        m.n_ch, m.ch_site, m.ch_spin, m.ch_mu = (len(t["chemi_mu"]), _p(t["chemi_site"], _i32p),
                                                _p(t["chemi_spin"], _i32p), _p(t["chemi_mu"], _f64p))
        m.n_intra, m.intra_site, m.intra_U = len(t["intra_U"]), _p(t["intra_site"], _i32p), _p(t["intra_U"], _f64p)
        m.n_inter, m.inter_ij, m.inter_V = len(t["inter_V"]), _p(t["inter_ij"], _i32p), _p(t["inter_V"], _f64p)
        m.n_hund, m.hund_ij, m.hund_J = len(t["hund_J"]), _p(t["hund_ij"], _i32p), _p(t["hund_J"], _f64p)
        self.m = m
        d = [C.c_uint64() for _ in range(6)]
        if lib().or_dims(C.byref(m), *[C.byref(x) for x in d]) != 0:
            raise ValueError("unsupported model")
        self.dim, self.size21, self.size22, self.irght, self.ilft, self.ihfbit = [x.value for x in d]
        self.list_1 = self.list_2_1 = self.list_2_2 = None
        self.diag = None

    def sz(self):
        canon = self.spec.model in (0, 1)
        self.list_1 = np.zeros(self.dim + 1 if canon else 1, np.uint64)
        self.list_2_1 = np.zeros(max(self.size21, 1), np.uint64)
        self.list_2_2 = np.zeros(max(self.size22, 1), np.uint64)
        rc = lib().or_sz(C.byref(self.m), _p(self.list_1, _u64p), _p(self.list_2_1, _u64p),
                         _p(self.list_2_2, _u64p))
        if rc != 0:
            raise RuntimeError("or_sz failed")
        return self.list_1, self.list_2_1, self.list_2_2

    def diagonalcalc(self):
        if self.list_1 is None:
            self.sz()
        self.diag = np.zeros(self.dim + 1, np.float64)
        lib().or_diagonalcalc(C.byref(self.m), _p(self.list_1, _u64p), self.dim, _p(self.diag, _f64p))
        return self.diag

    def prepare(self):
        self.sz()
        self.diagonalcalc()
        return self

    def mltply(self, v0: np.ndarray, v1: np.ndarray, jlo: int = 1, jhi: Optional[int] = None) -> complex:
        """v0[1..dim] += H v1[1..dim] (complex128, 1-indexed, length dim+1); returns prdct."""
        assert v0.dtype == np.complex128 and v1.dtype == np.complex128
        assert v0.shape == (self.dim + 1,) and v1.shape == (self.dim + 1,)
        pr = np.zeros(2, np.float64)
        rc = lib().or_mltply(C.byref(self.m), _p(self.list_1, _u64p), _p(self.list_2_1, _u64p),
                             _p(self.list_2_2, _u64p), _p(self.diag, _f64p), self.dim,
                             v0.ctypes.data_as(_f64p), v1.ctypes.data_as(_f64p), _p(pr, _f64p),
                             jlo, self.dim if jhi is None else jhi)
        if rc != 0:
            raise RuntimeError("or_mltply failed")
        return complex(pr[0], pr[1])

    def hv(self, v: np.ndarray) -> Tuple[np.ndarray, complex]:
        """H v for a 0-based vector; returns (Hv, <v|H|v>)."""
        v1 = np.zeros(self.dim + 1, np.complex128)
        v1[1:] = v
        v0 = np.zeros_like(v1)
        pr = self.mltply(v0, v1)
        return v0[1:].copy(), pr


def fill_vector(seed: int, dim: int) -> np.ndarray:
    """Synthetic input vector (0-based) from the documented SplitMix64 generator."""
    v = np.zeros(dim + 1, np.complex128)
    lib().or_fill_vector(seed, dim, v.ctypes.data_as(_f64p))
    return v[1:].copy()


def splitmix_vector_np(seed: int, dim: int) -> np.ndarray:
    """Same generator in numpy (independent restatement, used to pin or_fill_vector)."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        c = np.arange(2 * dim, dtype=np.uint64)
        x = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)) ^ c
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    del M
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u = 2.0 * u - 1.0
    return u[0::2] + 1j * u[1::2]


def run_reference(def_dir: str, mode: str, out_dir: str, seed: int = 0, reps: int = 3,
                  threads: Optional[int] = None, timeout: int = 3600) -> str:
    """Run the compiled reference harness (oracle/_ref) — only where it was built."""
    if not os.path.exists(REF_HARNESS):
        raise FileNotFoundError(REF_HARNESS)
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    r = subprocess.run([REF_HARNESS, os.path.abspath(def_dir), mode, os.path.abspath(out_dir),
                        str(seed), str(reps)], env=env, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"ref_harness failed rc={r.returncode}: {r.stdout[-2000:]} {r.stderr[-2000:]}")
    return r.stdout
