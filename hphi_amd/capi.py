"""ctypes binding of libhphi_amd.so (the C ABI in include/hphi_amd.h).

The shared library is the product; this module only marshals arguments.  It
raises when the library is missing: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhphi_amd.so")

_i32p = C.POINTER(C.c_int)
_f64p = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p

# every symbol include/hphi_amd.h declares
EXPORTED = [
    "hphi_last_error", "hphi_engine_create", "hphi_engine_create_from_def", "hphi_engine_destroy",
    "hphi_engine_info", "hphi_engine_device_bytes", "hphi_sz", "hphi_diagonalcalc", "hphi_mltply",
    "hphi_mltply_dev", "hphi_mltply_block_dev", "hphi_dev_alloc", "hphi_dev_free", "hphi_memcpy_h2d",
    "hphi_memcpy_d2h", "hphi_dev_zero", "hphi_synchronize", "hphi_fill_random",
    "hphi_vec_lanczos_beta", "hphi_vec_lanczos_update", "hphi_vec_norm2", "hphi_vec_dot",
    "hphi_vec_axpby", "hphi_engine_sync", "hphi_time_mltply_dev", "hphi_lanczos",
    "hphi_engine_tune", "hphi_engine_get_tune", "hphi_autotune", "hphi_tpq", "hphi_diag_maxabs",
    "hphi_read_def", "hphi_model_desc_free", "hphi_tridiag_lowest",
    "hphi_lanczos_good_eigenvalues", "hphi_lanczos_converged",
    "hphi_server_create", "hphi_server_destroy", "hphi_server_submit", "hphi_server_tick",
    "hphi_server_results", "hphi_server_now", "hphi_server_stats", "hphi_lobpcg", "hphi_herm_eig",
    "hphi_lanczos_eigenvector", "hphi_vec_lanczos_update_acc", "hphi_expec_energy_flct", "hphi_expec_cisajs",
    "hphi_expec_cisajscktalt", "hphi_write_vec_file", "hphi_read_vec_file", "hphi_write_restart_vecs",
    "hphi_read_restart_vecs", "hphi_write_tmcomponents", "hphi_read_tmcomponents", "hphi_write_energy",
    "hphi_write_cisajs", "hphi_write_cisajscktalt", "hphi_write_tpq_files", "hphi_lanczos_restart",
    "hphi_ctpq", "hphi_write_ctpq_files", "hphi_mltply_dev_scaled", "hphi_vec_lanczos_resid",
    "hphi_mltply_block_dev_noalloc", "hphi_model_nnz", "hphi_tpq_from", "hphi_ctpq_from",
    "hphi_append_tpq_files", "hphi_lobpcg_ex", "hphi_write_energy_states", "hphi_lanczos_ritz_residual",
    "hphi_lanczos_residuals_small", "hphi_expec_spin_nbody", "hphi_write_nbody_green",
]


class ModelDesc(C.Structure):
    _fields_ = [("model", C.c_int), ("nsite", C.c_int), ("nup", C.c_int), ("ndown", C.c_int),
                ("n_exchange", C.c_int), ("exchange_ij", _i32p), ("exchange_J", _f64p),
                ("n_pairlift", C.c_int), ("pairlift_ij", _i32p), ("pairlift_J", _f64p),
                ("n_transfer", C.c_int), ("transfer_idx", _i32p), ("transfer_t", _f64p),
                ("n_chemi", C.c_int), ("chemi_site", _i32p), ("chemi_spin", _i32p), ("chemi_mu", _f64p),
                ("n_intra", C.c_int), ("intra_site", _i32p), ("intra_U", _f64p),
                ("n_inter", C.c_int), ("inter_ij", _i32p), ("inter_V", _f64p),
                ("n_hund", C.c_int), ("hund_ij", _i32p), ("hund_J", _f64p),
                ("n_interall", C.c_int), ("interall_idx", _i32p), ("interall_V", _f64p),
                ("n_interall_diag", C.c_int), ("interall_diag_idx", _i32p), ("interall_diag_V", _f64p),
                ("n_pairhop", C.c_int), ("pairhop_ij", _i32p), ("pairhop_J", _f64p)]


class ModPara(C.Structure):
    _fields_ = [("calc_type", C.c_int), ("lanczos_max", C.c_int), ("initial_iv", C.c_int),
                ("exct", C.c_int), ("lanczos_eps", C.c_int), ("lanczos_target", C.c_int),
                ("num_ave", C.c_int), ("expec_interval", C.c_int), ("large_value", C.c_double),
                ("restart", C.c_int), ("initial_vec_type", C.c_int), ("calc_eigenvec", C.c_int),
                ("output_eigenvec", C.c_int), ("input_eigenvec", C.c_int), ("expand_coef", C.c_int)]


class LanczosParams(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("n_eigen", C.c_int), ("eps", C.c_double),
                ("target", C.c_int), ("seed", C.c_uint64), ("check_every", C.c_int), ("mode", C.c_int)]


class Phys(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("energy", "var", "doublon", "doublon2", "num", "num2", "sz", "sz2")]

    def as_dict(self) -> Dict[str, float]:
        return {k: getattr(self, k) for k, _ in self._fields_}


class LanczosResult(C.Structure):
    _fields_ = [("eigenvalues", C.c_double * 16), ("n_found", C.c_int), ("iterations", C.c_int),
                ("converged", C.c_int), ("seconds", C.c_double), ("hv_seconds", C.c_double)]


class TpqParams(C.Structure):
    _fields_ = [("k", C.c_int), ("steps", C.c_int), ("init_kind", C.c_int), ("layout", C.c_int),
                ("use_graph", C.c_int), ("large_value", C.c_double), ("seed", C.c_uint64)]


class CtpqParams(C.Structure):
    _fields_ = [("k", C.c_int), ("steps", C.c_int), ("init_kind", C.c_int), ("expand_coef", C.c_int),
                ("large_value", C.c_double), ("seed", C.c_uint64)]


class TpqResult(C.Structure):
    _fields_ = [("seconds", C.c_double), ("large_value", C.c_double), ("steps", C.c_int)]


class LobpcgParams(C.Structure):
    _fields_ = [("k", C.c_int), ("max_iter", C.c_int), ("eps", C.c_double), ("precg", C.c_int),
                ("large_value", C.c_double), ("seed", C.c_uint64)]


class LobpcgResult(C.Structure):
    _fields_ = [("eigenvalues", C.c_double * 16), ("residuals", C.c_double * 16), ("max_residual", C.c_double),
                ("iterations", C.c_int), ("converged", C.c_int), ("seconds", C.c_double),
                ("hv_seconds", C.c_double)]


class ServerParams(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_jobs", C.c_int), ("n_eigen", C.c_int),
                ("eps", C.c_double), ("max_iter", C.c_int), ("check_every", C.c_int)]


class JobResult(C.Structure):
    _fields_ = [("id", C.c_int), ("n_found", C.c_int), ("iterations", C.c_int), ("converged", C.c_int),
                ("dim", C.c_uint64), ("eigenvalues", C.c_double * 16), ("t_arrival", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double)]


_lib = None


class HphiError(RuntimeError):
    pass


def lib():
    """Load libhphi_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HphiError(f"{LIB_PATH} is missing: build it with `python -m hphi_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.hphi_last_error.restype = C.c_char_p
        L.hphi_engine_create.argtypes = [C.POINTER(ModelDesc), C.POINTER(_vp)]
        L.hphi_engine_create_from_def.argtypes = [C.c_char_p, C.POINTER(_vp), C.POINTER(ModPara)]
        L.hphi_engine_destroy.argtypes = [_vp]
        L.hphi_engine_destroy.restype = None
        L.hphi_engine_info.argtypes = [_vp] + [_u64p] * 6
        L.hphi_engine_device_bytes.argtypes = [_vp]
        L.hphi_model_nnz.argtypes = [C.POINTER(ModelDesc), _u64p, _u64p]
        L.hphi_engine_device_bytes.restype = C.c_uint64
        L.hphi_sz.argtypes = [_vp, _u64p, _u64p, _u64p]
        L.hphi_diagonalcalc.argtypes = [_vp, _f64p]
        L.hphi_mltply.argtypes = [_vp, _vp, _vp, _f64p]
        L.hphi_mltply_dev.argtypes = [_vp, _vp, _vp, _f64p]
        L.hphi_mltply_dev_scaled.argtypes = [_vp, _vp, _vp, C.c_double, C.c_double, C.c_int, _f64p]
        L.hphi_mltply_block_dev.argtypes = [_vp, C.c_int, _vp, _vp, _f64p]
        L.hphi_mltply_block_dev_noalloc.argtypes = [_vp, C.c_int, _vp, _vp, _f64p]
        L.hphi_dev_alloc.argtypes = [C.c_size_t, C.POINTER(_vp)]
        L.hphi_dev_free.argtypes = [_vp]
        L.hphi_memcpy_h2d.argtypes = [_vp, _vp, C.c_size_t]
        L.hphi_memcpy_d2h.argtypes = [_vp, _vp, C.c_size_t]
        L.hphi_dev_zero.argtypes = [_vp, C.c_size_t]
        L.hphi_fill_random.argtypes = [_vp, C.c_uint64, C.c_int, C.c_uint64]
        L.hphi_vec_lanczos_beta.argtypes = [_vp, _vp, _vp, C.c_double, _f64p]
        L.hphi_vec_lanczos_resid.argtypes = [_vp, _vp, _vp, _vp, C.c_double, C.c_double, _f64p]
        L.hphi_vec_lanczos_update.argtypes = [_vp, _vp, _vp, C.c_double, C.c_double]
        L.hphi_vec_norm2.argtypes = [_vp, _vp, C.c_int, _f64p]
        L.hphi_vec_dot.argtypes = [_vp, _vp, _vp, C.c_int, _f64p]
        L.hphi_vec_axpby.argtypes = [_vp, _vp, _vp, C.c_int, _f64p]
        L.hphi_engine_sync.argtypes = [_vp]
        L.hphi_time_mltply_dev.argtypes = [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int, _f64p]
        L.hphi_lanczos.argtypes = [_vp, C.POINTER(LanczosParams), C.POINTER(LanczosResult)]
        L.hphi_engine_tune.argtypes = [_vp, C.c_char_p, C.c_int]
        L.hphi_engine_get_tune.argtypes = [_vp, C.c_char_p, C.POINTER(C.c_int)]
        L.hphi_autotune.argtypes = [_vp, C.c_int, _vp, _vp, C.c_int, C.c_char_p, C.c_size_t]
        L.hphi_tpq.argtypes = [_vp, C.POINTER(TpqParams), _f64p, C.POINTER(TpqResult)]
        L.hphi_diag_maxabs.argtypes = [_vp, _f64p]
        L.hphi_ctpq.argtypes = [_vp, C.POINTER(CtpqParams), _f64p, C.POINTER(TpqResult)]
        L.hphi_lobpcg.argtypes = [_vp, C.POINTER(LobpcgParams), _f64p, C.POINTER(LobpcgResult)]
        L.hphi_herm_eig.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.hphi_server_create.argtypes = [C.POINTER(ServerParams), C.POINTER(_vp)]
        L.hphi_server_destroy.argtypes = [_vp]
        L.hphi_server_destroy.restype = None
        L.hphi_server_submit.argtypes = [_vp, C.POINTER(ModelDesc), C.c_uint64, C.c_double, C.POINTER(C.c_int)]
        L.hphi_server_tick.argtypes = [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.hphi_server_results.argtypes = [_vp, C.POINTER(JobResult), C.c_int, C.POINTER(C.c_int)]
        L.hphi_server_now.argtypes = [_vp]
        L.hphi_server_now.restype = C.c_double
        L.hphi_server_stats.argtypes = [_vp, _f64p]
        L.hphi_lanczos_eigenvector.argtypes = [_vp, C.POINTER(LanczosParams), C.c_int, _vp,
                                               C.POINTER(LanczosResult)]
        L.hphi_vec_lanczos_update_acc.argtypes = [_vp, _vp, _vp, _vp, C.c_double, C.c_double, C.c_double]
        L.hphi_expec_energy_flct.argtypes = [_vp, _vp, C.POINTER(Phys)]
        L.hphi_expec_cisajs.argtypes = [_vp, _vp, C.c_int, C.POINTER(C.c_int), _f64p]
        L.hphi_expec_cisajscktalt.argtypes = [_vp, _vp, C.c_int, C.POINTER(C.c_int), _f64p]
        L.hphi_write_vec_file.argtypes = [C.c_char_p, C.c_int, C.c_uint64, _f64p]
        L.hphi_read_vec_file.argtypes = [C.c_char_p, C.c_uint64, _f64p, C.POINTER(C.c_int)]
        L.hphi_write_restart_vecs.argtypes = [C.c_char_p, C.c_uint64, C.c_uint64, _f64p, _f64p]
        L.hphi_read_restart_vecs.argtypes = [C.c_char_p, C.c_uint64, _f64p, _f64p, _u64p]
        L.hphi_write_tmcomponents.argtypes = [C.c_char_p, C.c_int, C.c_double, _f64p, _f64p]
        L.hphi_read_tmcomponents.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), _f64p, _f64p, _f64p]
        L.hphi_write_energy.argtypes = [C.c_char_p, C.POINTER(Phys)]
        L.hphi_write_cisajs.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), _f64p]
        L.hphi_write_cisajscktalt.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int), _f64p]
        L.hphi_write_tpq_files.argtypes = [C.c_char_p, C.c_int, C.c_int, _f64p, C.c_int]
        L.hphi_append_tpq_files.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _f64p, C.c_int, C.c_double]
        L.hphi_write_energy_states.argtypes = [C.c_char_p, C.c_int, C.POINTER(Phys)]
        L.hphi_tpq_from.argtypes = [_vp, C.POINTER(TpqParams), _vp, C.c_int, _f64p, C.POINTER(TpqResult), _vp]
        L.hphi_ctpq_from.argtypes = [_vp, C.POINTER(CtpqParams), _vp, C.c_int, _f64p, C.POINTER(TpqResult), _vp]
        L.hphi_expec_spin_nbody.argtypes = [_vp, _vp, C.c_int, C.c_int, _i32p, _f64p, _i32p]
        L.hphi_write_nbody_green.argtypes = [C.c_char_p, C.c_int, C.c_int, _i32p, _f64p, _i32p]
        L.hphi_lobpcg_ex.argtypes = [_vp, C.POINTER(LobpcgParams), _vp, _vp, _f64p, C.POINTER(LobpcgResult)]
        L.hphi_write_ctpq_files.argtypes = [C.c_char_p, C.c_int, C.c_int, _f64p, C.c_int]
        L.hphi_lanczos_restart.argtypes = [_vp, C.POINTER(LanczosParams), C.c_char_p, C.c_char_p, C.c_char_p,
                                           C.c_char_p, C.POINTER(LanczosResult)]
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise HphiError(lib().hphi_last_error().decode())


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None and a.size else None


class DeviceArray:
    """A raw device allocation (complex128 elements) owned by Python."""

    def __init__(self, n_complex: int):
        self.n = int(n_complex)
        self.nbytes = 16 * self.n
        p = _vp()
        check(lib().hphi_dev_alloc(max(self.nbytes, 16), C.byref(p)))
        self.ptr = p

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.size == self.n
        check(lib().hphi_memcpy_h2d(self.ptr, a.ctypes.data_as(_vp), self.nbytes))
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.n, np.complex128)
        check(lib().hphi_memcpy_d2h(out.ctypes.data_as(_vp), self.ptr, self.nbytes))
        return out

    def zero(self) -> "DeviceArray":
        check(lib().hphi_dev_zero(self.ptr, self.nbytes))
        return self

    def free(self) -> None:
        if self.ptr:
            lib().hphi_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Engine:
    """One Hamiltonian on the GPU: tables, diagonal and H·v through the C ABI."""

    def __init__(self, handle: _vp, modpara: Optional[ModPara] = None):
        self.h = handle
        self.modpara = modpara
        d = [C.c_uint64() for _ in range(6)]
        check(lib().hphi_engine_info(self.h, *[C.byref(x) for x in d]))
        self.dim, self.size21, self.size22, self.irght, self.ilft, self.ihfbit = [x.value for x in d]

    @classmethod
    def from_terms(cls, model: int, nsite: int, nup: int, ndown: int, t: Dict[str, np.ndarray]) -> "Engine":
        d, keep = cls.desc_from_terms(model, nsite, nup, ndown, t)
        h = _vp()
        check(lib().hphi_engine_create(C.byref(d), C.byref(h)))
        return cls(h)

    @staticmethod
    def desc_from_terms(model, nsite, nup, ndown, t):
        """(ModelDesc, keep-alive arrays) for the C ABI."""
        keep = {k: np.ascontiguousarray(v) for k, v in t.items()}
        d = ModelDesc()
        d.model, d.nsite, d.nup, d.ndown = model, nsite, nup, ndown
        d.n_exchange, d.exchange_ij, d.exchange_J = len(keep["exchange_J"]), _p(keep["exchange_ij"], _i32p), _p(keep["exchange_J"], _f64p)
        d.n_pairlift, d.pairlift_ij, d.pairlift_J = len(keep["pairlift_J"]), _p(keep["pairlift_ij"], _i32p), _p(keep["pairlift_J"], _f64p)
        d.n_transfer, d.transfer_idx, d.transfer_t = len(keep["trans_t"]), _p(keep["trans_idx"], _i32p), _p(keep["trans_t"], _f64p)
        d.n_chemi, d.chemi_site, d.chemi_spin, d.chemi_mu = (len(keep["chemi_mu"]), _p(keep["chemi_site"], _i32p),
                                                             _p(keep["chemi_spin"], _i32p), _p(keep["chemi_mu"], _f64p))
        d.n_intra, d.intra_site, d.intra_U = len(keep["intra_U"]), _p(keep["intra_site"], _i32p), _p(keep["intra_U"], _f64p)
        d.n_inter, d.inter_ij, d.inter_V = len(keep["inter_V"]), _p(keep["inter_ij"], _i32p), _p(keep["inter_V"], _f64p)
        d.n_hund, d.hund_ij, d.hund_J = len(keep["hund_J"]), _p(keep["hund_ij"], _i32p), _p(keep["hund_J"], _f64p)
        for k, w in (("interall_idx", 8), ("interall_diag_idx", 4), ("pairhop_ij", 2)):
            keep.setdefault(k, np.zeros((0, w), np.int32))
        for k, w in (("interall_V", 2),):
            keep.setdefault(k, np.zeros((0, w), np.float64))
        for k in ("interall_diag_V", "pairhop_J"):
            keep.setdefault(k, np.zeros(0, np.float64))
        d.n_interall, d.interall_idx, d.interall_V = (len(keep["interall_V"]), _p(keep["interall_idx"], _i32p),
                                                      _p(keep["interall_V"], _f64p))
        d.n_interall_diag, d.interall_diag_idx, d.interall_diag_V = (
            len(keep["interall_diag_V"]), _p(keep["interall_diag_idx"], _i32p), _p(keep["interall_diag_V"], _f64p))
        d.n_pairhop, d.pairhop_ij, d.pairhop_J = (len(keep["pairhop_J"]), _p(keep["pairhop_ij"], _i32p),
                                                  _p(keep["pairhop_J"], _f64p))
        return d, keep

    @classmethod
    def from_spec(cls, spec) -> "Engine":
        from .lattice import processed_terms
        return cls.from_terms(spec.model, spec.nsite, spec.nup, spec.ndown, processed_terms(spec))

    @classmethod
    def model_nnz(cls, spec) -> Tuple[int, int]:
        """(dim, off-diagonal elements of H) of a canonical sector, host only."""
        from .lattice import processed_terms
        d, keep = cls.desc_from_terms(spec.model, spec.nsite, spec.nup, spec.ndown, processed_terms(spec))
        dim, nnz = C.c_uint64(), C.c_uint64()
        check(lib().hphi_model_nnz(C.byref(d), C.byref(dim), C.byref(nnz)))
        return int(dim.value), int(nnz.value)

    @classmethod
    def from_def(cls, namelist: str) -> "Engine":
        h = _vp()
        mp = ModPara()
        check(lib().hphi_engine_create_from_def(namelist.encode(), C.byref(h), C.byref(mp)))
        return cls(h, mp)

    def close(self):
        if self.h:
            lib().hphi_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self) -> int:
        return int(lib().hphi_engine_device_bytes(self.h))

    # -- drop-in entry points ------------------------------------------------
    def sz(self, copy: bool = True) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        if not copy:
            check(lib().hphi_sz(self.h, None, None, None))
            return None, None, None
        l1 = np.zeros(self.dim + 1, np.uint64)
        l21 = np.zeros(max(self.size21, 1), np.uint64)
        l22 = np.zeros(max(self.size22, 1), np.uint64)
        check(lib().hphi_sz(self.h, _p(l1, _u64p), _p(l21, _u64p) if self.size21 else None,
                            _p(l22, _u64p) if self.size22 else None))
        return l1, l21, l22

    def diagonalcalc(self, copy: bool = True) -> Optional[np.ndarray]:
        if not copy:
            check(lib().hphi_diagonalcalc(self.h, None))
            return None
        d = np.zeros(self.dim + 1, np.float64)
        check(lib().hphi_diagonalcalc(self.h, _p(d, _f64p)))
        return d

    def mltply(self, v0, v1) -> complex:
        """Reference-shaped call: v0[1..dim] += H v1[1..dim] on 1-indexed host arrays."""
        assert v0.dtype == np.complex128 and v1.dtype == np.complex128 and v0.flags.c_contiguous
        pr = np.zeros(2)
        check(lib().hphi_mltply(self.h, v0.ctypes.data_as(_vp), v1.ctypes.data_as(_vp), _p(pr, _f64p)))
        return complex(pr[0], pr[1])

    def mltply_dev(self, v0: DeviceArray, v1: DeviceArray) -> complex:
        pr = np.zeros(2)
        check(lib().hphi_mltply_dev(self.h, v0.ptr, v1.ptr, _p(pr, _f64p)))
        return complex(pr[0], pr[1])

    def mltply_dev_scaled(self, v0: DeviceArray, v1: DeviceArray, c0: float, c1: float,
                          overwrite: bool = False) -> complex:
        """v0 = c0 v0 + c1 H v1 (v0 = c1 H v1 with overwrite); returns <v1|H|v1>."""
        pr = np.zeros(2)
        check(lib().hphi_mltply_dev_scaled(self.h, v0.ptr, v1.ptr, c0, c1, 1 if overwrite else 0, _p(pr, _f64p)))
        return complex(pr[0], pr[1])

    def lanczos_resid(self, r: DeviceArray, w: DeviceArray, vg: Optional[DeviceArray], f: float,
                      cg: float = 0.0) -> float:
        """r <- r - f w (and vg += cg w); returns |r|^2."""
        out = C.c_double()
        check(lib().hphi_vec_lanczos_resid(self.h, r.ptr, w.ptr, vg.ptr if vg is not None else None, f, cg,
                                           C.byref(out)))
        return out.value

    def mltply_block_dev(self, k: int, V0: DeviceArray, V1: DeviceArray) -> np.ndarray:
        pr = np.zeros(2 * k)
        check(lib().hphi_mltply_block_dev(self.h, k, V0.ptr, V1.ptr, _p(pr, _f64p)))
        return pr[0::2] + 1j * pr[1::2]

    def mltply_block_dev_noalloc(self, k: int, V0: DeviceArray, V1: DeviceArray) -> np.ndarray:
        """The allocation-free block H·v (tables and diagonal must be built)."""
        pr = np.zeros(2 * k)
        check(lib().hphi_mltply_block_dev_noalloc(self.h, k, V0.ptr, V1.ptr, _p(pr, _f64p)))
        return pr[0::2] + 1j * pr[1::2]

    def hv(self, v: np.ndarray) -> Tuple[np.ndarray, complex]:
        """H v of a 0-based host vector through device memory."""
        a = DeviceArray(self.dim).upload(v)
        b = DeviceArray(self.dim).zero()
        pr = self.mltply_dev(b, a)
        return b.download(), pr

    def time_mltply(self, k: int, V0: DeviceArray, V1: DeviceArray, warmup: int, reps: int) -> float:
        ms = C.c_double()
        check(lib().hphi_time_mltply_dev(self.h, k, V0.ptr, V1.ptr, warmup, reps, C.byref(ms)))
        return ms.value

    def tune(self, **knobs) -> None:
        for k, v in knobs.items():
            check(lib().hphi_engine_tune(self.h, k.encode(), int(v)))

    def get_tune(self, knob: str) -> int:
        v = C.c_int()
        check(lib().hphi_engine_get_tune(self.h, knob.encode(), C.byref(v)))
        return v.value

    def autotune(self, k: int, V0: DeviceArray, V1: DeviceArray, reps: int = 2) -> str:
        buf = C.create_string_buffer(8192)
        check(lib().hphi_autotune(self.h, k, V0.ptr, V1.ptr, reps, buf, 8192))
        return buf.value.decode()

    def lobpcg(self, k: int, max_iter: int = 500, eps: float = 1e-7, precg: bool = False,
               large_value: float = 0.0, seed: int = 0):
        """Block LOBPCG for the k lowest eigenpairs; returns (LobpcgResult,
        history[iterations, 1 + k] = (max residual, mu_0..mu_{k-1}) per step)."""
        p = LobpcgParams(k, max_iter, eps, 1 if precg else 0, large_value, seed)
        hist = np.zeros((max(max_iter, 1), k + 1))
        r = LobpcgResult()
        check(lib().hphi_lobpcg(self.h, C.byref(p), _p(hist, _f64p), C.byref(r)))
        return r, hist[: r.iterations].copy()

    def lobpcg_ex(self, k: int, X0: Optional[DeviceArray] = None, Xout: Optional[DeviceArray] = None,
                  max_iter: int = 500, eps: float = 1e-7, precg: bool = False, large_value: float = 0.0,
                  seed: int = 0):
        """lobpcg from the start vectors X0 (k vectors one after another, or None
        for the seeded ones), the final normalised eigenvectors into Xout."""
        p = LobpcgParams(k, max_iter, eps, 1 if precg else 0, large_value, seed)
        hist = np.zeros((max(max_iter, 1), k + 1))
        r = LobpcgResult()
        check(lib().hphi_lobpcg_ex(self.h, C.byref(p), X0.ptr if X0 is not None else None,
                                   Xout.ptr if Xout is not None else None, _p(hist, _f64p), C.byref(r)))
        return r, hist[: r.iterations].copy()

    def tpq_from(self, W0: DeviceArray, k: int, steps: int, step0: int = 0, W_end: Optional[DeviceArray] = None,
                 layout: int = 1, use_graph: bool = True, large_value: float = 0.0):
        """Microcanonical TPQ from the vectors W0 (k vectors one after another);
        step0 > 0 resumes (rows of steps step0..steps-1); returns (out, TpqResult)."""
        p = TpqParams(k, steps, 0, layout, 1 if use_graph else 0, large_value, 0)
        out = np.zeros((steps - step0, k, 12))
        r = TpqResult()
        check(lib().hphi_tpq_from(self.h, C.byref(p), W0.ptr, step0, _p(out, _f64p), C.byref(r),
                                  W_end.ptr if W_end is not None else None))
        return out, r

    def ctpq_from(self, W0: DeviceArray, k: int, steps: int, step0: int = 0, W_end: Optional[DeviceArray] = None,
                  expand_coef: int = 0, large_value: float = 0.0):
        """Canonical TPQ from the vectors W0, as tpq_from."""
        p = CtpqParams(k, steps, 0, expand_coef, large_value, 0)
        out = np.zeros((steps - step0, k, 12))
        r = TpqResult()
        check(lib().hphi_ctpq_from(self.h, C.byref(p), W0.ptr, step0, _p(out, _f64p), C.byref(r),
                                   W_end.ptr if W_end is not None else None))
        return out, r

    def tpq(self, k: int, steps: int, seed: int = 0, init_kind: int = -1, layout: int = 1,
            use_graph: bool = True, large_value: float = 0.0):
        """Microcanonical TPQ for k samples; returns (out[steps, k, 12], TpqResult)."""
        p = TpqParams(k, steps, init_kind, layout, 1 if use_graph else 0, large_value, seed)
        out = np.zeros((steps, k, 12))
        r = TpqResult()
        check(lib().hphi_tpq(self.h, C.byref(p), _p(out, _f64p), C.byref(r)))
        return out, r

    def ctpq(self, k: int, steps: int, seed: int = 0, init_kind: int = -1, expand_coef: int = 0,
             large_value: float = 0.0):
        """Canonical TPQ for k samples; returns (out[steps, k, 12], TpqResult)."""
        p = CtpqParams(k, steps, init_kind, expand_coef, large_value, seed)
        out = np.zeros((steps, k, 12))
        r = TpqResult()
        check(lib().hphi_ctpq(self.h, C.byref(p), _p(out, _f64p), C.byref(r)))
        return out, r

    def diag_maxabs(self) -> float:
        m = C.c_double()
        check(lib().hphi_diag_maxabs(self.h, C.byref(m)))
        return m.value

    def fill_random(self, v: DeviceArray, k: int, seed: int) -> None:
        check(lib().hphi_fill_random(v.ptr, self.dim, k, seed))

    def lanczos(self, n_eigen: int = 1, max_iter: int = 2000, eps: float = 1e-14, target: int = 2,
                seed: int = 0, check_every: int = 1, mode: int = 0) -> LanczosResult:
        p = LanczosParams(max_iter, n_eigen, eps, target, seed, check_every, mode)
        r = LanczosResult()
        check(lib().hphi_lanczos(self.h, C.byref(p), C.byref(r)))
        return r

    def lanczos_restart(self, restart_in: Optional[Tuple[str, str]] = None,
                        restart_out: Optional[Tuple[str, str]] = None, n_eigen: int = 1, max_iter: int = 2000,
                        eps: float = 1e-14, target: int = 2, seed: int = 0, check_every: int = 1,
                        mode: int = 0) -> LanczosResult:
        """Lanczos continued from / writing the reference's restart files:
        restart_in / restart_out = (TMComponents path, recalcvec path)."""
        p = LanczosParams(max_iter, n_eigen, eps, target, seed, check_every, mode)
        r = LanczosResult()
        enc = lambda x: x.encode() if x else None
        ti, vi = restart_in or (None, None)
        to, vo = restart_out or (None, None)
        check(lib().hphi_lanczos_restart(self.h, C.byref(p), enc(ti), enc(vi), enc(to), enc(vo), C.byref(r)))
        return r

    def lanczos_eigenvector(self, k: int = 0, n_eigen: int = 1, max_iter: int = 2000, eps: float = 1e-14,
                            target: int = 2, seed: int = 0, check_every: int = 1,
                            mode: int = 0) -> Tuple["DeviceArray", LanczosResult]:
        """Eigenvector of the k-th lowest Ritz value (a DeviceArray, normalised)."""
        p = LanczosParams(max_iter, n_eigen, eps, target, seed, check_every, mode)
        r = LanczosResult()
        v = DeviceArray(self.dim)
        check(lib().hphi_lanczos_eigenvector(self.h, C.byref(p), k, v.ptr, C.byref(r)))
        return v, r

    def expec_spin_nbody(self, v: "DeviceArray", nbody: int, idx):
        """SpinGC three/four/six-body Green's functions: (values, ok flags)."""
        q = np.ascontiguousarray(np.asarray(idx, np.int32).reshape(-1, 4 * nbody))
        out = np.zeros(2 * q.shape[0])
        ok = np.zeros(q.shape[0], np.int32)
        check(lib().hphi_expec_spin_nbody(self.h, v.ptr, nbody, q.shape[0], _p(q, _i32p), _p(out, _f64p),
                                          _p(ok, _i32p)))
        return out[0::2] + 1j * out[1::2], ok

    def expec_energy_flct(self, v: "DeviceArray") -> Dict[str, float]:
        """energy, var, doublon, doublon2, num, num2, sz, sz2 of the device vector v."""
        out = Phys()
        check(lib().hphi_expec_energy_flct(self.h, v.ptr, C.byref(out)))
        return out.as_dict()

    def _green(self, fn, v: "DeviceArray", idx, width: int) -> np.ndarray:
        idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int32).reshape(-1, width))
        n = idx.shape[0]
        out = np.zeros(2 * n)
        check(fn(self.h, v.ptr, n, idx.ctypes.data_as(C.POINTER(C.c_int)), _p(out, _f64p)))
        return out[0::2] + 1j * out[1::2]

    def cisajs(self, v: "DeviceArray", idx) -> np.ndarray:
        """<v|c+_{i s1} c_{j s2}|v> for rows (i, s1, j, s2)."""
        return self._green(lib().hphi_expec_cisajs, v, idx, 4)

    def cisajscktalt(self, v: "DeviceArray", idx) -> np.ndarray:
        """<v|c+_{i s1} c_{j s2} c+_{k s3} c_{l s4}|v> for rows of 8 indices."""
        return self._green(lib().hphi_expec_cisajscktalt, v, idx, 8)


def herm_eig(a: np.ndarray):
    """Eigenvalues (ascending) and eigenvectors of a small Hermitian matrix
    (the library's subspace solver)."""
    a = np.asfortranarray(np.asarray(a, np.complex128))
    n = a.shape[0]
    w = np.zeros(n)
    v = np.zeros((n, n), np.complex128, order="F")
    check(lib().hphi_herm_eig(n, a.ctypes.data_as(_f64p), _p(w, _f64p), v.ctypes.data_as(_f64p)))
    return w, v


def tridiag_lowest(a: np.ndarray, b: np.ndarray, k: int) -> np.ndarray:
    """Lowest k eigenvalues of the symmetric tridiagonal (host C routine of the library)."""
    L = lib()
    L.hphi_tridiag_lowest.argtypes = [_f64p, _f64p, C.c_int, C.c_int, _f64p]
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    ev = np.zeros(k)
    L.hphi_tridiag_lowest(_p(a, _f64p), _p(b, _f64p) if b.size else None, len(a), k, _p(ev, _f64p))
    return ev


class Server:
    """Continuous-batching Lanczos service (hphi_server_* in the C ABI)."""

    def __init__(self, hbm_budget: int = 16 << 30, max_jobs: int = 64, n_eigen: int = 4,
                 eps: float = 1e-12, max_iter: int = 2000, check_every: int = 5):
        p = ServerParams(hbm_budget, max_jobs, n_eigen, eps, max_iter, check_every)
        self.h = _vp()
        check(lib().hphi_server_create(C.byref(p), C.byref(self.h)))

    def now(self) -> float:
        return lib().hphi_server_now(self.h)

    def submit(self, spec, seed: int, t_arrival: float) -> int:
        from .lattice import processed_terms
        d, keep = Engine.desc_from_terms(spec.model, spec.nsite, spec.nup, spec.ndown, processed_terms(spec))
        i = C.c_int()
        check(lib().hphi_server_submit(self.h, C.byref(d), seed, t_arrival, C.byref(i)))
        del keep
        return i.value

    def tick(self):
        a, w = C.c_int(), C.c_int()
        check(lib().hphi_server_tick(self.h, C.byref(a), C.byref(w)))
        return a.value, w.value

    def stats(self):
        out = np.zeros(10)
        check(lib().hphi_server_stats(self.h, _p(out, _f64p)))
        return {"admit_s": out[0], "gpu_s": out[1], "check_s": out[2], "ticks": int(out[3]),
                "job_steps": int(out[4]), "slab_steps": int(out[5]), "hv_s": out[6], "hv_bytes": out[7],
                "hv_launches": int(out[8]), "moves": int(out[9])}

    def results(self, max_n: int = 4096):
        buf = (JobResult * max_n)()
        n = C.c_int()
        check(lib().hphi_server_results(self.h, buf, max_n, C.byref(n)))
        return [buf[i] for i in range(n.value)]

    def close(self):
        if self.h:
            lib().hphi_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- the reference's output and restart files (host-side C writers/readers) ----

def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def write_vec_file(path: str, step: int, v: np.ndarray) -> None:
    """<head>_eigenvec_<k>_rank_<r>.dat / tmpvec_set<i>_rank_<r>.dat layout."""
    v = _c128(v)
    check(lib().hphi_write_vec_file(path.encode(), step, v.size, v.ctypes.data_as(_f64p)))


def read_vec_file(path: str, dim: int) -> Tuple[np.ndarray, int]:
    v = np.zeros(dim, np.complex128)
    st = C.c_int()
    check(lib().hphi_read_vec_file(path.encode(), dim, v.ctypes.data_as(_f64p), C.byref(st)))
    return v, st.value


def write_restart_vecs(path: str, step: int, v0: np.ndarray, v1: np.ndarray) -> None:
    v0, v1 = _c128(v0), _c128(v1)
    check(lib().hphi_write_restart_vecs(path.encode(), step, v0.size, v0.ctypes.data_as(_f64p),
                                        v1.ctypes.data_as(_f64p)))


def read_restart_vecs(path: str, dim: int) -> Tuple[np.ndarray, np.ndarray, int]:
    v0, v1 = np.zeros(dim, np.complex128), np.zeros(dim, np.complex128)
    st = C.c_uint64()
    check(lib().hphi_read_restart_vecs(path.encode(), dim, v0.ctypes.data_as(_f64p), v1.ctypes.data_as(_f64p),
                                       C.byref(st)))
    return v0, v1, st.value


def write_tmcomponents(path: str, norm: float, alpha: np.ndarray, beta: np.ndarray) -> None:
    a, b = np.ascontiguousarray(alpha, np.float64), np.ascontiguousarray(beta, np.float64)
    check(lib().hphi_write_tmcomponents(path.encode(), a.size, norm, _p(a, _f64p), _p(b, _f64p)))


def read_tmcomponents(path: str, cap: int = 100000) -> Tuple[float, np.ndarray, np.ndarray]:
    a, b = np.zeros(cap), np.zeros(cap)
    n, nrm = C.c_int(), C.c_double()
    check(lib().hphi_read_tmcomponents(path.encode(), cap, C.byref(n), C.byref(nrm), _p(a, _f64p), _p(b, _f64p)))
    return nrm.value, a[: n.value].copy(), b[: n.value].copy()


def write_energy(path: str, phys: Dict[str, float]) -> None:
    p = Phys(**{k: phys.get(k, 0.0) for k, _ in Phys._fields_})
    check(lib().hphi_write_energy(path.encode(), C.byref(p)))


def write_cisajs(path: str, idx, vals) -> None:
    idx = np.ascontiguousarray(np.asarray(idx, np.int32).reshape(-1, 4))
    v = np.ascontiguousarray(np.column_stack([np.real(vals), np.imag(vals)]).reshape(-1), np.float64)
    check(lib().hphi_write_cisajs(path.encode(), idx.shape[0], idx.ctypes.data_as(C.POINTER(C.c_int)),
                                  _p(v, _f64p)))


def write_cisajscktalt(path: str, model: int, idx, vals) -> None:
    idx = np.ascontiguousarray(np.asarray(idx, np.int32).reshape(-1, 8))
    v = np.ascontiguousarray(np.column_stack([np.real(vals), np.imag(vals)]).reshape(-1), np.float64)
    check(lib().hphi_write_cisajscktalt(path.encode(), model, idx.shape[0], idx.ctypes.data_as(C.POINTER(C.c_int)),
                                        _p(v, _f64p)))


def write_tpq_files(directory: str, rand_i: int, rows: np.ndarray) -> None:
    """rows[steps, 12] of one sample (a slice out[:, c] of Engine.tpq)."""
    r = np.ascontiguousarray(rows, np.float64)
    check(lib().hphi_write_tpq_files(directory.encode(), rand_i, r.shape[0], _p(r, _f64p), 12))


def write_ctpq_files(directory: str, rand_i: int, rows: np.ndarray) -> None:
    """rows[steps, 12] of one canonical TPQ sample (a slice out[:, c] of Engine.ctpq)."""
    r = np.ascontiguousarray(rows, np.float64)
    check(lib().hphi_write_ctpq_files(directory.encode(), rand_i, r.shape[0], _p(r, _f64p), 12))


def append_tpq_files(directory: str, rand_i: int, step0: int, rows: np.ndarray, norm1: float = 0.0) -> None:
    """rows[steps - step0, 12] of one sample of a resumed run, appended."""
    r = np.ascontiguousarray(rows, np.float64)
    check(lib().hphi_append_tpq_files(directory.encode(), rand_i, step0, step0 + r.shape[0], _p(r, _f64p), 12,
                                      norm1))


def write_energy_states(path: str, phys) -> None:
    """The block solvers' energy file from a list of phys dicts."""
    arr = (Phys * len(phys))(*[Phys(**{k: p.get(k, 0.0) for k, _ in Phys._fields_}) for p in phys])
    check(lib().hphi_write_energy_states(path.encode(), len(phys), arr))


def write_nbody_green(path: str, nbody: int, idx, vals, ok) -> None:
    q = np.ascontiguousarray(np.asarray(idx, np.int32).reshape(-1, 4 * nbody))
    v = np.ascontiguousarray(np.column_stack([np.real(vals), np.imag(vals)]).reshape(-1), np.float64)
    k = np.ascontiguousarray(np.asarray(ok, np.int32))
    check(lib().hphi_write_nbody_green(path.encode(), nbody, q.shape[0], _p(q, _i32p), _p(v, _f64p), _p(k, _i32p)))
