"""ctypes binding of `libdalia_hip.so` (the C ABI declared in include/dalia_hip.h).

The library is built in-tree by `dalia_amd._build.build()`.  There is no
fallback: if the shared object is missing or a GPU is absent, loading fails
with an error that says so.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdalia_hip.so")

SPARSITY = {"bt": 0, "bta": 1}

c_double_p = ctypes.POINTER(ctypes.c_double)


class DhLayout(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("b", ctypes.c_int), ("a", ctypes.c_int),
        ("bp", ctypes.c_int), ("ap", ctypes.c_int),
        ("diag", ctypes.c_void_p), ("lower", ctypes.c_void_p), ("arrow", ctypes.c_void_p),
        ("tip", ctypes.c_void_p), ("dinv", ctypes.c_void_p),
    ]


class DhModel(ctypes.Structure):
    _fields_ = [
        ("ns", ctypes.c_int), ("nt", ctypes.c_int), ("nb", ctypes.c_int),
        ("fem_nnz", ctypes.c_int),
        ("fem_rowptr", ctypes.c_void_p), ("fem_col", ctypes.c_void_p), ("fem_val", ctypes.c_void_p),
        ("m_diag", ctypes.c_void_p), ("m_off", ctypes.c_void_p),
        ("ata_nnz", ctypes.c_longlong),
        ("ata_rowptr", ctypes.c_void_p), ("ata_col", ctypes.c_void_p), ("ata_val", ctypes.c_void_p),
        ("ata_arrow", ctypes.c_void_p), ("ata_tip", ctypes.c_void_p),
        ("fe_prec", ctypes.c_double),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_LL = ctypes.c_longlong

#: name -> (restype, argtypes); every symbol include/dalia_hip.h declares.
SIGNATURES = {
    "dh_version": (_I, []),
    "dh_create": (_I, [_I, _I, _I, _P, ctypes.POINTER(_P)]),
    "dh_destroy": (_I, [_P]),
    "dh_marginal_variances": (_I, [_P, _P]),
    "dh_assemble_coreg": (_I, [_P, ctypes.POINTER(DhModel), _I, _P, _I, _I]),
    "dh_scale_indexed": (_I, [_P, _P, _P, _P, _P, _LL]),
    "dh_obs_variances": (_I, [_P, _LL, _P, _P, _P, _P, _P]),
    "dh_get_layout": (_I, [_P, ctypes.POINTER(DhLayout)]),
    "dh_set_stream": (_I, [_P, _P]),
    "dh_memory_bytes": (_LL, [_P]),
    "dh_assemble": (_I, [_P, ctypes.POINTER(DhModel), _P, _I, _I]),
    "dh_scatter_coo": (_I, [_P, _LL, _P, _P, _P, _I]),
    "dh_cholesky": (_I, [_P, _I]),
    "dh_solve": (_I, [_P, _P, _I]),
    "dh_logdet": (_I, [_P, _I, _P]),
    "dh_logdet_noalloc": (_I, [_P, _I, _I, _I, _LL, _P, _I, _I, _P, _P]),
    "dh_selected_inversion": (_I, [_P, _I]),
    "dh_get_info": (_I, [_P, ctypes.POINTER(_I)]),
    "dh_qc_matvec": (_I, [_P, ctypes.POINTER(DhModel), _P, _P, _P, _P]),
    "dh_scale": (_I, [_P, _P, _D, _P, _LL]),
    "dh_scale_dev": (_I, [_P, _P, _P, _P, _LL]),
    "dh_dot": (_I, [_P, _P, _P, _LL, _P]),
    "dh_contract_fem": (_I, [_P, ctypes.POINTER(DhModel), _P]),
    "dh_contract_fem_vec": (_I, [_P, ctypes.POINTER(DhModel), _P, _P]),
    "dh_contract_ata": (_I, [_P, ctypes.POINTER(DhModel), _P]),
    "dh_contract_fem_v": (_I, [_P, ctypes.POINTER(DhModel), _I, _I, _P]),
    "dh_contract_fem_vec_v": (_I, [_P, ctypes.POINTER(DhModel), _P, _I, _I, _I, _P]),
    "dh_contract_ata_v": (_I, [_P, ctypes.POINTER(DhModel), _I, _I, _P]),
    "dh_ata_quad_v": (_I, [_P, ctypes.POINTER(DhModel), _P, _I, _I, _I, _P]),
    "dh_prior_eig": (_I, [_P, ctypes.POINTER(DhModel), _P, c_double_p, _P]),
    "dh_prior_eig_dev": (_I, [_P, ctypes.POINTER(DhModel), _P, _P, _P]),
    "dh_set_tuning": (_I, [_P, _I, _I]),
    "dh_get_tile_profile": (_I, [_P, _P, _I]),
    "dh_dag_table": (_I, [_I, _I, _I, _P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "dh_multi_create": (_I, [ctypes.POINTER(_P), _I, ctypes.POINTER(_P)]),
    "dh_multi_destroy": (_I, [_P]),
    "dh_multi_cholesky": (_I, [_P, _I, _P]),
    "dh_ckpt_create": (_I, [_I, _I, _I, _I, _P, ctypes.POINTER(_P)]),
    "dh_ckpt_destroy": (_I, [_P]),
    "dh_ckpt_memory_bytes": (_LL, [_P]),
    "dh_ckpt_vectors": (_P, [_P]),
    "dh_ckpt_set_stream": (_I, [_P, _P]),
    "dh_ckpt_get_info": (_I, [_P, ctypes.POINTER(_I)]),
    "dh_ckpt_set_tuning": (_I, [_P, _I, _I]),
    "dh_ckpt_timing_enable": (_I, [_P, _I]),
    "dh_ckpt_timing_read": (_I, [_P, c_double_p, ctypes.POINTER(_LL)]),
    "dh_ckpt_forward": (_I, [_P, ctypes.POINTER(DhModel), _P, _I, _P, _P]),
    "dh_ckpt_backward": (_I, [_P, ctypes.POINTER(DhModel), _P, _P, _I, _P, _P, _P]),
    "dh_timing_enable": (_I, [_P, _I]),
    "dh_timing_read": (_I, [_P, c_double_p, ctypes.POINTER(_LL)]),
}

#: kernel kinds of dh_timing_enable
TIMED_KERNELS = {"k_syrk": 0, "k_update": 1, "k_sel": 2, "k_panel": 3, "k_leaf": 4, "k_assemble_rows": 5,
                 "k_syrk_tail": 6, "k_pobtaf_dag": 7}

_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m dalia_amd._build` "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    # torch first: the library then binds to the HIP runtime that torch has loaded (torch ships
    # its own), so that torch's streams and device pointers are valid in the library's calls
    import torch  # noqa: F401
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class DhError(RuntimeError):
    pass


#: bits of the NaN that dh_logdet writes when a readiness wait of the factorization gave up
NAN_GIVEUP = 0x7FF80000DEAD0004


def is_giveup_nan(value) -> bool:
    return bool(np.isnan(value)) and int(np.float64(value).view(np.uint64)) == NAN_GIVEUP


def raise_nan_logdet(lib, h, *values) -> None:
    """A NaN log-determinant: raise DhError when a readiness wait of the factorization gave up
    (the factor is invalid; the NaN carries NAN_GIVEUP, or handle h still reports it), else the
    reference's ValueError for an input that is not SPD (`structured_solver.py:163-164`).
    Synchronises the handle's stream."""
    v = ctypes.c_int()
    rc = lib.dh_get_info(h, ctypes.byref(v))
    if rc == -4 or any(is_giveup_nan(x) for x in values):
        raise DhError("dh_cholesky failed: a fused launch gave up waiting for a factored tile "
                      "(the factor, log-determinant and solve are invalid)")
    raise ValueError("Logdet is NaN. Check the input matrix.")


def check(rc: int, what: str) -> None:
    if rc != 0:
        names = {-1: "bad argument", -2: "HIP error", -3: "wrong state",
                 -4: "a fused launch gave up waiting for a factored tile"}
        raise DhError(f"{what} failed: {names.get(rc, rc)}")
