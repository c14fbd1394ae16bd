"""The C ABI library loads and exports every symbol include/dalia_hip.h declares (no GPU calls)."""

import os
import re

from dalia_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "dalia_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"^\s*(?:int|long long|dh_handle\s*\*)\s*(dh_\w+)\s*\(", txt, flags=re.M))


def test_header_declares_expected_surface():
    syms = header_symbols()
    for s in ("dh_create", "dh_destroy", "dh_assemble", "dh_cholesky", "dh_solve", "dh_logdet",
              "dh_selected_inversion", "dh_scatter_coo", "dh_contract_fem", "dh_prior_eig"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = header_symbols()
    assert syms == set(_lib.SIGNATURES), syms ^ set(_lib.SIGNATURES)
    for s in syms:
        assert hasattr(lib, s), s
    assert lib.dh_version() == 1


def test_bad_arguments_rejected_without_gpu():
    lib = _lib.load()
    import ctypes
    h = ctypes.c_void_p()
    assert lib.dh_create(0, 10, 0, None, ctypes.byref(h)) == -1
    assert lib.dh_create(2, 10, 65, None, ctypes.byref(h)) == -1
    assert lib.dh_create(2, 10, 0, None, None) == -1
    assert lib.dh_destroy(None) == -1
