"""Host-side argument checks of the drop-in solver, raised before any device work (no GPU needed).

The reference fails on the host for a matrix or right-hand side that does not match its blocks
(`src/dalia/solvers/structured_solver.py:192-250`, `:102-149`); the drop-in must too, instead of
handing the library a buffer that is too short or entries outside its blocks.
"""

import struct

import numpy as np
import pytest
import scipy.sparse as sps

from dalia_amd import _lib
from dalia_amd.solver import HipBTASolver


def _bare_solver(n, b, a):
    """A solver object with its sizes set and no device handle (the checks run before any)."""
    s = HipBTASolver.__new__(HipBTASolver)
    s.n_diag_blocks, s.diagonal_blocksize, s.arrowhead_blocksize = n, b, a
    s.h = None
    return s


@pytest.mark.parametrize("shape", [(3 * 10 + 2 + 1, 3 * 10 + 2 + 1), (3 * 10 + 1, 3 * 10 + 1), (32, 31)])
def test_cholesky_rejects_wrong_matrix_order(shape):
    s = _bare_solver(3, 10, 2)
    A = sps.identity(shape[0], format="csr")[:, : shape[1]]
    with pytest.raises(ValueError, match="does not match"):
        s._spmatrix_to_structured(A, "bta")


def test_bt_accepts_both_orders_and_rejects_others():
    s = _bare_solver(3, 10, 2)
    with pytest.raises(ValueError, match="does not match"):
        s._spmatrix_to_structured(sps.identity(29, format="csr"), "bt")
    with pytest.raises(ValueError, match="does not match"):
        s._spmatrix_to_structured(sps.identity(30, format="csr"), "bta")


@pytest.mark.parametrize("sparsity,length", [("bta", 31), ("bta", 33), ("bta", 30), ("bt", 32), ("bt", 29)])
def test_solve_rejects_wrong_rhs_length(sparsity, length):
    s = _bare_solver(3, 10, 2)
    with pytest.raises(ValueError, match="does not match"):
        s._rhs_view(np.zeros(length), sparsity)


def test_solve_rejects_non_float64_and_unknown_types():
    s = _bare_solver(3, 10, 2)
    with pytest.raises(ValueError):
        s._rhs_view(np.zeros(32, dtype=np.float32), "bta")
    with pytest.raises(ValueError):
        s._rhs_view([0.0] * 32, "bta")


def test_unknown_sparsity_is_rejected_first():
    s = _bare_solver(3, 10, 2)
    with pytest.raises(ValueError, match="Unknown sparsity"):
        s._spmatrix_to_structured(sps.identity(32, format="csr"), "dense")


def test_giveup_nan_is_recognised():
    giveup = struct.unpack("<d", struct.pack("<Q", _lib.NAN_GIVEUP))[0]
    assert np.isnan(giveup) and _lib.is_giveup_nan(giveup)
    assert _lib.is_giveup_nan(np.float64(giveup))
    assert not _lib.is_giveup_nan(float("nan"))
    assert not _lib.is_giveup_nan(1.0)


def test_header_and_binding_agree_on_the_giveup_nan():
    import os
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                            "dalia_hip.h")).read()
    assert f"#define DH_NAN_GIVEUP 0x{_lib.NAN_GIVEUP:016x}ULL" in hdr
