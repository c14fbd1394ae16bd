"""GPU parity of the BTA kernels (through the C ABI) against the CPU oracle.

Bar: logdet within 1e-10 relative; factor, solve and selected-inverse blocks
within 1e-9 relative Frobenius.
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import bta as obta
from tests.bta_cases import random_bta

pytestmark = pytest.mark.gpu

SHAPES = [(1, 64, 0), (3, 64, 0), (4, 100, 3), (3, 130, 5), (2, 200, 6), (5, 64, 64), (1, 70, 2), (6, 256, 4), (17, 96, 3),
          (5, 300, 4)]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _oracle(M, n, b, a):
    D, L, R, T = obta.dense_to_bta(M, n, b, a)
    if a == 0:
        R = T = None
    obta.pobtaf(D, L, R, T)
    return D, L, R, T


@pytest.mark.parametrize("n,b,a", SHAPES)
def test_pobtaf_logdet_solve(n, b, a):
    from dalia_amd.solver import HipBTASolver
    print(f"case pobtaf n={n} b={b} a={a}", flush=True)
    M, S = random_bta(n, b, a, seed=n * 1000 + b + a)
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    s.cholesky(S, sparsity="bta")
    assert s.info() == 0
    D, L, R, T = _oracle(M, n, b, a)
    Dg = s.A_diagonal_blocks.cpu().numpy()
    for i in range(n):
        assert rel(np.tril(Dg[i]), np.tril(D[i])) < 1e-9
    if n > 1:
        assert rel(s.A_lower_diagonal_blocks.cpu().numpy(), L) < 1e-9
    if a:
        assert rel(s.A_arrow_bottom_blocks.cpu().numpy(), R) < 1e-9
        assert rel(np.tril(s.A_arrow_tip_block.cpu().numpy()), np.tril(T)) < 1e-9
    ld = s.logdet("bta")
    sign, ld_ref = np.linalg.slogdet(M)
    assert sign > 0
    assert abs(ld - ld_ref) <= 1e-10 * abs(ld_ref) + 1e-12
    rng = np.random.default_rng(7)
    rhs = rng.standard_normal(n * b + a)
    x = torch.as_tensor(rhs.copy(), device="cuda")
    s.solve(x, sparsity="bta")
    xr = np.linalg.solve(M, rhs)
    assert rel(x.cpu().numpy(), xr) < 1e-9


@pytest.mark.parametrize("side", [1, 0])
@pytest.mark.parametrize("n,b,a", SHAPES)
def test_pobtasi(n, b, a, side):
    """Selected inversion with the products of the factor on a second stream or not (side)."""
    from dalia_amd import _lib
    from dalia_amd.solver import HipBTASolver
    print(f"case pobtasi n={n} b={b} a={a} side={side}", flush=True)
    M, S = random_bta(n, b, a, seed=n * 7 + b + a)
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    lib = _lib.load()
    _lib.check(lib.dh_set_tuning(s.h, 3, side), "dh_set_tuning")
    s.cholesky(S, sparsity="bta")
    s.selected_inversion("bta")
    info = ctypes.c_int(-1)
    assert lib.dh_get_info(s.h, ctypes.byref(info)) == 0 and info.value == 0  # no flag wait gave up
    Minv = np.linalg.inv(M)
    Xd, Xl, Xr, Xt = obta.dense_to_bta(Minv, n, b, a)
    assert rel(s.A_diagonal_blocks.cpu().numpy(), Xd) < 1e-9
    if n > 1:
        assert rel(s.A_lower_diagonal_blocks.cpu().numpy(), Xl) < 1e-9
    if a:
        assert rel(s.A_arrow_bottom_blocks.cpu().numpy(), Xr) < 1e-9
        assert rel(s.A_arrow_tip_block.cpu().numpy(), Xt) < 1e-9


@pytest.mark.parametrize("overlap", [1, 2])
@pytest.mark.parametrize("n,b,a", SHAPES)
def test_overlap_factor_solve_selinv(n, b, a, overlap):
    """The inverse of the diagonal factors (1) and the products of the factor (2) formed on a second
    stream beside the persistent factorization; two factorizations in a row on one handle."""
    from dalia_amd import _lib
    from dalia_amd.solver import HipBTASolver
    lib = _lib.load()
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    assert s.set_overlap(overlap)
    for rep in range(2):
        print(f"case overlap={overlap} n={n} b={b} a={a} rep={rep}", flush=True)
        M, S = random_bta(n, b, a, seed=n * 31 + b + a + rep)
        s.cholesky(S, sparsity="bta")
        info = ctypes.c_int(-1)
        assert lib.dh_get_info(s.h, ctypes.byref(info)) == 0 and info.value == 0
        ld = s.logdet("bta")
        ld_ref = np.linalg.slogdet(M)[1]
        assert abs(ld - ld_ref) <= 1e-10 * abs(ld_ref) + 1e-12
        rhs = np.random.default_rng(rep).standard_normal(n * b + a)
        x = torch.as_tensor(rhs.copy(), device="cuda")
        s.solve(x, sparsity="bta")
        assert rel(x.cpu().numpy(), np.linalg.solve(M, rhs)) < 1e-9
        s.selected_inversion("bta")
        assert lib.dh_get_info(s.h, ctypes.byref(info)) == 0 and info.value == 0
        Xd, Xl, Xr, Xt = obta.dense_to_bta(np.linalg.inv(M), n, b, a)
        assert rel(s.A_diagonal_blocks.cpu().numpy(), Xd) < 1e-9
        if n > 1:
            assert rel(s.A_lower_diagonal_blocks.cpu().numpy(), Xl) < 1e-9
        if a:
            assert rel(s.A_arrow_bottom_blocks.cpu().numpy(), Xr) < 1e-9
            assert rel(s.A_arrow_tip_block.cpu().numpy(), Xt) < 1e-9


@pytest.mark.parametrize("overlap", [0, 2])
@pytest.mark.parametrize("n,b", [(4, 100), (3, 192)])
def test_bt_path_ignores_arrow(n, b, overlap):
    """sparsity 'bt' factors the block-tridiagonal part only (pobtf / pobts / pobtsi)."""
    from dalia_amd.solver import HipBTASolver
    a = 4
    print(f"case bt n={n} b={b} overlap={overlap}", flush=True)
    M, S = random_bta(n, b, a, seed=11 + b)
    Mbt = M[: n * b, : n * b]
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    assert s.set_overlap(overlap)
    s.cholesky(S, sparsity="bt")
    ld = s.logdet("bt")
    ld_ref = np.linalg.slogdet(Mbt)[1]
    assert abs(ld - ld_ref) <= 1e-10 * abs(ld_ref)
    rhs = np.random.default_rng(3).standard_normal(n * b)
    x = torch.as_tensor(rhs.copy(), device="cuda")
    s.solve(x, "bt")
    assert rel(x.cpu().numpy(), np.linalg.solve(Mbt, rhs)) < 1e-9
    s.selected_inversion("bt")
    Xd, Xl, _, _ = obta.dense_to_bta(np.linalg.inv(Mbt), n, b, 0)
    assert rel(s.A_diagonal_blocks.cpu().numpy(), Xd) < 1e-9
    assert rel(s.A_lower_diagonal_blocks.cpu().numpy(), Xl) < 1e-9


def test_not_spd_reports_and_logdet_raises():
    """An indefinite diagonal block: NaN logdet raises ValueError (structured_solver.py:163-164)."""
    from dalia_amd.solver import HipBTASolver
    n, b, a = 3, 64, 2
    M, _ = random_bta(n, b, a, seed=5)
    M[b + 10, b + 10] = -50.0
    import scipy.sparse as sps
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    s.cholesky(sps.csr_matrix(M), sparsity="bta")
    assert s.info() > 0
    with pytest.raises(ValueError):
        s.logdet("bta")


def test_unknown_sparsity_raises():
    from dalia_amd.solver import HipBTASolver
    s = HipBTASolver(diagonal_blocksize=64, n_diag_blocks=2, arrowhead_blocksize=0)
    with pytest.raises(ValueError):
        s.logdet("dense")


def test_structured_to_spmatrix_marginal_variances():
    """Diagonal of the selected inverse read back on Q's pattern (core/dalia.py:938-962)."""
    from dalia_amd.solver import HipBTASolver
    n, b, a = 3, 80, 3
    M, S = random_bta(n, b, a, seed=99)
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    s.cholesky(S, "bta")
    s.selected_inversion("bta")
    import scipy.sparse as sps
    out = s._structured_to_spmatrix(sps.eye(n * b + a), "bta")
    assert rel(out.diagonal(), np.diag(np.linalg.inv(M))) < 1e-9
