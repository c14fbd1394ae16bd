"""GPU parity of the BTA kernels at block sizes of the benchmark shapes.

With 13 or more 64-column tiles per block the persistent factorization runs every kind of task
and every range its schedule has (early tasks of the sub-diagonal tiles, own-term ranges of the
early tasks and helpers, helpers queued inside the column loop), which smaller blocks do not.
(3, 2025, 6) has the block size of the ns = 2025 benchmark (32 tile columns).

Bar (against dense LAPACK on the densified matrix): logdet within 1e-10 relative; factor, solve
and selected-inverse blocks within 1e-9 relative Frobenius.  The same dense brute-force pattern
as the reference's example check (`examples/gst_small/run.py:116-123`).
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import bta as obta
from tests.bta_cases import random_bta

pytestmark = pytest.mark.gpu

#: (n, b, a): tile columns per block 13, 16, 32 and 16 without arrowhead
LARGE = [(4, 832, 6), (2, 1024, 4), (3, 2025, 6), (5, 1000, 0)]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module", params=LARGE, ids=lambda s: "n%d_b%d_a%d" % s)
def case(request):
    n, b, a = request.param
    M, S = random_bta(n, b, a, seed=n * 1000 + b + a)
    return n, b, a, M, S


def _check_factor(s, M, n, b, a):
    D, L, R, T = obta.dense_to_bta(M, n, b, a)
    if a == 0:
        R = T = None
    obta.pobtaf(D, L, R, T)
    Dg = s.A_diagonal_blocks.cpu().numpy()
    for i in range(n):
        assert rel(np.tril(Dg[i]), np.tril(D[i])) < 1e-9, i
    if n > 1:
        assert rel(s.A_lower_diagonal_blocks.cpu().numpy(), L) < 1e-9
    if a:
        assert rel(s.A_arrow_bottom_blocks.cpu().numpy(), R) < 1e-9
        assert rel(np.tril(s.A_arrow_tip_block.cpu().numpy()), np.tril(T)) < 1e-9


def _check_selinv(s, Minv, n, b, a):
    Xd, Xl, Xr, Xt = obta.dense_to_bta(Minv, n, b, a)
    assert rel(s.A_diagonal_blocks.cpu().numpy(), Xd) < 1e-9
    if n > 1:
        assert rel(s.A_lower_diagonal_blocks.cpu().numpy(), Xl) < 1e-9
    if a:
        assert rel(s.A_arrow_bottom_blocks.cpu().numpy(), Xr) < 1e-9
        assert rel(s.A_arrow_tip_block.cpu().numpy(), Xt) < 1e-9


@pytest.mark.parametrize("overlap", [0, 2])
def test_factor_logdet_solve_selinv(case, overlap):
    from dalia_amd import _lib
    from dalia_amd.solver import HipBTASolver
    n, b, a, M, S = case
    print(f"case large n={n} b={b} a={a} overlap={overlap}", flush=True)
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    assert s.set_overlap(overlap)
    s.cholesky(S, sparsity="bta")
    info = ctypes.c_int(-1)
    assert _lib.load().dh_get_info(s.h, ctypes.byref(info)) == 0 and info.value == 0
    if overlap == 0:  # with overlap 2 the factor is the same launch; its blocks are checked here
        _check_factor(s, M, n, b, a)
    ld = s.logdet("bta")
    sign, ld_ref = np.linalg.slogdet(M)
    assert sign > 0 and abs(ld - ld_ref) <= 1e-10 * abs(ld_ref)
    rhs = np.random.default_rng(7).standard_normal(n * b + a)
    x = torch.as_tensor(rhs.copy(), device="cuda")
    s.solve(x, sparsity="bta")
    assert rel(x.cpu().numpy(), np.linalg.solve(M, rhs)) < 1e-9
    s.selected_inversion("bta")
    assert _lib.load().dh_get_info(s.h, ctypes.byref(info)) == 0 and info.value == 0
    _check_selinv(s, np.linalg.inv(M), n, b, a)


def test_bt_factor_and_selinv(case):
    """pobtf / pobts / pobtsi on the block-tridiagonal part (the prior's factorization)."""
    from dalia_amd.solver import HipBTASolver
    n, b, a, M, S = case
    print(f"case large bt n={n} b={b}", flush=True)
    Mbt = M[: n * b, : n * b]
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    s.cholesky(S, sparsity="bt")
    _check_factor(s, Mbt, n, b, 0)
    ld = s.logdet("bt")
    assert abs(ld - np.linalg.slogdet(Mbt)[1]) <= 1e-10 * abs(ld)
    s.selected_inversion("bt")
    Xd, Xl, _, _ = obta.dense_to_bta(np.linalg.inv(Mbt), n, b, 0)
    assert rel(s.A_diagonal_blocks.cpu().numpy(), Xd) < 1e-9
    if n > 1:
        assert rel(s.A_lower_diagonal_blocks.cpu().numpy(), Xl) < 1e-9


@pytest.mark.parametrize("alg,syrk", [(0, 0), (0, 1)])
def test_per_panel_schedule_at_2025(alg, syrk):
    """The autotuner's alternative schedule (one launch per 64-column panel) at b = 2025."""
    from dalia_amd import _lib
    from dalia_amd.solver import HipBTASolver
    n, b, a = 3, 2025, 6
    print(f"case per-panel alg={alg} syrk={syrk}", flush=True)
    M, S = random_bta(n, b, a, seed=n * 1000 + b + a)
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    lib = _lib.load()
    _lib.check(lib.dh_set_tuning(s.h, 2, alg), "dh_set_tuning")
    _lib.check(lib.dh_set_tuning(s.h, 0, syrk), "dh_set_tuning")
    s.cholesky(S, sparsity="bta")
    assert s.info() == 0
    ld = s.logdet("bta")
    assert abs(ld - np.linalg.slogdet(M)[1]) <= 1e-10 * abs(ld)
    s.selected_inversion("bta")
    _check_selinv(s, np.linalg.inv(M), n, b, a)


def test_device_coo_input_path():
    """A sparse matrix whose COO arrays live on the device (the cupyx input of DALIA's CuPy
    backend) is scattered in place; CUDA-array right-hand sides are solved in place."""
    from dalia_amd.solver import HipBTASolver
    n, b, a = 4, 832, 6
    print("case device COO input", flush=True)
    M, S = random_bta(n, b, a, seed=21)
    coo = S.tocoo()

    class DevCOO:  # the parts of cupyx.scipy.sparse.coo_matrix the drop-in uses
        def __init__(self):
            self.shape = coo.shape
            self.row = torch.as_tensor(coo.row.astype(np.int32), device="cuda")
            self.col = torch.as_tensor(coo.col.astype(np.int32), device="cuda")
            self.data = torch.as_tensor(coo.data, device="cuda")

        def tocoo(self):
            return self

    class DevVec:  # an array exposing only __cuda_array_interface__ (as CuPy's do)
        def __init__(self, t):
            self.t = t
            self.__cuda_array_interface__ = t.__cuda_array_interface__

    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    s.cholesky(DevCOO(), sparsity="bta")
    ld = s.logdet("bta")
    assert abs(ld - np.linalg.slogdet(M)[1]) <= 1e-10 * abs(ld)
    rhs = np.random.default_rng(5).standard_normal(n * b + a)
    t = torch.as_tensor(rhs.copy(), device="cuda")
    v = DevVec(t)
    assert s.solve(v, "bta") is v
    assert rel(t.cpu().numpy(), np.linalg.solve(M, rhs)) < 1e-9
