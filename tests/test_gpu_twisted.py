"""GPU: the two-ended factorization (blocks 0 .. m-1 eliminated downward, n-1 .. m+1 upward in one
persistent launch, block m = (n - 1) // 2 last) through the C ABI.

The factor is pinned against dense LAPACK: it is the Cholesky factor of the matrix with its blocks
in the order 0 .. m-1, n-1 .. m+1, m, arrowhead (`_perm_factor_blocks`).  logdet, solve and the
selected inverse are those of the matrix itself, checked against dense slogdet / solve / inv.
Shapes cover n = 3 (one block per side), sides of unequal length, an arrowhead of a full tile,
no arrowhead, and block sizes of 13, 16 and 32 tile columns (every task kind of the schedule).

Bar: logdet within 1e-10 relative; factor, solve and selected-inverse blocks within 1e-9 relative
Frobenius; engine values within 1e-12 (f) and 1e-10 (gradient, mu, variances) of the one-ended
engine.
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import bta as obta
from tests.bta_cases import random_bta

pytestmark = pytest.mark.gpu

SHAPES = [(3, 64, 0), (3, 130, 5), (4, 100, 3), (5, 64, 64), (6, 256, 4), (7, 96, 3), (17, 96, 3), (5, 300, 4),
          (4, 832, 6), (3, 2025, 6), (5, 1000, 0)]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _order(n):
    m = (n - 1) // 2
    return m, list(range(m)) + list(range(n - 1, m, -1)) + [m]


def _perm_factor_blocks(M, n, b, a):
    """Blocks of the Cholesky factor of M with its blocks reordered for the two-ended elimination,
    in the handle's slots: diag[i] and arrow[i] of block i, lower[i] the coupling of blocks i and
    i + 1 with the rows of the one eliminated later."""
    m, order = _order(n)
    idx = np.concatenate([np.arange(j * b, (j + 1) * b) for j in order] + [np.arange(n * b, n * b + a)])
    L = np.linalg.cholesky(M[np.ix_(idx, idx)])
    pos = {j: p for p, j in enumerate(order)}

    def blk(p, q):
        return L[p * b:(p + 1) * b, q * b:(q + 1) * b]

    D = np.stack([blk(pos[j], pos[j]) for j in range(n)])
    Lo = []
    for i in range(n - 1):
        lo, hi = (i, i + 1) if i < m else (i + 1, i)  # (earlier, later) in the elimination
        Lo.append(blk(pos[hi], pos[lo]))
    R = np.stack([L[n * b:, pos[j] * b:(pos[j] + 1) * b] for j in range(n)]) if a else None
    T = L[n * b:, n * b:] if a else None
    return D, np.stack(Lo), R, T


def _check_selinv(s, Minv, n, b, a):
    Xd, Xl, Xr, Xt = obta.dense_to_bta(Minv, n, b, a)
    assert rel(s.A_diagonal_blocks.cpu().numpy(), Xd) < 1e-9
    assert rel(s.A_lower_diagonal_blocks.cpu().numpy(), Xl) < 1e-9
    if a:
        assert rel(s.A_arrow_bottom_blocks.cpu().numpy(), Xr) < 1e-9
        assert rel(s.A_arrow_tip_block.cpu().numpy(), Xt) < 1e-9


@pytest.mark.parametrize("n,b,a", SHAPES)
def test_twisted_factor_logdet_solve_selinv(n, b, a):
    from dalia_amd import _lib
    from dalia_amd.solver import HipBTASolver
    lib = _lib.load()
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    s.set_twisted(True)
    for rep in range(2):  # the handle is reused: flags, tip partials and the launch order reset
        print(f"case twisted n={n} b={b} a={a} rep={rep}", flush=True)
        M, S = random_bta(n, b, a, seed=n * 1000 + b + a + 17 * rep)
        s.cholesky(S, sparsity="bta")
        info = ctypes.c_int(-1)
        assert lib.dh_get_info(s.h, ctypes.byref(info)) == 0 and info.value == 0
        D, Lo, R, T = _perm_factor_blocks(M, n, b, a)
        Dg = s.A_diagonal_blocks.cpu().numpy()
        for i in range(n):
            assert rel(np.tril(Dg[i]), D[i]) < 1e-9, i
        assert rel(s.A_lower_diagonal_blocks.cpu().numpy(), Lo) < 1e-9
        if a:
            assert rel(s.A_arrow_bottom_blocks.cpu().numpy(), R) < 1e-9
            assert rel(np.tril(s.A_arrow_tip_block.cpu().numpy()), T) < 1e-9
        ld = s.logdet("bta")
        ld_ref = np.linalg.slogdet(M)[1]
        assert abs(ld - ld_ref) <= 1e-10 * abs(ld_ref) + 1e-12
        rhs = np.random.default_rng(rep).standard_normal(n * b + a)
        x = torch.as_tensor(rhs.copy(), device="cuda")
        s.solve(x, sparsity="bta")
        assert rel(x.cpu().numpy(), np.linalg.solve(M, rhs)) < 1e-9
        s.selected_inversion("bta")
        assert lib.dh_get_info(s.h, ctypes.byref(info)) == 0 and info.value == 0
        _check_selinv(s, np.linalg.inv(M), n, b, a)


@pytest.mark.parametrize("n,b", [(3, 100), (6, 192), (4, 832)])
def test_twisted_bt_ignores_arrow(n, b):
    from dalia_amd.solver import HipBTASolver
    a = 4
    print(f"case twisted bt n={n} b={b}", flush=True)
    M, S = random_bta(n, b, a, seed=23 + b)
    Mbt = M[: n * b, : n * b]
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    s.set_twisted(True)
    s.cholesky(S, sparsity="bt")
    ld = s.logdet("bt")
    assert abs(ld - np.linalg.slogdet(Mbt)[1]) <= 1e-10 * abs(ld)
    rhs = np.random.default_rng(5).standard_normal(n * b)
    x = torch.as_tensor(rhs.copy(), device="cuda")
    s.solve(x, "bt")
    assert rel(x.cpu().numpy(), np.linalg.solve(Mbt, rhs)) < 1e-9
    s.selected_inversion("bt")
    Xd, Xl, _, _ = obta.dense_to_bta(np.linalg.inv(Mbt), n, b, 0)
    assert rel(s.A_diagonal_blocks.cpu().numpy(), Xd) < 1e-9
    assert rel(s.A_lower_diagonal_blocks.cpu().numpy(), Xl) < 1e-9


def test_twisted_needs_three_blocks():
    from dalia_amd import _lib
    from dalia_amd.solver import HipBTASolver
    s = HipBTASolver(diagonal_blocksize=64, n_diag_blocks=2, arrowhead_blocksize=2)
    with pytest.raises(_lib.DhError):
        s.set_twisted(True)


@pytest.mark.parametrize("shape", [(6, 3, 3, 0), (9, 4, 2, 3), (5, 7, 2, 1), (30, 5, 4, 8)])
def test_twisted_engine_matches_one_ended(shape):
    """f, mu, the adjoint gradient and the latent marginal variances of the engine, two-ended
    against one-ended (the model assembly writes the upward side's coupling blocks transposed)."""
    from dalia_amd import synth
    from dalia_amd.inla import STInla
    print(f"case twisted engine {shape}", flush=True)
    d = synth.make_st_model(*shape)
    th = d.theta + np.array([0.05, -0.1, 0.1, 0.2])
    ref, tw = STInla(d, twisted=False), STInla(d, twisted=True)
    assert not ref.twisted
    assert tw.twisted
    f_ref, f = ref.f(th), tw.f(th)
    assert abs(f - f_ref) <= 1e-12 * abs(f_ref)
    assert rel(tw.mu.cpu().numpy(), ref.mu.cpu().numpy()) < 1e-10
    (fg_ref, g_ref), (fg, g) = ref.value_and_grad(th), tw.value_and_grad(th)
    assert abs(fg - fg_ref) <= 1e-12 * abs(fg_ref)
    assert rel(g, g_ref) < 1e-10, (g, g_ref)
    z = np.zeros(d.n_latent)
    assert rel(tw.marginal_variances_latent(th, z), ref.marginal_variances_latent(th, z)) < 1e-10


def test_twisted_assembly_then_one_ended_factor():
    """A handle assembled for a two-ended factorization and then switched back (or factored in a
    multi-handle launch) is returned to the one-ended layout first."""
    from dalia_amd import _lib, synth
    from dalia_amd.inla import STInla
    from oracle import inla as oinla
    d = synth.make_st_model(6, 5, 3, 4)
    eng = STInla(d, twisted=True)
    lib = eng.lib
    f_ref = oinla.f_theta(d, d.theta)
    assert abs(eng.f(d.theta) - f_ref) <= 1e-10 * abs(f_ref)
    eng._set_params(d.theta)
    eng._assemble(None, "bta", True)          # written for the two-ended factorization
    eng.solver.set_twisted(False)
    _lib.check(lib.dh_cholesky(eng.h, 1), "dh_cholesky")
    ld = eng.solver.logdet("bta")
    eng.solver.set_twisted(True)
    eng._assemble(None, "bta", True)
    _lib.check(lib.dh_cholesky(eng.h, 1), "dh_cholesky")
    assert abs(eng.solver.logdet("bta") - ld) <= 1e-12 * abs(ld)
    eng._assemble(None, "bta", True)
    mc = ctypes.c_void_p()
    arr = (ctypes.c_void_p * 1)(eng.h.value)
    _lib.check(lib.dh_multi_create(arr, 1, ctypes.byref(mc)), "dh_multi_create")
    try:
        _lib.check(lib.dh_multi_cholesky(mc, 1, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                   "dh_multi_cholesky")
        assert abs(eng.solver.logdet("bta") - ld) <= 1e-12 * abs(ld)
    finally:
        lib.dh_multi_destroy(mc)
