"""GPU: the selected inversion in congruence form on the matrix's coupling tiles (`dh_set_tuning` key 10).

X_ii = inv(L_ii)^T (I + M) inv(L_ii), M = L^T Y + L_a^T Ya, Y = X_qq L + X_aq^T L_a, Ya = X_aq L + X_aa L_a,
X_{a,i} = -Ya inv(L_ii), X_{q,i} = -Y inv(L_ii) on the tiles of each coupling block between the first and
last nonzero tile of a tile row.  Checked against the dense inverse of the densified matrix (1e-9 relative
Frobenius): every diagonal block, the arrow and the tip in full, and the sub-diagonal blocks on every 64x64
tile where the matrix's coupling block has an entry (what callers reading the selected inverse on the
matrix's pattern use); one- and two-ended, with and without arrowhead, banded and ragged couplings; then
the spatio-temporal engine's f, gradient and marginal variances against the same engine with whole blocks.
The diagonal blocks must come out exactly symmetric: a rounding-level antisymmetric part in a diagonal
tile is turned by the next block's congruence into an error scaled by cond(L_ii)^2 that compounds over
the recursion (at ns = 4096 it grew about 2.7x per block); the long one-ended recursion at ns = 4096
checks that it stays at rounding level.
"""

import numpy as np
import pytest
import torch

from oracle import bta as obta
from tests.bta_cases import banded_first, random_bta, ragged_first

pytestmark = pytest.mark.gpu

CASES = [(5, 300, 4, "band"), (4, 1024, 4, "ragged"), (3, 2025, 6, "band"), (6, 832, 0, "band"), (1, 200, 3, "band")]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


_CACHE = {}


def _matrix(n, b, a, kind):
    key = (n, b, a, kind)
    if key not in _CACHE:
        _CACHE.clear()
        first = banded_first(b, 138) if kind == "band" else ragged_first(b)
        M, S = random_bta(n, b, a, seed=7 * n + b + a, first=first)
        _CACHE[key] = (M, S, np.linalg.inv(M))
    return _CACHE[key]


@pytest.mark.parametrize("twisted", [False, True])
@pytest.mark.parametrize("n,b,a,kind", CASES, ids=lambda v: str(v))
def test_pattern_selinv_matches_dense_inverse(n, b, a, kind, twisted):
    from dalia_amd.solver import HipBTASolver
    if twisted and n < 3:
        pytest.skip("two-ended needs n >= 3")
    print(f"case selinv pattern n={n} b={b} a={a} {kind} twisted={twisted}", flush=True)
    M, S, Minv = _matrix(n, b, a, kind)
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    s.set_twisted(twisted)
    s.set_selected_inversion_pattern(True)
    s.cholesky(S, sparsity="bta")
    assert s.info() == 0
    s.selected_inversion("bta")
    Xd, Xl, Xr, Xt = obta.dense_to_bta(Minv, n, b, a)
    Dg = s.A_diagonal_blocks.cpu().numpy()
    assert rel(Dg, Xd) < 1e-9
    assert np.array_equal(Dg, np.swapaxes(Dg, 1, 2))  # stored exactly symmetric, diagonal tiles included
    if a:
        assert rel(s.A_arrow_bottom_blocks.cpu().numpy(), Xr) < 1e-9
        assert rel(s.A_arrow_tip_block.cpu().numpy(), Xt) < 1e-9
    if n > 1:
        got = s.A_lower_diagonal_blocks.cpu().numpy()
        num = den = 0.0
        for i in range(n - 1):
            A = M[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b]
            for R in range(0, b, 64):
                for C in range(0, b, 64):
                    if np.any(A[R:R + 64, C:C + 64] != 0.0):
                        num += np.sum((got[i, R:R + 64, C:C + 64] - Xl[i, R:R + 64, C:C + 64]) ** 2)
                        den += np.sum(Xl[i, R:R + 64, C:C + 64] ** 2)
        assert np.sqrt(num / den) < 1e-9


@pytest.mark.parametrize("twisted", [True, False])
def test_engine_pattern_selinv_matches_whole_blocks(twisted):
    from dalia_amd import synth
    from dalia_amd.inla import STInla
    d = synth.make_st_model(24, 8, 3, 5)  # ns = 576 (9 tile columns), nt = 8
    print(f"case engine selinv pattern twisted={twisted}", flush=True)
    e1 = STInla(d, twisted=twisted, pattern_selinv=True)
    e0 = STInla(d, twisted=twisted, pattern_selinv=False)
    th = d.theta + np.array([0.05, -0.03, 0.02, 0.1])
    f1, g1 = e1.value_and_grad(th)
    f0, g0 = e0.value_and_grad(th)
    assert abs(f1 - f0) <= 1e-12 * abs(f0)
    assert np.linalg.norm(g1 - g0) <= 1e-10 * np.linalg.norm(g0)
    x1, x0 = e1.mu.cpu().numpy().copy(), e0.mu.cpu().numpy().copy()
    v1 = e1.marginal_variances_latent(th, x1)
    v0 = e0.marginal_variances_latent(th, x0)
    assert rel(v1, v0) < 1e-10
    o1 = e1.marginal_variances_observations(th, x1)
    o0 = e0.marginal_variances_observations(th, x0)
    assert rel(o1, o0) < 1e-10


def test_engine_pattern_selinv_long_recursion():
    """ns = 4096 (64 tile columns), nt = 24, one-ended: 23 congruence steps in a row on an
    ill-conditioned spatial block; the marginal variances and the gradient against whole blocks."""
    from dalia_amd import synth
    from dalia_amd.inla import STInla
    d = synth.make_st_model(64, 24, 6, 1)
    print("case engine selinv pattern ns=4096 nt=24 one-ended", flush=True)
    out = []
    for pat in (True, False):
        e = STInla(d, twisted=False, pattern_selinv=pat)
        f, g = e.value_and_grad(d.theta)
        x = e.mu.cpu().numpy().copy()
        out.append((f, g, e.marginal_variances_latent(d.theta, x)))
        del e
        torch.cuda.empty_cache()
    (f1, g1, v1), (f0, g0, v0) = out
    assert abs(f1 - f0) <= 1e-12 * abs(f0)
    assert np.linalg.norm(g1 - g0) <= 1e-10 * np.linalg.norm(g0)
    assert np.max(np.abs(v1 - v0) / np.abs(v0)) < 1e-11
