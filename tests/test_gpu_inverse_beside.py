"""GPU: the inverse of the diagonal factors formed on a second stream beside the solve's forward
sweep (`dh_set_tuning` key 12), and the selected inversion beside the solve's backward sweep (key 13).

The first solve after a factorization launches the inverse in chunks of up to 16 blocks, ordered as
the forward sweep reaches them (from both ends inward for a two-ended factor), and each sweep step
waits for the chunk that holds its block.  The kernels and blocks are those of the serial inverse,
so the solution, a second solve on the same factor and the selected inverse that follows must equal
the serial path's bit for bit; the solution is also checked against the dense solve.  With key 13
the recursion runs on the second stream, each of its steps after the backward-sweep step that reads
the blocks it overwrites: the selected inverse must again be bitwise the serial one, also when a
logdet between the solve and the recursion cancels the arrangement.  The shapes
give one chunk, chunks that meet in the middle with a short last chunk (37 blocks two-ended: blocks
0-15, 21-36, 16-20) and a one-ended factor with a short last chunk; then the spatio-temporal engine's
f and gradient with and without it.
"""

import numpy as np
import pytest
import torch

from tests.bta_cases import random_bta

pytestmark = pytest.mark.gpu

SHAPES = [(5, 130, 4, True), (20, 200, 2, True), (37, 64, 3, True), (37, 64, 0, False), (40, 96, 5, False)]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _run(n, b, a, twisted, beside, M, S, rhs, sel=False, logdet_between=False):
    from dalia_amd.solver import HipBTASolver
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    s.set_twisted(twisted)
    s.set_inverse_beside_solve(beside)
    s.set_selinv_beside_solve(sel)
    s.set_selected_inversion_pattern(True)
    s.cholesky(S, sparsity="bta")
    x1 = torch.as_tensor(rhs[0].copy(), device="cuda")
    s.solve(x1, sparsity="bta")
    x2 = torch.as_tensor(rhs[1].copy(), device="cuda")
    s.solve(x2, sparsity="bta")  # the inverse is formed: the plain path
    ld = s.logdet("bta") if logdet_between else None
    s.selected_inversion("bta")  # with sel: beside the second solve's backward sweep
    out = [x1.cpu().numpy(), x2.cpu().numpy(), s.A_diagonal_blocks.cpu().numpy(),
           s.A_lower_diagonal_blocks.cpu().numpy()]
    if ld is not None:
        out.append(np.array([ld]))
    if a:
        out += [s.A_arrow_bottom_blocks.cpu().numpy(), s.A_arrow_tip_block.cpu().numpy()]
    return out


@pytest.mark.parametrize("n,b,a,twisted", SHAPES, ids=lambda v: str(v))
def test_inverse_beside_solve_bitwise(n, b, a, twisted):
    print(f"case inverse beside solve n={n} b={b} a={a} twisted={twisted}", flush=True)
    M, S = random_bta(n, b, a, seed=31 * n + b + a)
    rng = np.random.default_rng(n + b)
    rhs = [rng.standard_normal(n * b + a) for _ in range(2)]
    ref = _run(n, b, a, twisted, False, M, S, rhs)
    for beside, sel in ((True, False), (False, True), (True, True)):
        got = _run(n, b, a, twisted, beside, M, S, rhs, sel)
        for k, (g, r) in enumerate(zip(got, ref)):
            assert np.array_equal(g, r), (beside, sel, k)
    assert rel(got[0], np.linalg.solve(M, rhs[0])) < 1e-9
    assert rel(got[1], np.linalg.solve(M, rhs[1])) < 1e-9
    ref_ld = _run(n, b, a, twisted, False, M, S, rhs, False, True)
    got_ld = _run(n, b, a, twisted, True, M, S, rhs, True, True)  # the logdet cancels key 13
    for k, (g, r) in enumerate(zip(got_ld, ref_ld)):
        assert np.array_equal(g, r), ("logdet between", k)


@pytest.mark.parametrize("twisted", [True, False])
def test_engine_inverse_beside_solve(twisted):
    from dalia_amd import synth
    from dalia_amd.inla import STInla
    d = synth.make_st_model(16, 40, 3, 5)  # ns = 256, nt = 40: three chunks of the inverse
    print(f"case engine inverse beside solve twisted={twisted}", flush=True)
    th = d.theta + np.array([0.04, -0.02, 0.03, 0.1])
    res = []
    for beside in (True, False):
        e = STInla(d, twisted=twisted, inverse_beside_solve=beside, selinv_beside_solve=beside)
        f, g = e.value_and_grad(th)
        f2, g2 = e.value_and_grad(th)  # a second evaluation on the same handle
        x = e.mu.cpu().numpy().copy()
        res.append((f, g, x, e.f(th), f2, g2, e.marginal_variances_latent(th, x)))
    for k, (u, v) in enumerate(zip(*res)):
        assert np.array_equal(np.asarray(u), np.asarray(v)), k
