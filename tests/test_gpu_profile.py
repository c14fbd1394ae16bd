"""GPU: zero-tile profiles of the sub-diagonal blocks in the persistent factorization.

Sub-diagonal blocks whose rows are zero left of some column (a banded spatial ordering, or uneven
rows, whole empty tile rows) keep those zeros in the factor: L_{i+1,i} = A_{i+1,i} L_ii^-T, and tile
(R, K) of it is A(R, K) minus products with tiles (R, k < K).  The factorization skips the products
with them (`dh_set_tuning` key 9).  Checked against dense LAPACK on the densified matrix (logdet
1e-10 relative, factor / solve / selected-inverse blocks 1e-9 relative Frobenius), one-ended and
two-ended, and against the same factorization with the profiles off, which must agree to the last
bit: the skipped products only ever added zeros.
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import bta as obta
from tests.bta_cases import banded_first, random_bta, ragged_first

pytestmark = pytest.mark.gpu

#: (n, b, a, kind): 32 tile columns with the benchmark's bandwidth (3 grid rows of 45 nodes), 16 with
#: uneven rows, and 13 without arrowhead
CASES = [(4, 2025, 6, "band"), (4, 1024, 4, "ragged"), (6, 832, 0, "ragged"), (3, 2025, 6, "ragged")]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


_CACHE = {}


def _matrix(n, b, a, kind):
    key = (n, b, a, kind)
    if key not in _CACHE:
        _CACHE.clear()
        first = banded_first(b, 3 * 45 + 3) if kind == "band" else ragged_first(b)
        M, S = random_bta(n, b, a, seed=n * 100 + b + a, first=first)
        _CACHE[key] = (M, S, np.linalg.inv(M))
    return _CACHE[key]


def _run(S, n, b, a, twisted, profile):
    from dalia_amd.solver import HipBTASolver
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    s.set_twisted(twisted)
    s.set_tile_profile(profile)
    s.cholesky(S, sparsity="bta")
    assert s.info() == 0
    out = {"D": s.A_diagonal_blocks.cpu().numpy().copy(), "L": s.A_lower_diagonal_blocks.cpu().numpy().copy()}
    if a:
        out["R"] = s.A_arrow_bottom_blocks.cpu().numpy().copy()
        out["T"] = s.A_arrow_tip_block.cpu().numpy().copy()
    out["ld"] = s.logdet("bta")
    rhs = np.random.default_rng(3).standard_normal(n * b + a)
    x = torch.as_tensor(rhs.copy(), device="cuda")
    s.solve(x, sparsity="bta")
    out["x"] = x.cpu().numpy()
    s.selected_inversion("bta")
    out["XD"] = s.A_diagonal_blocks.cpu().numpy().copy()
    out["XL"] = s.A_lower_diagonal_blocks.cpu().numpy().copy()
    if a:
        out["XR"] = s.A_arrow_bottom_blocks.cpu().numpy().copy()
        out["XT"] = s.A_arrow_tip_block.cpu().numpy().copy()
    return out, rhs


@pytest.mark.parametrize("twisted", [False, True])
@pytest.mark.parametrize("n,b,a,kind", CASES, ids=lambda v: str(v))
def test_profile_matches_dense_and_unprofiled(n, b, a, kind, twisted):
    print(f"case profile n={n} b={b} a={a} {kind} twisted={twisted}", flush=True)
    M, S, Minv = _matrix(n, b, a, kind)
    on, rhs = _run(S, n, b, a, twisted, True)
    off, _ = _run(S, n, b, a, twisted, False)
    for k in on:
        assert np.array_equal(np.asarray(on[k]), np.asarray(off[k])), k
    sign, ld_ref = np.linalg.slogdet(M)
    assert sign > 0 and abs(on["ld"] - ld_ref) <= 1e-10 * abs(ld_ref)
    assert rel(on["x"], np.linalg.solve(M, rhs)) < 1e-9
    if not twisted:
        D, L, R, T = obta.dense_to_bta(M, n, b, a)
        obta.pobtaf(D, L, R if a else None, T if a else None)
        for i in range(n):
            assert rel(np.tril(on["D"][i]), np.tril(D[i])) < 1e-9, i
        assert rel(on["L"], L) < 1e-9
        # the zero tiles are exactly zero in the factor
        Z = np.abs(L) == 0.0
        assert np.all(on["L"][Z] == 0.0)
    Xd, Xl, Xr, Xt = obta.dense_to_bta(Minv, n, b, a)
    assert rel(on["XD"], Xd) < 1e-9
    assert rel(on["XL"], Xl) < 1e-9
    if a:
        assert rel(on["XR"], Xr) < 1e-9
        assert rel(on["XT"], Xt) < 1e-9


@pytest.mark.parametrize("by_asm", [False, True])
@pytest.mark.parametrize("twisted", [False, True])
def test_device_profile_of_the_model(twisted, by_asm):
    """The profile the library reads from the assembled Qc of the spatio-temporal model (ns = 2025,
    45 x 45 grid in row-major node order), scanned before the factorization or recorded by the
    assembly (`dh_set_tuning` key 11), equals the one computed here from the oracle's sparse Qc:
    per sub-diagonal block, the first 64-column tile of each tile row with a nonzero entry (the
    two-ended layout stores blocks m .. n-2 transposed).  On this grid it is max(0, R - 3)."""
    from dalia_amd import synth
    from dalia_amd.inla import STInla
    from oracle import inla as oinla
    import bench
    d = synth.make_st_model(45, 6, 6, 1)
    eng = STInla(d, twisted=twisted, profile_from_assembly=by_asm)
    print(f"case device profile twisted={twisted} by_asm={by_asm}", flush=True)
    eng.f(d.theta)
    got = bench.tile_profile(eng)
    Q = oinla.qcond_sparse(d, d.theta).tocsr()
    ns, nt, nbt = d.ns, d.nt, (d.ns + 63) // 64
    m = (nt - 1) // 2
    tw = bool(eng.twisted)
    for s in range(nt - 1):
        B = Q[(s + 1) * ns:(s + 2) * ns, s * ns:(s + 1) * ns]
        if tw and s >= m:
            B = B.T
        B = B.tocoo()
        exp = np.full(nbt, nbt)
        nz = B.data != 0.0
        np.minimum.at(exp, B.row[nz] // 64, B.col[nz] // 64)
        assert np.array_equal(got[s], exp), (s, got[s], exp)
    assert np.array_equal(got[0], np.maximum(0, np.arange(nbt) - 3))


@pytest.mark.parametrize("twisted", [False, True])
def test_profile_from_assembly_bitwise(twisted):
    """The engine with the profiles recorded by the assembly and with the scan: f, the adjoint
    gradient, mu and the latent marginal variances agree to the last bit (the same profiles, first and
    last nonzero tile of each tile row, drive the factorization and the congruence-form selected
    inversion), over two evaluations at different theta."""
    from dalia_amd import synth
    from dalia_amd.inla import STInla
    d = synth.make_st_model(45, 7, 6, 3)
    print(f"case profile from assembly twisted={twisted}", flush=True)
    out = []
    for by_asm in (True, False):
        eng = STInla(d, twisted=twisted, profile_from_assembly=by_asm)
        res = []
        for th in (d.theta, d.theta + np.array([0.1, -0.05, 0.07, 0.2])):
            f, g = eng.value_and_grad(th)
            mu = eng.mu.cpu().numpy().copy()
            v = eng.marginal_variances_latent(th, mu)
            res.append((f, g, mu, v))
        out.append(res)
        del eng
    for (f1, g1, mu1, v1), (f0, g0, mu0, v0) in zip(*out):
        assert f1 == f0
        assert np.array_equal(g1, g0)
        assert np.array_equal(mu1, mu0)
        assert np.array_equal(v1, v0)

