"""GPU: graph replay, batched uneven-size evaluations, FD Hessians, kernel variants, autotuner."""

import numpy as np
import pytest
import torch

from dalia_amd import synth
from oracle import bta as obta
from oracle import inla as oinla
from tests.bta_cases import random_bta

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


THETAS = [np.array([-1.0, 0.7, 0.0, 1.5]), np.array([-0.8, 0.5, 0.2, 1.2]), np.array([-1.2, 0.9, -0.1, 1.8])]


@pytest.mark.parametrize("overlap", [0, 2])
def test_graph_replay_matches_eager_every_replay(overlap):
    """overlap 2: the eager evaluations use the second stream, the captured one stays on one stream."""
    from dalia_amd.inla import STInla
    d = synth.make_st_model(7, 5, 3, 2)
    print(f"case graph ns={d.ns} nt={d.nt} overlap={overlap}", flush=True)
    eng = STInla(d, overlap=overlap)
    ref = [eng.f(t) for t in THETAS]
    eng.capture_f()
    for rep in range(3):
        for t, r in zip(THETAS, ref):
            fg = eng.f_graph(t)
            assert abs(fg - r) <= 1e-12 * abs(r)
            eng.f(THETAS[(rep + 1) % 3])  # eager work between replays
            torch.cuda.synchronize()
    f_or = oinla.f_theta(d, THETAS[1])
    assert abs(eng.f_graph(THETAS[1]) - f_or) <= 1e-10 * abs(f_or)


def test_batched_uneven_instances_one_graph():
    from dalia_amd.batch import BatchedF
    from dalia_amd.inla import STInla
    shapes = [(5, 3, 2, 0), (8, 6, 3, 1), (6, 4, 4, 2)]
    datas = [synth.make_st_model(*s) for s in shapes]
    engines = [STInla(d) for d in datas]
    print("case batch shapes", [(d.ns, d.nt, d.nb) for d in datas], flush=True)
    assignment = [0, 1, 2, 1, 0, 2, 2, 1]
    bf = BatchedF(engines, assignment)
    rng = np.random.default_rng(0)
    for rep in range(2):
        thetas = [THETAS[0] + 0.1 * rng.standard_normal(4) for _ in assignment]
        fv = bf(thetas)
        for e, inst in enumerate(assignment):
            ref = oinla.f_theta(datas[inst], thetas[e])
            assert abs(fv[e] - ref) <= 1e-10 * abs(ref), (rep, e)


def test_batched_hessians_match_oracle():
    from dalia_amd.batch import batched_hessians
    from dalia_amd.inla import STInla
    datas = [synth.make_st_model(5, 3, 2, 0), synth.make_st_model(6, 4, 2, 1)]
    engines = [STInla(d) for d in datas]
    thetas = [d.theta for d in datas]
    H, _ = batched_hessians(engines, thetas, h=5e-3, offdiag_points=2)
    for d, t, Hg in zip(datas, thetas, H):
        Hr = oinla.fd_hessian(d, t, 5e-3, offdiag_points=2)
        assert np.abs(Hg - Hr).max() <= 1e-6 * np.abs(Hr).max()
        Hs = engines[0].__class__  # noqa: F841
    # the reference's 4-point stencil through the plain engine
    H4 = engines[0].hessian(thetas[0], h=5e-3, offdiag_points=4)
    H4r = oinla.fd_hessian(datas[0], thetas[0], 5e-3, offdiag_points=4)
    assert np.abs(H4 - H4r).max() <= 1e-6 * np.abs(H4r).max()


@pytest.mark.parametrize("alg,syrk", [(1, 0), (0, 1), (0, 0)])
@pytest.mark.parametrize("n,b,a", [(3, 256, 4), (4, 128, 0), (2, 200, 3), (5, 64, 64)])
def test_kernel_variants_parity(alg, syrk, n, b, a):
    from dalia_amd.solver import HipBTASolver
    print(f"case variants alg={alg} syrk={syrk} n={n} b={b} a={a}", flush=True)
    M, S = random_bta(n, b, a, seed=123 + b)
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    s.lib.dh_set_tuning(s.h, 2, alg)
    s.lib.dh_set_tuning(s.h, 0, syrk)
    s.cholesky(S, "bta")
    assert abs(s.logdet("bta") - np.linalg.slogdet(M)[1]) <= 1e-10 * abs(np.linalg.slogdet(M)[1])
    s.selected_inversion("bta")
    Xd, Xl, Xr, Xt = obta.dense_to_bta(np.linalg.inv(M), n, b, a)
    assert rel(s.A_diagonal_blocks.cpu().numpy(), Xd) < 1e-9
    assert rel(s.A_lower_diagonal_blocks.cpu().numpy(), Xl) < 1e-9
    if a:
        assert rel(s.A_arrow_bottom_blocks.cpu().numpy(), Xr) < 1e-9


def test_autotune_choice_applied(tmp_path):
    from dalia_amd import autotune
    from dalia_amd.inla import STInla
    d = synth.make_st_model(12, 4, 3, 0)  # ns = 144 -> bp = 192 (odd tile count), and b=256 below
    eng = STInla(d)
    path = str(tmp_path / "tune.json")
    c = autotune.tune(eng, path=path, reps=1)
    assert c["alg"] in (0, 1) and c["syrk"] in (0, 1)
    c2 = autotune.tune(eng, path=path)  # from the file
    assert c2["syrk"] == c["syrk"]
    f_ref = oinla.f_theta(d, d.theta)
    assert abs(eng.f(d.theta) - f_ref) <= 1e-10 * abs(f_ref)
    f, g = eng.value_and_grad(d.theta)
    _, g_ref = oinla.adjoint_gradient(d, d.theta)
    assert rel(g, g_ref) < 1e-8


def test_batched_hessian_diagonals_and_covariance():
    """The diagonal 3-point stencil (`utils/finite_difference_stencils.py:81-120`) in one graph over
    a univariate and a coregional instance, and the covariance of theta (`core/dalia.py:747-765`)."""
    from dalia_amd.batch import batched_hessian_diags
    from dalia_amd.coreg import CoregInla
    from dalia_amd.inla import STInla, compute_covariance_hp
    from oracle import coreg as ocoreg
    du = synth.make_st_model(5, 3, 2, 0)
    dc = synth.make_coreg_model(4, 3, 2, 2, 9)
    engines = [STInla(du), CoregInla(dc)]
    thetas = [du.theta, dc.theta]
    print("case batched diagonal Hessians", flush=True)
    diags, bf = batched_hessian_diags(engines, thetas, h=5e-3, points=3)
    assert len(bf.assignment) == (1 + 2 * 4) + (1 + 2 * 9)
    funs = [lambda t: oinla.f_theta(du, t), lambda t: ocoreg.f_theta_coreg(dc, t)]
    for dg, t, fun in zip(diags, thetas, funs):
        f0 = fun(t)
        ref = np.array([(fun(t + 5e-3 * e) - 2 * f0 + fun(t - 5e-3 * e)) / 5e-3 ** 2 for e in np.eye(t.size)])
        assert np.abs(dg - ref).max() <= 1e-5 * np.abs(ref).max()
    C = compute_covariance_hp(engines[0], du.theta)
    Cr = np.linalg.inv(oinla.fd_hessian(du, du.theta, 5e-3, offdiag_points=4))
    assert np.abs(C - Cr).max() <= 1e-5 * np.abs(Cr).max()


def test_captured_adjoint_gradient_every_replay():
    """The whole adjoint evaluation captured once and replayed for new theta, every replay against
    the eager evaluation (eager work between the replays)."""
    from dalia_amd.inla import STInla
    d = synth.make_st_model(7, 6, 3, 4)
    print(f"case captured gradient ns={d.ns} nt={d.nt}", flush=True)
    eng = STInla(d)
    ref = [eng.value_and_grad(t) for t in THETAS]
    eng.capture_grad()
    for rep in range(3):
        for t, (fr, gr) in zip(THETAS, ref):
            f, g = eng.value_and_grad_graph(t)
            assert abs(f - fr) <= 1e-12 * abs(fr)
            assert rel(g, gr) < 1e-10
            eng.f(THETAS[(rep + 1) % 3])  # eager work between replays
            torch.cuda.synchronize()
    f_or, g_or = oinla.adjoint_gradient(d, THETAS[2])
    f, g = eng.value_and_grad_graph(THETAS[2])
    assert abs(f - f_or) <= 1e-10 * abs(f_or) and rel(g, g_or) < 1e-8


@pytest.mark.parametrize("assignment", [[0, 1, 2, 1, 0, 2, 2, 1], [0, 0, 0, 1, 2]])
def test_fused_rounds_one_launch_per_round(assignment):
    """Uneven instances (univariate and coregional) whose factorizations of each round run as one
    persistent launch over all instances (`dh_multi_cholesky`); rounds with fewer instances when
    their evaluation counts differ.  Every replay against the oracle."""
    from dalia_amd.batch import BatchedF
    from dalia_amd.coreg import CoregInla
    from dalia_amd.inla import STInla
    from oracle import coreg as ocoreg
    du = [synth.make_st_model(5, 3, 2, 0), synth.make_st_model(9, 6, 3, 1)]
    dc = synth.make_coreg_model(5, 4, 2, 2, 6)
    engines = [STInla(du[0]), CoregInla(dc), STInla(du[1])]
    datas = [du[0], dc, du[1]]
    print("case fused rounds", assignment, flush=True)
    bf = BatchedF(engines, assignment, fused=True)
    rng = np.random.default_rng(2)
    for rep in range(2):
        thetas = [datas[a].theta + 0.05 * rng.standard_normal(datas[a].theta.size) for a in assignment]
        fv = bf(thetas)
        for e, a in enumerate(assignment):
            ref = ocoreg.f_theta_coreg(dc, thetas[e]) if a == 1 else oinla.f_theta(datas[a], thetas[e])
            assert abs(fv[e] - ref) <= 1e-10 * abs(ref), (rep, e)
        for eng, d in zip(engines, datas):  # the engines still evaluate eagerly afterwards
            ref = ocoreg.f_theta_coreg(d, d.theta) if eng is engines[1] else oinla.f_theta(d, d.theta)
            assert abs(eng.f(d.theta) - ref) <= 1e-10 * abs(ref)
