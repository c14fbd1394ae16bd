"""GPU parity of assembly, f(theta) and the adjoint gradient against the CPU oracle."""

import numpy as np
import pytest
import torch

from dalia_amd import synth
from oracle import inla as oinla

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module", params=[(6, 5, 3, 0), (9, 4, 2, 3)])
def small(request):
    m, nt, nb, seed = request.param
    return synth.make_st_model(m, nt, nb, seed)


def test_assembly_matches_oracle(small):
    from dalia_amd.inla import STInla
    d = small
    print(f"case assembly ns={d.ns} nt={d.nt} nb={d.nb}", flush=True)
    eng = STInla(d)
    th = d.theta
    Qc = oinla.qcond_sparse(d, th)
    D, L, R, T = oinla.sparse_to_bta(Qc, d.nt, d.ns, d.nb, "bta")
    eng._assemble(th, "bta", True)
    s = eng.solver
    assert rel(s.A_diagonal_blocks.cpu().numpy(), D) < 1e-14
    assert rel(s.A_lower_diagonal_blocks.cpu().numpy(), L) < 1e-14
    assert rel(s.A_arrow_bottom_blocks.cpu().numpy(), R) < 1e-14
    assert rel(s.A_arrow_tip_block.cpu().numpy(), T) < 1e-14
    Qp = oinla.qprior_sparse(d, th)
    D, L, _, _ = oinla.sparse_to_bta(Qp, d.nt, d.ns, d.nb, "bt")
    eng._assemble(th, "bt", False)
    assert rel(s.A_diagonal_blocks.cpu().numpy(), D) < 1e-14
    assert rel(s.A_lower_diagonal_blocks.cpu().numpy(), L) < 1e-14


def test_f_theta_matches_oracle(small):
    from dalia_amd.inla import STInla
    d = small
    print(f"case f ns={d.ns} nt={d.nt}", flush=True)
    eng = STInla(d)
    for th in (d.theta, d.theta + np.array([0.1, -0.2, 0.05, 0.3])):
        f_ref, parts = oinla.f_theta(d, th, return_parts=True)
        f = eng.f(th)
        assert abs(f - f_ref) <= 1e-10 * abs(f_ref)
        assert rel(eng.mu.cpu().numpy(), parts["mu"]) < 1e-9


@pytest.mark.parametrize("overlap", [0, 2])
def test_adjoint_gradient(small, overlap):
    """overlap 2: the inverse and the products of the factor formed beside the factorization."""
    from dalia_amd.inla import STInla
    d = small
    print(f"case adjoint ns={d.ns} nt={d.nt} overlap={overlap}", flush=True)
    eng = STInla(d, overlap=overlap)
    th = d.theta
    f_ref, g_ref = oinla.adjoint_gradient(d, th)
    f, g = eng.value_and_grad(th)
    assert abs(f - f_ref) <= 1e-10 * abs(f_ref)
    assert rel(g, g_ref) < 1e-8
    f0, gfd = eng.fd_gradient(th, h=1e-4)
    assert rel(g, gfd) < 1e-5


@pytest.mark.parametrize("overlap", [0, 2])
def test_adjoint_without_closed_form(small, overlap):
    """General FEM inputs: prior term by pobtf + pobtsi instead of the eigen form."""
    from dalia_amd.inla import STInla
    d = small
    print(f"case adjoint-bt ns={d.ns} nt={d.nt} overlap={overlap}", flush=True)
    eng = STInla(d, closed_form_prior=False, overlap=overlap)
    f_ref, g_ref = oinla.adjoint_gradient(d, d.theta)
    f, g = eng.value_and_grad(d.theta)
    assert abs(f - f_ref) <= 1e-10 * abs(f_ref)
    assert rel(g, g_ref) < 1e-8


def test_autograd_function(small):
    from dalia_amd.inla import FTheta, STInla
    d = small
    eng = STInla(d)
    th = torch.tensor(d.theta, dtype=torch.float64, requires_grad=True)
    f = FTheta.apply(th, eng)
    f.backward()
    f_ref, g_ref = oinla.adjoint_gradient(d, d.theta)
    assert abs(float(f) - f_ref) <= 1e-10 * abs(f_ref)
    assert rel(th.grad.numpy(), g_ref) < 1e-8


@pytest.mark.parametrize("shape", [(5, 4, 2, 0), (6, 5, 3, 1), (4, 3, 0, 2)])
def test_marginal_variances(shape):
    """diag(Qc^-1) and diag(A Qc^-1 A^T) through the HIP selected inversion vs the oracle."""
    from dalia_amd.inla import STInla
    d = synth.make_st_model(*shape)
    print(f"case marginal variances ns={d.ns} nt={d.nt} nb={d.nb}", flush=True)
    eng = STInla(d)
    theta = d.theta + np.array([0.05, -0.1, 0.1, 0.2])
    mv = eng.marginal_variances_latent(theta, np.zeros(d.ns * d.nt + d.nb))
    ref = oinla.marginal_variances_latent(d, theta)
    assert np.abs(mv - ref).max() <= 1e-9 * np.abs(ref).max()
    ov = eng.marginal_variances_observations(theta, np.zeros(d.ns * d.nt + d.nb))
    oref = oinla.marginal_variances_observations(d, theta)
    assert np.abs(ov - oref).max() <= 1e-9 * np.abs(oref).max()
    eng.f(theta)
    mv2 = eng.marginal_variances_latent()  # stored theta
    assert np.abs(mv2 - ref).max() <= 1e-9 * np.abs(ref).max()
    with pytest.raises(ValueError):
        eng.marginal_variances_latent(theta, None)


def test_hessian_from_adjoint_gradients_matches_fd_hessian():
    from dalia_amd.inla import STInla, hessian_adjoint
    d = synth.make_st_model(5, 4, 2, 21)
    print("case hessian_adjoint st", flush=True)
    eng = STInla(d)
    H = hessian_adjoint(eng, d.theta, h=1e-3)
    H_ref = oinla.fd_hessian(d, d.theta, 1e-3)
    assert np.abs(H - H_ref).max() <= 1e-4 * np.abs(H_ref).max(), (H, H_ref)


def test_sphere_manifold_f_and_adjoint():
    """The sphere branch of the hyperparameter map (`spatio_temporal.py:171-180`) through the HIP
    path: f and the adjoint gradient against the oracle."""
    import dataclasses

    from dalia_amd.inla import STInla
    d = dataclasses.replace(synth.make_st_model(7, 5, 2, 4), manifold="sphere")
    print("case sphere manifold ns=49 nt=5", flush=True)
    eng = STInla(d)
    th = d.theta + np.array([-0.2, 0.1, 0.05, 0.1])
    f_ref, g_ref = oinla.adjoint_gradient(d, th)
    assert abs(eng.f(th) - f_ref) <= 1e-10 * abs(f_ref)
    f, g = eng.value_and_grad(th)
    assert abs(f - f_ref) <= 1e-10 * abs(f_ref)
    assert rel(g, g_ref) < 1e-8


def test_loaded_input_tree_through_the_hip_path(tmp_path):
    """A model written in DALIA's on-disk layout (`core/model.py:239-252`,
    `submodels/spatio_temporal.py:27-35`) and read back by `io.load_st_model` runs through the HIP
    path: f, the adjoint gradient and the marginal variances against the oracle on the loaded data;
    the coregional tree likewise for f."""
    from dalia_amd import io
    from dalia_amd.coreg import CoregInla
    from dalia_amd.inla import STInla
    from oracle import coreg as ocoreg
    d = synth.make_st_model(9, 6, 3, 5)
    io.save_st_model(d, str(tmp_path / "st"))
    e = io.load_st_model(str(tmp_path / "st"), d.theta, d.priors)
    print(f"case loaded tree ns={e.ns} nt={e.nt} nb={e.nb}", flush=True)
    eng = STInla(e)
    f_ref, g_ref = oinla.adjoint_gradient(e, e.theta)
    assert abs(eng.f(e.theta) - f_ref) <= 1e-10 * abs(f_ref)
    f, g = eng.value_and_grad(e.theta)
    assert rel(g, g_ref) < 1e-8
    mv = eng.marginal_variances_latent(e.theta, np.zeros(e.n_latent))
    assert rel(mv, oinla.marginal_variances_latent(e, e.theta)) < 1e-9
    dc = synth.make_coreg_model(5, 4, 2, 2, 7)
    io.save_coreg_model(dc, str(tmp_path / "coreg"))
    ec = io.load_coreg_model(str(tmp_path / "coreg"), 2, dc.theta, dc.priors)
    fc_ref = ocoreg.f_theta_coreg(ec, ec.theta)
    assert abs(CoregInla(ec).f(ec.theta) - fc_ref) <= 1e-10 * abs(fc_ref)
