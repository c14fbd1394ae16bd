"""Pin the CPU oracle's model functions by mathematics that fixes them."""

import math

import numpy as np
import pytest
import scipy.integrate as si
import scipy.sparse as sps

from dalia_amd import synth
from oracle import inla


@pytest.fixture(scope="module")
def tiny():
    return synth.make_st_model(5, 4, 2, 0)


def test_marginal_likelihood_identity(tiny):
    """For a Gaussian likelihood the Laplace approximation is exact:
    f + log pi(theta) = -log N(y; 0, A Qp^-1 A^T + e^-tau I) + 1/2 nb log(fe_prec) - n/2 log 2pi."""
    d = tiny
    for th in (d.theta, d.theta + np.array([0.2, -0.1, 0.3, -0.4])):
        f = inla.f_theta(d, th)
        Qp = inla.qprior_sparse(d, th).toarray()
        A = d.a_full().toarray()
        Sig = A @ np.linalg.solve(Qp, A.T) + math.exp(-th[3]) * np.eye(d.n_obs)
        ld = np.linalg.slogdet(Sig)[1]
        logpy = -0.5 * ld - 0.5 * d.y @ np.linalg.solve(Sig, d.y) - 0.5 * d.n_obs * math.log(2 * math.pi)
        const = 0.5 * d.nb * math.log(d.fe_prec) - 0.5 * d.n_obs * math.log(2 * math.pi)
        lhs = f + inla.log_prior(th, d.priors)
        assert abs(lhs - (-logpy + const)) < 1e-8 * max(1.0, abs(lhs))


def test_assembly_equals_dense_kron(tiny):
    d = tiny
    th = d.theta
    gs, gt, gst = inla.interpretable2compute(*th[:3])
    k = math.exp(gs)
    C0, G1, G2, G3 = (x.toarray() for x in (d.c0, d.g1, d.g2, d.g3))
    q1 = k ** 2 * C0 + G1
    q2 = k ** 4 * C0 + 2 * k ** 2 * G1 + G2
    q3 = k ** 6 * C0 + 3 * k ** 4 * G1 + 3 * k ** 2 * G2 + G3
    M0, M1, M2 = (x.toarray() for x in (d.m0, d.m1, d.m2))
    Q = math.exp(2 * gst) * (np.kron(M0, q3) + math.exp(gt) * np.kron(M1, q2) + math.exp(2 * gt) * np.kron(M2, q1))
    assert np.allclose(inla.qprior_st_sparse(d, th).toarray(), Q, rtol=1e-13, atol=1e-13 * abs(Q).max())
    # kron coefficients reproduce the same matrix
    c = inla.kron_coefficients(th)
    Q2 = sum(c[kk, m] * np.kron([M0, M1, M2][kk], [C0, G1, G2, G3][m]) for kk in range(3) for m in range(4))
    assert np.allclose(Q2, Q, rtol=1e-13, atol=1e-13 * abs(Q).max())
    # time-major layout gives a block-tridiagonal pattern
    n, b = d.nt, d.ns
    for i in range(n):
        for j in range(n):
            if abs(i - j) > 1:
                assert not np.any(Q[i * b:(i + 1) * b, j * b:(j + 1) * b])


def test_fem_identities(tiny):
    d = tiny
    ci = sps.diags(1.0 / d.c0.diagonal())
    assert abs(d.g2 - d.g1 @ ci @ d.g1).max() < 1e-12 * abs(d.g2).max()
    assert abs(d.g3 - d.g1 @ ci @ d.g2).max() < 1e-12 * abs(d.g3).max()
    # P1 stiffness rows sum to zero, lumped mass sums to the domain area
    assert abs(d.g1 @ np.ones(d.ns)).max() < 1e-12
    assert abs(d.c0.diagonal().sum() - 1.0) < 1e-12
    # stiffness reproduces the Dirichlet energy of a linear function exactly
    m = int(round(math.sqrt(d.ns)))
    pts, _ = synth.grid_triangles(m)
    u = 2.0 * pts[:, 0] - 3.0 * pts[:, 1]
    assert abs(u @ (d.g1 @ u) - 13.0) < 1e-12


@pytest.mark.parametrize("kind,alpha,u", [("r_s", 0.01, 0.5), ("r_t", 0.01, 5.0), ("sigma_st", 0.01, 3.0),
                                         ("prec_o", 0.05, 4.0)])
def test_pc_prior_normalised_and_tail(kind, alpha, u):
    lam = inla.pc_lambda(kind, alpha, u)
    dens = lambda t: math.exp(inla.pc_log_prior(kind, lam, t))
    total, _ = si.quad(dens, -40, 90, limit=800)
    assert abs(total - 1.0) < 1e-8
    if kind in ("r_s", "r_t"):
        tail, _ = si.quad(dens, -40, math.log(u), limit=400)   # P(range < u) = alpha
    else:
        tail, _ = si.quad(dens, math.log(u), 90, limit=800)    # P(sigma > u) = alpha
    assert abs(tail - alpha) < 1e-8
    # gradient
    h = 1e-6
    for t in (-1.0, 0.3, 2.0):
        g = inla.pc_log_prior_grad(kind, lam, t)
        gn = (inla.pc_log_prior(kind, lam, t + h) - inla.pc_log_prior(kind, lam, t - h)) / (2 * h)
        assert abs(g - gn) < 1e-6 * max(1, abs(g))


def test_interpretable_roundtrip():
    for th in ([-1.0, 0.7, 0.0], [0.3, -0.2, 1.1]):
        back = inla.compute2interpretable(*inla.interpretable2compute(*th))
        assert np.allclose(back, th, atol=1e-13)


def test_adjoint_gradient_matches_central_differences(tiny):
    d = tiny
    for th in (d.theta, d.theta + np.array([-0.3, 0.2, 0.1, 0.2])):
        f, g = inla.adjoint_gradient(d, th)
        _, gfd = inla.fd_gradient(d, th, 1e-4)
        assert np.linalg.norm(g - gfd) / np.linalg.norm(g) < 1e-6


def test_closed_form_prior_logdet(tiny):
    d = tiny
    th = d.theta
    _, parts = inla.f_theta(d, th, return_parts=True)
    v, g = inla.prior_logdet_closed_form(d, th, with_grad=True)
    assert abs(v - parts["logdet_qp"]) < 1e-11 * abs(v)
    e = 1e-5
    for j in range(3):
        tp, tm = th.copy(), th.copy()
        tp[j] += e
        tm[j] -= e
        gn = (inla.prior_logdet_closed_form(d, tp) - inla.prior_logdet_closed_form(d, tm)) / (2 * e)
        assert abs(g[j] - gn) < 1e-6 * max(1.0, abs(gn))


def test_fd_stencils_exact_on_quadratic():
    rng = np.random.default_rng(0)
    H = rng.standard_normal((3, 3))
    H = H + H.T
    gv = rng.standard_normal(3)
    fun = lambda t: 0.5 * t @ H @ t + gv @ t + 1.0
    th = rng.standard_normal(3)
    _, g = inla.fd_gradient(None, th, 1e-3, fun=fun)
    assert np.allclose(g, H @ th + gv, atol=1e-9)
    for k in (4, 2):
        Hn = inla.fd_hessian(None, th, 1e-3, offdiag_points=k, fun=fun)
        assert np.allclose(Hn, H, atol=1e-6)
    assert inla.hessian_stencil(np.zeros(7), 0.1, 4).shape[0] == 1 + 14 + 4 * 21
    assert inla.hessian_stencil(np.zeros(7), 0.1, 2).shape[0] == 57


def test_marginal_variances_match_dense_inverse():
    """Selected-inversion marginal variances equal the diagonal of the dense inverse,
    and diag(A X A^T) from the BTA pattern equals the dense covariance of the observations."""
    d = synth.make_st_model(4, 3, 2, 5)
    theta = d.theta + np.array([0.1, -0.2, 0.05, 0.3])
    Qc = inla.qcond_sparse(d, theta).toarray()
    C = np.linalg.inv(Qc)
    mv = inla.marginal_variances_latent(d, theta)
    assert mv.shape == (d.ns * d.nt + d.nb,)
    assert np.abs(mv - np.diag(C)).max() <= 1e-12 * np.abs(np.diag(C)).max()
    A = d.a_full().toarray()
    ov = inla.marginal_variances_observations(d, theta)
    ref = np.einsum("ij,jk,ik->i", A, C, A)
    assert np.abs(ov - ref).max() <= 1e-11 * np.abs(ref).max()


def test_sphere_adjoint_gradient_and_closed_form(tiny):
    """The sphere branch of the hyperparameter map through f, the adjoint gradient and the
    closed-form prior term (planar FEM inputs: only the map differs)."""
    import dataclasses
    d = dataclasses.replace(tiny, manifold="sphere")
    th = d.theta + np.array([-0.2, 0.1, 0.05, 0.1])
    f, g = inla.adjoint_gradient(d, th)
    f_plane = inla.f_theta(tiny, th)
    assert abs(f - f_plane) > 1e-6 * abs(f_plane)  # the map matters
    _, gfd = inla.fd_gradient(d, th, 1e-4)
    assert np.linalg.norm(g - gfd) / np.linalg.norm(g) < 1e-6
    _, parts = inla.f_theta(d, th, return_parts=True)
    v, gv = inla.prior_logdet_closed_form(d, th, with_grad=True)
    assert abs(v - parts["logdet_qp"]) < 1e-11 * abs(v)
    e = 1e-5
    for j in range(3):
        tp, tm = th.copy(), th.copy()
        tp[j] += e
        tm[j] -= e
        gn = (inla.prior_logdet_closed_form(d, tp) - inla.prior_logdet_closed_form(d, tm)) / (2 * e)
        assert abs(gv[j] - gn) < 1e-6 * max(1.0, abs(gn))
