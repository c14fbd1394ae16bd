"""CPU: pin the coregional oracle (no GPU)."""

import math

import numpy as np
import pytest

from dalia_amd import synth
from oracle import coreg
from oracle import inla


@pytest.fixture(scope="module", params=[2, 3])
def small(request):
    return synth.make_coreg_model(4, 3, 2, request.param, 5)


def univariate_view(d, i):
    """Variate i of a coregional model as a univariate model (prior sigma_st = sigma_i)."""
    return synth.STModelData(ns=d.ns, nt=d.nt, nb=d.nb, c0=d.c0, g1=d.g1, g2=d.g2, g3=d.g3, m0=d.m0, m1=d.m1,
                             m2=d.m2, a_st=d.a_st[i], x_cov=d.x_cov[i], y=d.y[i],
                             theta=None, priors=synth.DEFAULT_PRIORS)


def test_lambda_zero_decouples_into_univariate_models(small):
    """With every lambda = 0 the variates are independent: f + log pi of the coregional
    model equals the sum of the univariate f + log pi, variate i's prior scaled by
    sigma_i exactly as a univariate sigma_st (gamma_st shifts by -sigma)."""
    d = small
    th = d.theta.copy()
    th[4 * d.nv:] = 0.0
    lhs = coreg.f_theta_coreg(d, th) + coreg.log_prior(d, th)
    rhs = 0.0
    for i in range(d.nv):
        u = univariate_view(d, i)
        thu = np.array([th[3 * i], th[3 * i + 1], d.sigma_st + th[3 * d.nv + i], th[3 * i + 2]])
        rhs += inla.f_theta(u, thu) + inla.log_prior(thu, u.priors)
    assert abs(lhs - rhs) <= 1e-9 * max(1.0, abs(rhs))


def test_marginal_likelihood_identity(small):
    """f + log pi(theta) = -log N(y; 0, A Qp^-1 A^T + D^-1) + 1/2 a log(fe_prec) - n/2 log 2 pi."""
    d = small
    for th in (d.theta, d.theta + 0.1 * np.sin(np.arange(d.theta.size) + 1.0)):
        f = coreg.f_theta_coreg(d, th)
        Qp = coreg.qprior_coreg_sparse(d, th).toarray()
        A = d.a_full().toarray()
        Sig = A @ np.linalg.solve(Qp, A.T) + np.diag(1.0 / coreg.obs_precisions(d, th))
        y = np.concatenate(d.y)
        logpy = -0.5 * np.linalg.slogdet(Sig)[1] - 0.5 * y @ np.linalg.solve(Sig, y) - 0.5 * y.size * math.log(2 * math.pi)
        const = 0.5 * d.a * math.log(d.fe_prec) - 0.5 * y.size * math.log(2 * math.pi)
        lhs = f + coreg.log_prior(d, th)
        assert abs(lhs - (-logpy + const)) < 1e-8 * max(1.0, abs(lhs))


def test_prior_is_unit_triangular_mixing(small):
    """logdet of the assembled, permuted prior = sum_i (logdet Qu_i - 2 ns nt log sigma_i),
    and the prior is SPD for a nonzero lambda."""
    d = small
    th = d.theta
    Q = coreg.qprior_st_coreg_sparse(d, th).toarray()
    sign, ld = np.linalg.slogdet(Q)
    assert sign > 0
    assert abs(ld - coreg.prior_logdet_mixing(d, th)) < 1e-9 * abs(ld)
    np.linalg.cholesky(Q)


def test_permutation_places_variates_in_time_blocks(small):
    """Entry (t, i, s) x (t', j, s') of the permuted prior is entry (i, t, s) x (j, t', s')
    of the variate-major block matrix, and the pattern is block tridiagonal in t."""
    d = small
    th = d.theta
    Q = coreg.qprior_st_coreg_sparse(d, th).toarray()
    Qu = [coreg.variate_qu(d, k, th).toarray() for k in range(d.nv)]
    w = coreg.mixing_weights(d, th)
    nst, ns, b = d.ns * d.nt, d.ns, d.b
    rng = np.random.default_rng(0)
    for _ in range(200):
        t, tp = rng.integers(0, d.nt, 2)
        i, j = rng.integers(0, d.nv, 2)
        s, sp = rng.integers(0, ns, 2)
        ref = sum(w[(i, j)][k] * Qu[k][t * ns + s, tp * ns + sp] for k in range(d.nv))
        assert abs(Q[t * b + i * ns + s, tp * b + j * ns + sp] - ref) <= 1e-12 * max(1.0, abs(ref))
    for t in range(d.nt):
        for tp in range(d.nt):
            if abs(t - tp) > 1:
                assert not np.any(Q[t * b:(t + 1) * b, tp * b:(tp + 1) * b])
