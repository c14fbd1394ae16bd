"""CPU oracle for one f(theta) evaluation of DALIA's coregional Gaussian model.

TEST INFRASTRUCTURE ONLY (see `oracle/bta.py`): the checker of the HIP
coregional path, never part of it.  numpy/scipy restatement:

* `qprior_coreg_sparse`  <- `models/coregional_model.py:279-551` (q_ij blocks,
                            symmetric permutation, regression tip)
* `qcond_coreg_sparse`   <- `models/coregional_model.py:553-639`
                            (Q - A^T D A with per-variate D = -exp(prec_o_i) I)
* `f_theta_coreg`        <- `core/dalia.py:603-745` (Gaussian branch) with the
                            coregional model's pieces; the prior log-determinant
                            is the BT factorization of the permuted
                            spatio-temporal block (`core/dalia.py:1074-1108`).

Pinning (tests/test_oracle_coreg.py): a dense restatement of the block
formulas with explicit permutation matrices, the unit-triangular mixing
identity logdet Qp_st = sum_i (logdet Qu_i - 2 ns nt log sigma_i), the
Gaussian marginal-likelihood identity, and the reduction to independent
univariate models when every lambda is 0.
"""

from __future__ import annotations

import math
import types

import numpy as np
import scipy.sparse as sps

from oracle import bta
from oracle import inla as oinla


def variate_qu(data, i: int, theta) -> sps.csc_matrix:
    """Qu_i: the spatio-temporal prior of variate i with its (r_s, r_t) and the
    fixed sigma_st (`coregional_model.py:283-298`, `spatio_temporal.py:114-149`)."""
    th = np.array([theta[3 * i], theta[3 * i + 1], data.sigma_st])
    return oinla.qprior_st_sparse(data, th)


def mixing_weights(data, theta) -> dict:
    """w[(i, j)][k]: q_ij = sum_k w[(i,j)][k] Qu_k, as written in
    `coregional_model.py:304-497` (lambdas raw, sigmas exp(theta))."""
    nv = data.nv
    s = [math.exp(theta[3 * nv + k]) for k in range(nv)]
    lam = theta[4 * nv:]
    w = {}
    if nv == 2:
        l01 = lam[0]
        w[(0, 0)] = [1 / s[0] ** 2, l01 ** 2 / s[1] ** 2]
        w[(1, 0)] = [0.0, -l01 / s[1] ** 2]
        w[(1, 1)] = [0.0, 1 / s[1] ** 2]
    else:
        l01, l02, l12 = lam[0], lam[1], lam[2]
        w[(0, 0)] = [1 / s[0] ** 2, l01 ** 2 / s[1] ** 2, l12 ** 2 / s[2] ** 2]
        w[(1, 0)] = [0.0, -l01 / s[1] ** 2, l02 * l12 / s[2] ** 2]
        w[(2, 0)] = [0.0, 0.0, -l12 / s[2] ** 2]
        w[(1, 1)] = [0.0, 1 / s[1] ** 2, l02 ** 2 / s[2] ** 2]
        w[(2, 1)] = [0.0, 0.0, -l02 / s[2] ** 2]
        w[(2, 2)] = [0.0, 0.0, 1 / s[2] ** 2]
    for i in range(nv):
        for j in range(i + 1, nv):
            w[(i, j)] = w[(j, i)]
    return w


def st_permutation(data) -> np.ndarray:
    """Rows of the variate-major spatio-temporal prior in time-block order
    (`coregional_model.py:785-815`)."""
    nst = data.ns * data.nt
    blocks = [np.arange(nst).reshape(data.nt, data.ns) + i * nst for i in range(data.nv)]
    return np.hstack(blocks).ravel()


def qprior_st_coreg_sparse(data, theta) -> sps.csc_matrix:
    """Permuted coregional spatio-temporal prior (size nv ns nt)."""
    nv = data.nv
    Qu = [variate_qu(data, k, theta) for k in range(nv)]
    w = mixing_weights(data, theta)
    rows = []
    for i in range(nv):
        row = []
        for j in range(nv):
            q = sum(w[(i, j)][k] * Qu[k] for k in range(nv) if w[(i, j)][k] != 0.0)
            row.append(sps.csc_matrix(q) if not np.isscalar(q) else sps.csc_matrix(Qu[0].shape))
        rows.append(row)
    Q = sps.bmat(rows, format="csc")
    p = st_permutation(data)
    return sps.csc_matrix(Q[p, :][:, p])


def qprior_coreg_sparse(data, theta) -> sps.csc_matrix:
    """[permuted spatio-temporal block, fe_prec I_{nv nb}] (`coregional_model.py:521-551`)."""
    Qst = qprior_st_coreg_sparse(data, theta)
    if data.a == 0:
        return Qst
    return sps.csc_matrix(sps.block_diag([Qst, data.fe_prec * sps.eye(data.a)]))


def obs_precisions(data, theta) -> np.ndarray:
    """exp(prec_o_i) for every observation row of a_full (`coregional_model.py:595-639`)."""
    tau = np.array([math.exp(theta[3 * i + 2]) for i in range(data.nv)])
    return tau[data.obs_precision_index()]


def qcond_coreg_sparse(data, theta, Qp=None) -> sps.csc_matrix:
    """Qc = Qp - A^T D A, D = -exp(prec_o) per observation."""
    if Qp is None:
        Qp = qprior_coreg_sparse(data, theta)
    A = data.a_full()
    d = obs_precisions(data, theta)
    return sps.csc_matrix(Qp + A.T @ sps.diags(d) @ A)


def log_prior(data, theta) -> float:
    return oinla.log_prior(theta, data.priors)


def likelihood(data, theta) -> float:
    """sum_i 1/2 prec_o_i n_i - 1/2 exp(prec_o_i) y_i^T y_i at eta = 0 (`likelihoods/gaussian.py:19-62`)."""
    s = 0.0
    for i in range(data.nv):
        p = theta[3 * i + 2]
        s += 0.5 * p * data.y[i].size - 0.5 * math.exp(p) * float(data.y[i] @ data.y[i])
    return s


def f_theta_coreg(data, theta, return_parts: bool = False):
    """f(theta) of the coregional model, exactly as the univariate one is built
    (`core/dalia.py:631-698`): BTA Cholesky of Qc (b = nv ns, a = nv nb), solve,
    quadratic form by a product with Qc, BT Cholesky of the prior."""
    theta = np.asarray(theta, dtype=np.float64)
    n, b, a = data.nt, data.b, data.a
    Qp = qprior_coreg_sparse(data, theta)
    Qc = qcond_coreg_sparse(data, theta, Qp)
    A = data.a_full()
    y = np.concatenate(data.y)
    rhs = A.T @ (obs_precisions(data, theta) * y)
    D, L, R, T = oinla.sparse_to_bta(Qc, n, b, a, "bta")
    bta.pobtaf(D, L, R, T)
    ld_c = bta.logdet(D, T if a else None)
    mu = bta.pobtas(D, L, R, T, rhs.copy(), "N")
    mu = bta.pobtas(D, L, R, T, mu, "C")
    quad = float(mu @ (Qc @ mu))
    Dp, Lp, _, _ = oinla.sparse_to_bta(Qp, n, b, a, "bt")
    bta.pobtf(Dp, Lp)
    ld_p = bta.logdet(Dp)
    lp = log_prior(data, theta)
    lik = likelihood(data, theta)
    f = 0.5 * ld_c - 0.5 * quad - (lp + lik + 0.5 * ld_p)
    if return_parts:
        return f, dict(ld_c=ld_c, quad=quad, ld_p=ld_p, log_prior=lp, lik=lik, mu=mu, rhs=rhs)
    return f


def prior_logdet_mixing(data, theta) -> float:
    """sum_i (logdet Qu_i - 2 ns nt log sigma_i): the mixing of the q_ij blocks is
    unit lower triangular (no counterpart in the reference; used to pin the
    assembled prior)."""
    s = 0.0
    nst = data.ns * data.nt
    for i in range(data.nv):
        Qu = variate_qu(data, i, theta).toarray()
        s += np.linalg.slogdet(Qu)[1] - 2.0 * nst * theta[3 * data.nv + i]
    return s
