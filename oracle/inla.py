"""CPU oracle for one f(theta) evaluation of DALIA's Gaussian spatio-temporal model.

TEST INFRASTRUCTURE ONLY (see `oracle/bta.py`).  A numpy/scipy restatement of
the reference path, function by function:

* `interpretable2compute`  <- `submodels/spatio_temporal.py:151-197` (plane and sphere branches)
* `qprior_sparse`          <- `submodels/spatio_temporal.py:114-149`,
                              `submodels/regression.py:28-35`, `core/model.py:306-382`
* `qcond_sparse`           <- `core/model.py:384-423`, `likelihoods/gaussian.py:98-131`
* `sparse_to_bta`          <- `solvers/structured_solver.py:192-250` (CPU branch)
* `log_prior`              <- `core/model.py:455-478`,
                              `prior_hyperparameters/penalized_complexity.py:13-85`,
                              `prior_hyperparameters/gaussian.py:26-29`
* `likelihood`             <- `likelihoods/gaussian.py:19-62` at eta = 0
* `f_theta`                <- `core/dalia.py:603-745` (Gaussian branch) with
                              `:1074-1108` and `:1110-1168`
* `fd_gradient`            <- `core/dalia.py:513-601`
* `fd_hessian`             <- `core/dalia.py:767-910`
* `adjoint_gradient`       <- no counterpart in the reference (it replaces the
                              FD gradient); restated from the derivative of f.
* `prior_logdet_closed_form` <- no counterpart; Kronecker-eigen identity.

Pinning (tests/test_oracle_inla.py): the Gaussian marginal likelihood
identity (f + log pi(theta) equals -log N(y; 0, A Qp^-1 A^T + e^-tau I) up to
a theta-independent constant), dense np.kron assembly, PC-prior
normalisation and tail probabilities, the interpretable/compute round trip,
central differences of f for the adjoint gradient, pobtf for the closed form.
"""

from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sps
from scipy.special import gamma as gamma_fn

from oracle import bta


# ---------------------------------------------------------------------------
# hyperparameter maps
# ---------------------------------------------------------------------------

def interpretable2compute(r_s, r_t, sigma_st, manifold: str = "plane"):
    """(r_s, r_t, sigma_st) -> (gamma_s, gamma_t, gamma_st) on the plane or the sphere.

    alpha_s=2, alpha_t=1, alpha_e=1 fixed (`spatio_temporal.py:158-160`); the sphere's spatial
    constant is the 50-term spherical-harmonic series of `:171-180`.  Written with numpy
    functions so that complex arguments work (complex-step derivatives of the sphere branch).
    """
    alpha_s, alpha_t, alpha_e = 2, 1, 1
    alpha = alpha_e + alpha_s * (alpha_t - 0.5)
    nu_s = alpha - 1
    nu_t = alpha_t - 0.5
    gamma_s = 0.5 * math.log(8 * nu_s) - r_s
    gamma_t = r_t - 0.5 * math.log(8 * nu_t) + alpha_s * gamma_s
    if manifold == "sphere":
        cr_t = gamma_fn(nu_t) / (gamma_fn(alpha_t) * pow(4 * math.pi, 0.5))
        c_s = 0.0
        for k in range(50):
            c_s = c_s + (2.0 * k + 1.0) / (4.0 * math.pi * (np.exp(gamma_s) ** 2 + k * (k + 1)) ** alpha)
        gamma_st = 0.5 * math.log(cr_t) + 0.5 * np.log(c_s) - 0.5 * gamma_t - sigma_st
    elif manifold == "plane":
        c1 = gamma_fn(nu_t) * gamma_fn(nu_s) / (gamma_fn(alpha_t) * gamma_fn(alpha) * pow(4 * math.pi, 1.5))
        gamma_st = 0.5 * math.log(c1) - 0.5 * gamma_t - nu_s * gamma_s - sigma_st
    else:
        raise ValueError(f"Manifold not supported: {manifold}")
    return gamma_s, gamma_t, gamma_st


def dgamma_dtheta(theta, manifold: str = "plane") -> np.ndarray:
    """Rows d gamma_s, d gamma_t, d gamma_st over columns (r_s, r_t, sigma_st): exact on the plane
    (the map is affine), by a complex step in r_s for gamma_st on the sphere."""
    J = np.array([[-1.0, 0.0, 0.0], [-2.0, 1.0, 0.0], [2.0, -0.5, -1.0]])
    if manifold == "sphere":
        h = 1e-30
        _, _, g = interpretable2compute(complex(theta[0], h), theta[1], theta[2], "sphere")
        J[2, 0] = g.imag / h
    return J


def compute2interpretable(gamma_s: float, gamma_t: float, gamma_st: float):
    """Inverse map, plane manifold (`spatio_temporal.py:199-257`)."""
    alpha_s, alpha_t, alpha_e = 2, 1, 1
    alpha = alpha_e + alpha_s * (alpha_t - 0.5)
    nu_s = alpha - 1
    nu_t = alpha_t - 0.5
    es, et, est = math.exp(gamma_s), math.exp(gamma_t), math.exp(gamma_st)
    r_s = math.log(math.sqrt(8 * nu_s) / es)
    r_t = math.log(et * math.sqrt(8 * nu_t) / pow(es, alpha_s))
    c1s = pow(4 * math.pi, 2 / 2.0) * pow(4 * math.pi, 0.5)
    c1 = gamma_fn(nu_t) * gamma_fn(nu_s) / (gamma_fn(alpha_t) * gamma_fn(alpha) * c1s)
    sigma_st = math.log(math.sqrt(c1) / ((est * math.sqrt(et)) * pow(es, alpha - 2 / 2.0)))
    return r_s, r_t, sigma_st


def kron_coefficients(theta, manifold: str = "plane") -> np.ndarray:
    """c[k, m] with Qp_ST = sum_k sum_m c[k,m] (M_k kron G_m), G = (c0, g1, g2, g3)."""
    gs, gt, gst = interpretable2compute(theta[0], theta[1], theta[2], manifold)
    k2 = math.exp(2 * gs)
    e2st = math.exp(2 * gst)
    et = math.exp(gt)
    c = np.zeros((3, 4))
    c[0] = e2st * np.array([k2 ** 3, 3 * k2 ** 2, 3 * k2, 1.0])
    c[1] = e2st * et * np.array([k2 ** 2, 2 * k2, 1.0, 0.0])
    c[2] = e2st * et * et * np.array([k2, 1.0, 0.0, 0.0])
    return c


# ---------------------------------------------------------------------------
# assembly
# ---------------------------------------------------------------------------

def qprior_st_sparse(data, theta) -> sps.csc_matrix:
    """Spatio-temporal prior precision, time-major `kron(m, q)` (`spatio_temporal.py:114-149`)."""
    gs, gt, gst = interpretable2compute(theta[0], theta[1], theta[2], getattr(data, "manifold", "plane"))
    es, et, est = math.exp(gs), math.exp(gt), math.exp(gst)
    q1 = es ** 2 * data.c0 + data.g1
    q2 = es ** 4 * data.c0 + 2 * es ** 2 * data.g1 + data.g2
    q3 = es ** 6 * data.c0 + 3 * es ** 4 * data.g1 + 3 * es ** 2 * data.g2 + data.g3
    Q = est ** 2 * (sps.kron(data.m0, q3) + et * sps.kron(data.m1, q2) + et ** 2 * sps.kron(data.m2, q1))
    return sps.csc_matrix(Q)


def qprior_sparse(data, theta) -> sps.csc_matrix:
    """Block-diagonal [Qp_ST, fe_prec I_nb] (`core/model.py:306-382`)."""
    Qst = qprior_st_sparse(data, theta)
    Qfe = data.fe_prec * sps.eye(data.nb)
    return sps.csc_matrix(sps.block_diag([Qst, Qfe]))


def qcond_sparse(data, theta, Qp=None) -> sps.csc_matrix:
    """Qc = Qp - A^T (-e^tau I) A (`core/model.py:421`)."""
    if Qp is None:
        Qp = qprior_sparse(data, theta)
    A = data.a_full()
    return sps.csc_matrix(Qp + math.exp(theta[-1]) * (A.T @ A))


def sparse_to_bta(Q, n: int, b: int, a: int, sparsity: str):
    """Dense BTA blocks of a sparse matrix (`structured_solver.py:192-250`)."""
    Qc = sps.csc_matrix(Q)
    D = np.zeros((n, b, b))
    L = np.zeros((max(n - 1, 0), b, b))
    R = np.zeros((n, a, b)) if a else None
    T = np.zeros((a, a)) if a else None
    for i in range(n):
        D[i] = Qc[i * b:(i + 1) * b, i * b:(i + 1) * b].toarray()
        if i < n - 1:
            L[i] = Qc[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b].toarray()
        if sparsity == "bta" and a:
            R[i] = Qc[n * b:n * b + a, i * b:(i + 1) * b].toarray()
    if sparsity == "bta" and a:
        T[:] = Qc[n * b:, n * b:].toarray()
    return D, L, R, T


# ---------------------------------------------------------------------------
# scalar terms
# ---------------------------------------------------------------------------

_PC_TYPES = ("r_s", "r_t", "sigma_st", "prec_o")


def pc_lambda(kind: str, alpha: float, u: float) -> float:
    """PC-prior rate (`penalized_complexity.py:28-43`, spatial dimension 2)."""
    if kind == "r_s":
        return -math.log(alpha) * pow(u, 0.5 * 2)
    if kind == "r_t":
        return -math.log(alpha) * pow(u, 0.5)
    return -math.log(alpha) / u


def pc_log_prior(kind: str, lam: float, th: float) -> float:
    """`penalized_complexity.py:51-85`."""
    if kind == "r_s":
        return math.log(lam) - lam * math.exp(-th) - th
    if kind == "r_t":
        return math.log(lam) - lam * math.exp(-0.5 * th) + math.log(0.5) - 0.5 * th
    return math.log(lam) - lam * math.exp(th) + th


def pc_log_prior_grad(kind: str, lam: float, th: float) -> float:
    if kind == "r_s":
        return lam * math.exp(-th) - 1.0
    if kind == "r_t":
        return 0.5 * lam * math.exp(-0.5 * th) - 0.5
    return -lam * math.exp(th) + 1.0


def log_prior(theta, priors) -> float:
    """Sum of hyper-prior log densities in theta order (`core/model.py:455-478`)."""
    s = 0.0
    for i, p in enumerate(priors):
        if p.kind == "penalized_complexity":
            s += pc_log_prior(_PC_TYPES[i], pc_lambda(_PC_TYPES[i], p.alpha, p.u), theta[i])
        elif p.kind == "gaussian":
            s += -0.5 * p.precision * (theta[i] - p.mean) ** 2
        else:
            raise ValueError(p.kind)
    return s


def log_prior_grad(theta, priors) -> np.ndarray:
    g = np.zeros(len(priors))
    for i, p in enumerate(priors):
        if p.kind == "penalized_complexity":
            g[i] = pc_log_prior_grad(_PC_TYPES[i], pc_lambda(_PC_TYPES[i], p.alpha, p.u), theta[i])
        else:
            g[i] = -p.precision * (theta[i] - p.mean)
    return g


def likelihood(theta, y: np.ndarray) -> float:
    """Gaussian log-likelihood at eta = 0, without the 2 pi constant (`gaussian.py:58-60`)."""
    tau = theta[-1]
    return 0.5 * tau * y.shape[0] - 0.5 * math.exp(tau) * float(y @ y)


# ---------------------------------------------------------------------------
# f(theta)
# ---------------------------------------------------------------------------

def f_theta(data, theta, return_parts: bool = False):
    """One f(theta) evaluation exactly as the reference's Gaussian branch.

    f = [1/2 logdet Qc - 1/2 mu^T Qc mu]
        - [log pi(theta) + lik(eta=0) + 1/2 logdet Qp|BT]
    with mu = Qc^-1 (e^tau A^T y).
    """
    theta = np.asarray(theta, dtype=np.float64)
    n, b, a = data.nt, data.ns, data.nb
    Qp = qprior_sparse(data, theta)
    Qc = qcond_sparse(data, theta, Qp)
    D, L, R, T = sparse_to_bta(Qc, n, b, a, "bta")
    bta.pobtaf(D, L, R, T)
    A = data.a_full()
    rhs = math.exp(theta[-1]) * (A.T @ data.y)
    mu = bta.pobtas(D, L, R, T, rhs.copy(), "N")
    mu = bta.pobtas(D, L, R, T, mu, "C")
    ld_c = bta.logdet(D, T)
    quad = float(mu @ (Qc @ mu))
    cond = 0.5 * ld_c - 0.5 * quad
    lp = log_prior(theta, data.priors)
    lik = likelihood(theta, data.y)
    Dp, Lp, _, _ = sparse_to_bta(Qp, n, b, a, "bt")
    bta.pobtf(Dp, Lp)
    ld_p = bta.logdet(Dp)
    f = cond - (lp + lik + 0.5 * ld_p)
    if np.isnan(f):
        raise ValueError("f(theta) is NaN")
    if return_parts:
        return f, dict(logdet_qc=ld_c, quad=quad, log_prior=lp, lik=lik, logdet_qp=ld_p, mu=mu, rhs=rhs)
    return f


def fd_gradient(data, theta, h: float, fun=None):
    """Central differences with the reference's 2d+1 stencil (`core/dalia.py:545-581`).

    Evaluation order is reversed so that the last evaluation is the bare theta,
    as in `core/dalia.py:558-564`.  Returns (f(theta), grad).
    """
    fun = fun or (lambda th: f_theta(data, th))
    theta = np.asarray(theta, dtype=np.float64)
    d = theta.size
    n_eval = 2 * d + 1
    tm = np.repeat(theta.reshape(-1, 1), n_eval, axis=1)
    tm[:, 1:1 + d] += h * np.eye(d)
    tm[:, d + 1:n_eval] -= h * np.eye(d)
    fv = np.zeros(n_eval)
    for k in range(n_eval - 1, -1, -1):
        fv[k] = fun(tm[:, k])
    g = (fv[1:1 + d] - fv[d + 1:]) / (2 * h)
    return fv[0], g


def hessian_stencil(theta, h: float, offdiag_points: int = 4):
    """theta points of the FD Hessian stencil.

    offdiag_points=4 is the reference's stencil, 1 + 2d + 4 d(d-1)/2 points
    (`core/dalia.py:793-871`); offdiag_points=2 keeps only the (+,+) and
    (-,-) corners, 1 + 2d + 2 d(d-1)/2 points, and uses
    H_ij = (f++ + f-- - f+i - f-i - f+j - f-j + 2 f0) / (2 h^2).
    """
    theta = np.asarray(theta, dtype=np.float64)
    d = theta.size
    E = h * np.eye(d)
    pts = [theta.copy()]
    for i in range(d):
        pts.append(theta + E[i])
        pts.append(theta - E[i])
    for i in range(d):
        for j in range(i + 1, d):
            if offdiag_points == 4:
                pts += [theta + E[i] + E[j], theta + E[i] - E[j], theta - E[i] + E[j], theta - E[i] - E[j]]
            else:
                pts += [theta + E[i] + E[j], theta - E[i] - E[j]]
    return np.stack(pts)


def hessian_from_values(fv: np.ndarray, d: int, h: float, offdiag_points: int = 4) -> np.ndarray:
    """Assemble H from stencil values (`core/dalia.py:887-902` for the 4-point form)."""
    H = np.zeros((d, d))
    f0 = fv[0]
    fp = fv[1:1 + 2 * d:2]
    fm = fv[2:2 + 2 * d:2]
    for i in range(d):
        H[i, i] = (fp[i] - 2 * f0 + fm[i]) / h ** 2
    k = 1 + 2 * d
    for i in range(d):
        for j in range(i + 1, d):
            if offdiag_points == 4:
                fpp, fpm, fmp, fmm = fv[k:k + 4]
                k += 4
                H[i, j] = (fpp - fpm - fmp + fmm) / (4 * h * h)
            else:
                fpp, fmm = fv[k:k + 2]
                k += 2
                H[i, j] = (fpp + fmm - fp[i] - fm[i] - fp[j] - fm[j] + 2 * f0) / (2 * h * h)
            H[j, i] = H[i, j]
    return H


def fd_hessian(data, theta, h: float, offdiag_points: int = 4, fun=None) -> np.ndarray:
    fun = fun or (lambda th: f_theta(data, th))
    pts = hessian_stencil(theta, h, offdiag_points)
    fv = np.array([fun(p) for p in pts])
    return hessian_from_values(fv, len(theta), h, offdiag_points)


# ---------------------------------------------------------------------------
# derivatives of the Kronecker coefficients
# ---------------------------------------------------------------------------

def kron_coefficients_grad(theta, manifold: str = "plane") -> np.ndarray:
    """dc[j, k, m] = d c[k,m] / d theta_j for j in (r_s, r_t, sigma_st).

    c[k,m] = exp(2 gst + k gt + 2 p_m gs) * const with p = (3,2,1,0) for k=0,
    (2,1,0,-) for k=1, (1,0,-,-) for k=2; chain rule through
    gs = C - r_s, gt = r_t - C' + 2 gs, gst = C'' - gt/2 - gs - sigma_st.
    """
    c = kron_coefficients(theta, manifold)
    # exponent of exp(gs) in k2 powers: c[k,m] ~ k2^(p) with p listed below
    p = np.array([[3, 2, 1, 0], [2, 1, 0, 0], [1, 0, 0, 0]], dtype=np.float64)
    kk = np.array([0.0, 1.0, 2.0])[:, None]
    dg_s = 2.0 * p          # d log c / d gamma_s
    dg_t = kk * np.ones((1, 4))  # d log c / d gamma_t
    dg_st = 2.0 * np.ones((3, 4))  # d log c / d gamma_st
    # d gamma / d theta: rows (r_s, r_t, sigma)
    dgs, dgt, dgst = dgamma_dtheta(theta, manifold)
    out = np.zeros((3, 3, 4))
    for j in range(3):
        out[j] = c * (dg_s * dgs[j] + dg_t * dgt[j] + dg_st * dgst[j])
    return out


def dqprior_st_sparse(data, theta, j: int) -> sps.csc_matrix:
    dc = kron_coefficients_grad(theta, getattr(data, "manifold", "plane"))[j]
    G = [data.c0, data.g1, data.g2, data.g3]
    Mk = [data.m0, data.m1, data.m2]
    Q = None
    for k in range(3):
        qk = sum(dc[k, m] * G[m] for m in range(4))
        term = sps.kron(Mk[k], qk)
        Q = term if Q is None else Q + term
    return sps.csc_matrix(Q)


# ---------------------------------------------------------------------------
# adjoint gradient (oracle form: selected inversion of Qc and of Qp|BT)
# ---------------------------------------------------------------------------

def _trace_bta(Xblocks, dQ, n, b, a, sparsity):
    """tr(X dQ) with X given on the BTA pattern and dQ sparse symmetric."""
    D, L, R, T = Xblocks
    dD, dL, dR, dT = sparse_to_bta(dQ, n, b, a, sparsity)
    s = float(np.sum(D * dD)) + 2.0 * float(np.sum(L * dL))
    if sparsity == "bta" and a:
        s += 2.0 * float(np.sum(R * dR)) + float(np.sum(T * dT))
    return s


def adjoint_gradient(data, theta):
    """grad f by the adjoint formula; returns (f, grad).

    d f/d theta_j = 1/2 tr(Qc^-1 dQc) - (db)^T mu + 1/2 mu^T dQc mu
                    - 1/2 tr(Qp|BT^-1 dQp|BT) - d log pi - d lik
    """
    theta = np.asarray(theta, dtype=np.float64)
    n, b, a = data.nt, data.ns, data.nb
    f, parts = f_theta(data, theta, return_parts=True)
    mu = parts["mu"]
    Qp = qprior_sparse(data, theta)
    Qc = qcond_sparse(data, theta, Qp)
    Dc, Lc, Rc, Tc = sparse_to_bta(Qc, n, b, a, "bta")
    bta.pobtaf(Dc, Lc, Rc, Tc)
    bta.pobtasi(Dc, Lc, Rc, Tc)
    Dp, Lp, _, _ = sparse_to_bta(Qp, n, b, a, "bt")
    bta.pobtf(Dp, Lp)
    bta.pobtsi(Dp, Lp)
    d = theta.size
    g = np.zeros(d)
    N = n * b + a
    for j in range(3):
        dQst = dqprior_st_sparse(data, theta, j)
        dQ = sps.csc_matrix(sps.block_diag([dQst, sps.csc_matrix((a, a))])) if a else dQst
        tr_c = _trace_bta((Dc, Lc, Rc, Tc), dQ, n, b, a, "bta")
        tr_p = _trace_bta((Dp, Lp, None, None), dQ, n, b, a, "bt")
        qf = float(mu @ (dQ @ mu))
        g[j] = 0.5 * tr_c + 0.5 * qf - 0.5 * tr_p
    A = data.a_full()
    tau = theta[-1]
    AtA = sps.csc_matrix(A.T @ A)
    dQc = math.exp(tau) * AtA
    tr_c = _trace_bta((Dc, Lc, Rc, Tc), dQc, n, b, a, "bta")
    bvec = parts["rhs"]
    qf = float(mu @ (dQc @ mu))
    g[3] = 0.5 * tr_c - float(bvec @ mu) + 0.5 * qf
    g[3] -= 0.5 * data.n_obs - 0.5 * math.exp(tau) * float(data.y @ data.y)
    g -= log_prior_grad(theta, data.priors)
    assert N == mu.size
    return f, g


# ---------------------------------------------------------------------------
# marginal variances (core/dalia.py:912-1016)
# ---------------------------------------------------------------------------

def selected_inverse_blocks(data, theta):
    """Selected inverse of Qc(theta) on the BTA pattern (core/dalia.py:912-937)."""
    n, b, a = data.nt, data.ns, data.nb
    Qc = qcond_sparse(data, np.asarray(theta, dtype=np.float64))
    D, L, R, T = sparse_to_bta(Qc, n, b, a, "bta")
    bta.pobtaf(D, L, R, T)
    bta.pobtasi(D, L, R, T)
    return D, L, R, T


def marginal_variances_latent(data, theta) -> np.ndarray:
    """diag(Qc^-1): diagonals of the selected inverse blocks, then the tip
    (core/dalia.py:939-962, extract_diagonal of _structured_to_spmatrix(eye))."""
    D, _, _, T = selected_inverse_blocks(data, theta)
    parts = [np.diag(D[i]) for i in range(data.nt)]
    if data.nb:
        parts.append(np.diag(T))
    return np.concatenate(parts)


def marginal_variances_observations(data, theta) -> np.ndarray:
    """diag(A X A^T), X the selected inverse on Qc's pattern (core/dalia.py:964-1016).

    Every pair of latents that share an observation row is a nonzero of
    A^T A, hence of Qc, so X on the pattern is enough; the product is formed
    from the densified pattern (small inputs only).
    """
    D, L, R, T = selected_inverse_blocks(data, theta)
    X = bta.bta_to_dense(D, L, R if data.nb else None, T if data.nb else None)
    A = data.a_full().tocsr()
    AX = np.asarray(A @ X)
    return np.einsum("ij,ij->i", AX, A.toarray())


# ---------------------------------------------------------------------------
# Kronecker-eigen closed form for logdet Qp|BT
# ---------------------------------------------------------------------------

def spatial_eigenvalues(data) -> np.ndarray:
    """Eigenvalues of c0^-1/2 g1 c0^-1/2 (theta independent)."""
    c = data.c0.diagonal()
    s = 1.0 / np.sqrt(c)
    S = (sps.diags(s) @ data.g1 @ sps.diags(s)).toarray()
    return np.linalg.eigvalsh(0.5 * (S + S.T))


def prior_logdet_closed_form(data, theta, lam=None, with_grad: bool = False):
    """logdet Qp|BT = nt sum log c0 + sum_j logdet T_j (and d/d(r_s,r_t,sigma)).

    Valid when c0 is diagonal, g2 = g1 c0^-1 g1 and g3 = g1 c0^-1 g2.
    T_j = sum_k p_k(lt_j) M_k, p_0 = e^{2gst} lt^3, p_1 = e^{2gst+gt} lt^2,
    p_2 = e^{2gst+2gt} lt, lt_j = e^{2gs} + lambda_j.
    """
    if lam is None:
        lam = spatial_eigenvalues(data)
    manifold = getattr(data, "manifold", "plane")
    gs, gt, gst = interpretable2compute(theta[0], theta[1], theta[2], manifold)
    k2 = math.exp(2 * gs)
    lt = k2 + lam
    pw = np.array([3.0, 2.0, 1.0])
    coef = np.array([math.exp(2 * gst), math.exp(2 * gst + gt), math.exp(2 * gst + 2 * gt)])
    M = np.stack([data.m0.toarray(), data.m1.toarray(), data.m2.toarray()])
    Pk = coef[None, :] * lt[:, None] ** pw[None, :]  # (ns, 3)
    Tj = np.einsum("jk,kab->jab", Pk, M)
    sign, ld = np.linalg.slogdet(Tj)
    assert np.all(sign > 0)
    val = data.nt * float(np.sum(np.log(data.c0.diagonal()))) + float(np.sum(ld))
    if not with_grad:
        return val
    Tinv = np.linalg.inv(Tj)
    trM = np.einsum("jab,kba->jk", Tinv, M)  # tr(T_j^-1 M_k)
    # dP/dgamma: gs via lt (d lt/d gs = 2 k2), gt via coef exponent, gst
    dP_gs = coef[None, :] * pw[None, :] * lt[:, None] ** (pw[None, :] - 1) * 2 * k2
    dP_gt = Pk * np.array([0.0, 1.0, 2.0])[None, :]
    dP_gst = 2.0 * Pk
    d_gs = float(np.sum(trM * dP_gs))
    d_gt = float(np.sum(trM * dP_gt))
    d_gst = float(np.sum(trM * dP_gst))
    dgs, dgt, dgst = dgamma_dtheta(theta, manifold)
    grad = d_gs * dgs + d_gt * dgt + d_gst * dgst
    return val, grad
