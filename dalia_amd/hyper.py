"""Scalar hyperparameter algebra of the spatio-temporal Gaussian model (host side).

Only O(1) scalar work lives here: the map from the interpretable
hyperparameters to the SPDE parameters, the Kronecker coefficients of the
prior precision and their derivatives, the hyper-prior log densities and the
Gaussian likelihood constant.  Reference formulas:

* `interpretable2compute`  `submodels/spatio_temporal.py:151-197` (plane and sphere
  manifolds; alpha_s = 2, alpha_t = 1, alpha_e = 1 as fixed at `:158-160`)
* prior precision          `submodels/spatio_temporal.py:114-149`
* PC priors                `prior_hyperparameters/penalized_complexity.py:13-85`
* Gaussian priors          `prior_hyperparameters/gaussian.py:26-29`
* likelihood at eta = 0    `likelihoods/gaussian.py:58-60`
"""

from __future__ import annotations

import math

import numpy as np

# Gamma(1/2) Gamma(1) / (Gamma(1) Gamma(2) (4 pi)^1.5) for nu_t = 1/2, nu_s = 1,
# alpha_t = 1, alpha = 2.
_C1 = math.gamma(0.5) * math.gamma(1.0) / (math.gamma(1.0) * math.gamma(2.0) * pow(4 * math.pi, 1.5))

#: d(gamma_s, gamma_t, gamma_st) / d(r_s, r_t, sigma_st), rows = gamma, cols = theta
DGAMMA = np.array([
    [-1.0, 0.0, 0.0],
    [-2.0, 1.0, 0.0],
    [2.0, -0.5, -1.0],
])

_PC_KIND = ("r_s", "r_t", "sigma_st", "prec_o")


# sphere: Gamma(nu_t) / (Gamma(alpha_t) (4 pi)^1/2) and the spherical-harmonic series of the
# spatial variance, sum_{k < 50} (2k + 1) / (4 pi (kappa^2 + k (k + 1))^alpha), alpha = 2
_CRT = math.gamma(0.5) / (math.gamma(1.0) * math.sqrt(4 * math.pi))
_SPH_K = np.arange(50, dtype=np.float64)
_SPH_W = (2.0 * _SPH_K + 1.0) / (4.0 * math.pi)
_SPH_L = _SPH_K * (_SPH_K + 1.0)
MANIFOLDS = ("plane", "sphere")


def _sphere_series(gamma_s: float):
    """(c_s, d log c_s / d gamma_s) of the 50-term series (`spatio_temporal.py:171-180`)."""
    k2 = math.exp(2.0 * gamma_s)
    lt = k2 + _SPH_L
    c_s = float(np.sum(_SPH_W / lt ** 2))
    dc_s = float(np.sum(_SPH_W * (-2.0) / lt ** 3 * 2.0 * k2))
    return c_s, dc_s / c_s


def interpretable2compute(r_s: float, r_t: float, sigma_st: float, manifold: str = "plane"):
    gamma_s = 0.5 * math.log(8.0) - r_s
    gamma_t = r_t - 0.5 * math.log(4.0) + 2.0 * gamma_s
    if manifold == "plane":
        gamma_st = 0.5 * math.log(_C1) - 0.5 * gamma_t - gamma_s - sigma_st
    elif manifold == "sphere":
        c_s, _ = _sphere_series(gamma_s)
        gamma_st = 0.5 * math.log(_CRT) + 0.5 * math.log(c_s) - 0.5 * gamma_t - sigma_st
    else:
        raise ValueError(f"Manifold not supported: {manifold}")
    return gamma_s, gamma_t, gamma_st


def dgamma(theta, manifold: str = "plane") -> np.ndarray:
    """d(gamma_s, gamma_t, gamma_st) / d(r_s, r_t, sigma_st) at theta (rows gamma, cols theta)."""
    if manifold == "plane":
        return DGAMMA
    if manifold != "sphere":
        raise ValueError(f"Manifold not supported: {manifold}")
    gs, _, _ = interpretable2compute(float(theta[0]), float(theta[1]), float(theta[2]), manifold)
    _, dlog = _sphere_series(gs)
    # gamma_st = 1/2 log c_s(gamma_s) - 1/2 gamma_t - sigma + const, gamma_s = const - r_s
    return np.array([[-1.0, 0.0, 0.0], [-2.0, 1.0, 0.0], [1.0 - 0.5 * dlog, -0.5, -1.0]])


# powers of kappa^2 = exp(2 gamma_s) in c[k][m] for G = (c0, g1, g2, g3)
_PK2 = np.array([[3, 2, 1, 0], [2, 1, 0, 0], [1, 0, 0, 0]], dtype=np.float64)
_BINOM = np.array([[1, 3, 3, 1], [1, 2, 1, 0], [1, 1, 0, 0]], dtype=np.float64)


def kron_coefficients(theta, manifold: str = "plane") -> np.ndarray:
    """c[k][m]: Q_prior_ST = sum_k sum_m c[k][m] (M_k kron G_m).

    M0 kron q3 + e^gt M1 kron q2 + e^2gt M2 kron q1, all times e^2gst, with
    q1 = k^2 c0 + g1, q2 = k^4 c0 + 2k^2 g1 + g2, q3 = k^6 c0 + 3k^4 g1 + 3k^2 g2 + g3.
    """
    gs, gt, gst = interpretable2compute(float(theta[0]), float(theta[1]), float(theta[2]), manifold)
    k2 = math.exp(2.0 * gs)
    scale = np.array([math.exp(2 * gst), math.exp(2 * gst + gt), math.exp(2 * gst + 2 * gt)])
    return scale[:, None] * _BINOM * np.power(k2, _PK2) * (_BINOM != 0)


def kron_coefficients_grad(theta, manifold: str = "plane") -> np.ndarray:
    """dc[j][k][m] = d c[k][m] / d theta_j for j = r_s, r_t, sigma_st."""
    c = kron_coefficients(theta, manifold)
    dlog = np.stack([
        2.0 * _PK2,                                   # d log c / d gamma_s
        np.array([0.0, 1.0, 2.0])[:, None] * np.ones((1, 4)),  # / d gamma_t
        2.0 * np.ones((3, 4)),                        # / d gamma_st
    ])  # (3 gammas, 3, 4)
    out = np.einsum("gj,gkm->jkm", dgamma(theta, manifold), dlog) * c[None]
    return out


def pc_lambda(kind: str, alpha: float, u: float) -> float:
    if kind == "r_s":
        return -math.log(alpha) * u          # u^(d/2), d = 2
    if kind == "r_t":
        return -math.log(alpha) * math.sqrt(u)
    return -math.log(alpha) / u


def log_prior_and_grad(theta, priors):
    """Sum of hyper-prior log densities in theta order and its gradient."""
    lp = 0.0
    g = np.zeros(len(priors))
    for i, p in enumerate(priors):
        th = float(theta[i])
        if p.kind == "penalized_complexity":
            kind = _PC_KIND[i] if i < 4 else "sigma_st"
            lam = pc_lambda(kind, p.alpha, p.u)
            if kind == "r_s":
                lp += math.log(lam) - lam * math.exp(-th) - th
                g[i] = lam * math.exp(-th) - 1.0
            elif kind == "r_t":
                lp += math.log(lam) - lam * math.exp(-0.5 * th) + math.log(0.5) - 0.5 * th
                g[i] = 0.5 * lam * math.exp(-0.5 * th) - 0.5
            else:
                lp += math.log(lam) - lam * math.exp(th) + th
                g[i] = 1.0 - lam * math.exp(th)
        elif p.kind == "gaussian":
            lp += -0.5 * p.precision * (th - p.mean) ** 2
            g[i] = -p.precision * (th - p.mean)
        else:
            raise ValueError(f"unsupported hyper-prior {p.kind!r}")
    return lp, g


def likelihood_eta0(tau: float, n_obs: int, yty: float) -> float:
    """0.5 tau n - 0.5 e^tau y^T y (no 2 pi constant)."""
    return 0.5 * tau * n_obs - 0.5 * math.exp(tau) * yty
