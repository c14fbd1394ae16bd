"""Pin the hyperparameter map (r_s, r_t, sigma_st) -> (gamma_s, gamma_t, gamma_st) by mathematics.

The reference's map (`src/dalia/submodels/spatio_temporal.py:151-197`) is meant to make sigma_st
the log marginal standard deviation, r_s the log spatial range and r_t the log temporal range of
the separable-in-spirit diffusion field whose precision is

    Q = e^{2 gst} [M0 (x) q3 + e^{gt} M1 (x) q2 + e^{2 gt} M2 (x) q1],   q_p ~ (kappa^2 - Laplacian)^p,

(`:114-149`).  In the interior (M0 ~ identity in time, M2 ~ -d^2/dt^2, M1 a boundary term;
G_m ~ |k|^{2m} in space) the symbol is A(k) + w^2 B(k) with A = sum_m c[0][m] |k|^{2m} and
B = sum_m c[2][m] |k|^{2m}, c the Kronecker coefficients.  Integrating over the temporal
frequency w leaves 1 / (2 sqrt(A B)) per spatial frequency, so on the plane

* the marginal variance is int d^2k / (2 pi)^2 1 / (2 sqrt(A B)) and must equal e^{2 sigma_st};
* the temporal correlation at lag dt is the same integral weighted by exp(-sqrt(A / B) dt), and
  must equal E_2(2 dt / rho_t) (E_2 the exponential integral, rho_t = e^{r_t});
* the spatial correlation at distance d is the integral weighted by J_0(|k| d), a Matern
  correlation of smoothness 1, and must equal x K_1(x) with x = sqrt(8) d / rho_s.

These fix the constant c1 = Gamma(1/2) Gamma(1) / (Gamma(1) Gamma(2) (4 pi)^1.5), the gamma_t
shift and the gamma_s constant independently of the reference's text.  The tests below also
plant plausible mistakes in the oracle's map and check that each one fails.
"""

import math

import numpy as np
import pytest
from scipy import integrate, special

from oracle import inla

THETAS = [np.array([-1.0, 0.7, 0.0, 1.5]), np.array([-0.4, 1.6, 0.3, 0.2]), np.array([-2.0, 0.1, -0.5, 0.0])]


def _symbol(c):
    def AB(k):
        pw = np.array([1.0, k * k, k ** 4, k ** 6])
        return float(c[0] @ pw), float(c[2] @ pw)
    return AB


def plane_variance(c) -> float:
    AB = _symbol(c)
    return integrate.quad(lambda k: k / (2 * math.pi) / (2 * math.sqrt(AB(k)[0] * AB(k)[1])), 0, np.inf,
                          limit=500)[0]


def plane_temporal_corr(c, dt: float) -> float:
    AB = _symbol(c)

    def f(k):
        A, B = AB(k)
        return k / (2 * math.pi) * math.exp(-math.sqrt(A / B) * dt) / (2 * math.sqrt(A * B))
    return integrate.quad(f, 0, np.inf, limit=500)[0] / plane_variance(c)


def plane_spatial_corr(c, d: float) -> float:
    AB = _symbol(c)

    def f(k):
        A, B = AB(k)
        return k / (2 * math.pi) * special.j0(k * d) / (2 * math.sqrt(A * B))
    z = special.jn_zeros(0, 400) / d  # integrate between the zeros of J_0
    tot = integrate.quad(f, 0, z[0])[0] + sum(integrate.quad(f, a, b)[0] for a, b in zip(z[:-1], z[1:]))
    return tot / plane_variance(c)


def _checks(theta, coef_fn):
    c = coef_fn(theta)
    rho_s, rho_t, sigma = math.exp(theta[0]), math.exp(theta[1]), math.exp(theta[2])
    x = math.sqrt(8.0)
    return {
        "variance": (plane_variance(c), sigma ** 2),
        "temporal corr at rho_t": (plane_temporal_corr(c, rho_t), special.expn(2, 2.0)),
        "temporal corr at rho_t/2": (plane_temporal_corr(c, 0.5 * rho_t), special.expn(2, 1.0)),
        "spatial corr at rho_s": (plane_spatial_corr(c, rho_s), x * special.kv(1, x)),
    }


@pytest.mark.parametrize("theta", THETAS, ids=range(len(THETAS)))
def test_plane_map_gives_marginal_variance_and_ranges(theta):
    for name, (got, want) in _checks(theta, inla.kron_coefficients).items():
        assert got == pytest.approx(want, rel=1e-7), name


def _mutant(kind):
    """The oracle's plane map with one planted mistake."""
    def m(r_s, r_t, sigma_st, manifold="plane"):
        nu_s, nu_t, alpha, alpha_t = 1.0, 0.5, 2.0, 1.0
        gamma_s = 0.5 * math.log(8 * nu_s) - r_s
        if kind == "gamma_s constant log 4":
            gamma_s = 0.5 * math.log(4.0) - r_s
        gamma_t = r_t - 0.5 * math.log(8 * nu_t) + 2 * gamma_s
        if kind == "gamma_t shift log 8":
            gamma_t = r_t - 0.5 * math.log(8.0) + 2 * gamma_s
        scal = pow(4 * math.pi, 1.0 if kind == "c1 with (4 pi)^1" else 1.5)
        num = special.gamma(nu_t) * special.gamma(nu_s)
        if kind == "c1 without Gamma(nu_t)":
            num = special.gamma(nu_s)
        c1 = num / (special.gamma(alpha_t) * special.gamma(alpha) * scal)
        gamma_st = 0.5 * math.log(c1) - 0.5 * gamma_t - nu_s * gamma_s - sigma_st
        if kind == "gamma_st without the gamma_t term":
            gamma_st = 0.5 * math.log(c1) - nu_s * gamma_s - sigma_st
        if kind == "sigma_st sign":
            gamma_st = 0.5 * math.log(c1) - 0.5 * gamma_t - nu_s * gamma_s + sigma_st
        return gamma_s, gamma_t, gamma_st
    return m


@pytest.mark.parametrize("kind", ["gamma_s constant log 4", "gamma_t shift log 8", "c1 with (4 pi)^1",
                                  "c1 without Gamma(nu_t)", "gamma_st without the gamma_t term",
                                  "sigma_st sign"])
def test_planted_mistakes_fail(kind, monkeypatch):
    theta = np.array([-0.4, 1.6, 0.3, 0.2])
    monkeypatch.setattr(inla, "interpretable2compute", _mutant(kind))
    res = _checks(theta, inla.kron_coefficients)
    failed = [n for n, (got, want) in res.items() if abs(got - want) > 1e-3 * abs(want)]
    assert failed, f"planted mistake {kind!r} passed every check"


def test_unmutated_copy_passes(monkeypatch):
    """The mutant generator without a mutation reproduces the oracle's map."""
    monkeypatch.setattr(inla, "interpretable2compute", _mutant("none"))
    theta = THETAS[1]
    for name, (got, want) in _checks(theta, inla.kron_coefficients).items():
        assert got == pytest.approx(want, rel=1e-7), name


# ---------------------------------------------------------------------------
# sphere: the spatial spectrum is k (k + 1) with multiplicity 2k + 1 and sum_m |Y_km|^2 =
# (2k + 1) / (4 pi), so the marginal variance of the unit-sphere field is
# sum_k (2k + 1) / (4 pi) / (2 sqrt(A_k B_k)), A_k, B_k the symbol at |k|^2 -> k (k + 1).
# The reference's series stops at 50 terms (`spatio_temporal.py:175-178`), so the variance equals
# e^{2 sigma_st} times the ratio of the full series to the 50-term one.
# ---------------------------------------------------------------------------

def sphere_variance(c, kmax: int = 200000) -> float:
    k = np.arange(kmax, dtype=np.float64)
    L = k * (k + 1.0)
    pw = np.stack([np.ones_like(L), L, L ** 2, L ** 3])
    A, B = c[0] @ pw, c[2] @ pw
    return float(np.sum((2 * k + 1) / (4 * math.pi) / (2 * np.sqrt(A * B))))


def _series_ratio(theta, terms: int = 50, kmax: int = 200000) -> float:
    gs = 0.5 * math.log(8.0) - theta[0]
    k = np.arange(kmax, dtype=np.float64)
    w = (2 * k + 1) / (4 * math.pi * (math.exp(2 * gs) + k * (k + 1)) ** 2)
    return float(np.sum(w) / np.sum(w[:terms]))


@pytest.mark.parametrize("theta", THETAS, ids=range(len(THETAS)))
def test_sphere_map_gives_marginal_variance(theta):
    c = inla.kron_coefficients(theta, "sphere")
    want = math.exp(2 * theta[2]) * _series_ratio(theta)
    assert sphere_variance(c) == pytest.approx(want, rel=1e-9)


def _sphere_mutant(kind):
    def m(r_s, r_t, sigma_st, manifold="sphere"):
        gamma_s = 0.5 * math.log(8.0) - r_s
        gamma_t = r_t - 0.5 * math.log(4.0) + 2 * gamma_s
        cr_t = special.gamma(0.5) / (special.gamma(1.0) * pow(4 * math.pi, 1.5 if kind == "cR_t power" else 0.5))
        power = 1.0 if kind == "series power" else 2.0
        c_s = sum((2.0 * k + 1.0) / (4.0 * math.pi * (math.exp(2 * gamma_s) + k * (k + 1)) ** power)
                  for k in range(50))
        half_t = 0.5 if kind == "gamma_t sign" else -0.5
        gamma_st = 0.5 * math.log(cr_t) + 0.5 * math.log(c_s) + half_t * gamma_t - sigma_st
        return gamma_s, gamma_t, gamma_st
    return m


@pytest.mark.parametrize("kind", ["cR_t power", "series power", "gamma_t sign"])
def test_sphere_planted_mistakes_fail(kind, monkeypatch):
    theta = np.array([-0.4, 1.6, 0.3, 0.2])
    monkeypatch.setattr(inla, "interpretable2compute", _sphere_mutant(kind))
    c = inla.kron_coefficients(theta, "sphere")
    want = math.exp(2 * theta[2]) * _series_ratio(theta)
    assert abs(sphere_variance(c) - want) > 1e-3 * want
    monkeypatch.setattr(inla, "interpretable2compute", _sphere_mutant("none"))
    assert sphere_variance(inla.kron_coefficients(theta, "sphere")) == pytest.approx(want, rel=1e-9)
