"""Host-side scalar algebra of the product against the oracle."""

import numpy as np
import pytest

from dalia_amd import hyper, synth
from dalia_amd.model import fem_union, tridiag_parts, closed_form_ok
from oracle import inla


@pytest.mark.parametrize("th", [[-1.0, 0.7, 0.0, 1.5], [0.4, -0.3, 0.8, -0.2]])
def test_kron_coefficients_and_grad(th):
    th = np.array(th)
    assert np.allclose(hyper.kron_coefficients(th), inla.kron_coefficients(th), rtol=1e-14)
    assert np.allclose(hyper.kron_coefficients_grad(th), inla.kron_coefficients_grad(th), rtol=1e-13)
    e = 1e-6
    for j in range(3):
        tp, tm = th.copy(), th.copy()
        tp[j] += e
        tm[j] -= e
        num = (hyper.kron_coefficients(tp) - hyper.kron_coefficients(tm)) / (2 * e)
        assert np.allclose(hyper.kron_coefficients_grad(th)[j], num, rtol=1e-6, atol=1e-8)


def test_log_prior_matches_oracle():
    pri = synth.DEFAULT_PRIORS
    for th in ([-1.0, 0.7, 0.0, 1.5], [0.4, -0.3, 0.8, -0.2]):
        lp, g = hyper.log_prior_and_grad(th, pri)
        assert lp == pytest.approx(inla.log_prior(th, pri), rel=1e-14)
        assert np.allclose(g, inla.log_prior_grad(th, pri), rtol=1e-13)


def test_fem_union_values():
    d = synth.make_st_model(6, 3, 2, 1)
    rp, col, val = fem_union(d.c0, d.g1, d.g2, d.g3)
    for k, G in enumerate((d.c0, d.g1, d.g2, d.g3)):
        import scipy.sparse as sps
        M = sps.csr_matrix((val[k], col, rp), shape=G.shape)
        assert abs(M - G).max() == 0.0
    dg, off = tridiag_parts(d.m0, d.m1, d.m2, d.nt)
    assert np.allclose(dg[0], d.m0.diagonal()) and np.allclose(off[2], d.m2.diagonal(-1))
    assert closed_form_ok(d.c0, d.g1, d.g2, d.g3)
    assert not closed_form_ok(d.c0, d.g1, d.g2 * 1.01, d.g3)


@pytest.mark.parametrize("th", [[-1.0, 0.7, 0.0, 1.5], [0.4, -0.3, 0.8, -0.2], [-2.5, 1.0, -0.3, 0.0]])
def test_sphere_map_coefficients_and_grad(th):
    """Sphere branch (`spatio_temporal.py:171-180`): product against the oracle, and its
    gradient (an analytic series derivative) against central differences and against the
    oracle's complex-step Jacobian."""
    th = np.array(th)
    for a, b in zip(hyper.interpretable2compute(*th[:3], manifold="sphere"),
                    inla.interpretable2compute(*th[:3], manifold="sphere")):
        assert a == pytest.approx(float(np.real(b)), rel=1e-14, abs=1e-14)
    assert np.allclose(hyper.dgamma(th, "sphere"), inla.dgamma_dtheta(th, "sphere"), rtol=1e-12, atol=1e-14)
    assert np.allclose(hyper.kron_coefficients(th, "sphere"), inla.kron_coefficients(th, "sphere"), rtol=1e-14)
    assert np.allclose(hyper.kron_coefficients_grad(th, "sphere"), inla.kron_coefficients_grad(th, "sphere"),
                       rtol=1e-12)
    e = 1e-6
    for j in range(3):
        tp, tm = th.copy(), th.copy()
        tp[j] += e
        tm[j] -= e
        num = (hyper.kron_coefficients(tp, "sphere") - hyper.kron_coefficients(tm, "sphere")) / (2 * e)
        assert np.allclose(hyper.kron_coefficients_grad(th, "sphere")[j], num, rtol=1e-6, atol=1e-8)


def test_sphere_differs_from_plane_and_unknown_manifold_raises():
    th = np.array([-1.0, 0.7, 0.0, 1.5])
    assert not np.allclose(hyper.kron_coefficients(th, "sphere"), hyper.kron_coefficients(th, "plane"))
    with pytest.raises(ValueError):
        hyper.interpretable2compute(*th[:3], manifold="torus")
    with pytest.raises(ValueError):
        inla.interpretable2compute(*th[:3], manifold="torus")


def test_diagonal_hessian_stencils_are_exact_on_polynomials():
    """3-point diagonal stencil exact for quadratics, 5-point for quartics
    (`utils/finite_difference_stencils.py:81-171`); 31 points at d = 15 (1 + 2d)."""
    from dalia_amd.inla import hessian_diag_from_values, hessian_diag_stencil
    rng = np.random.default_rng(0)
    d = 15
    Q = rng.standard_normal((d, d))
    Q = Q + Q.T
    c4 = rng.standard_normal(d)
    th = rng.standard_normal(d)
    f2 = lambda x: 0.5 * x @ Q @ x + x.sum()
    f4 = lambda x: f2(x) + float(np.sum(c4 * x ** 4))
    pts = hessian_diag_stencil(th, 1e-2, 3)
    assert pts.shape == (31, d)
    H3 = hessian_diag_from_values(np.array([f2(p) for p in pts]), d, 1e-2, 3)
    assert np.allclose(H3, np.diag(Q), rtol=1e-8, atol=1e-8)
    pts5 = hessian_diag_stencil(th, 1e-2, 5)
    assert pts5.shape == (61, d)
    H5 = hessian_diag_from_values(np.array([f4(p) for p in pts5]), d, 1e-2, 5)
    assert np.allclose(H5, np.diag(Q) + 12 * c4 * th ** 2, rtol=1e-7, atol=1e-7)
    with pytest.raises(ValueError):
        hessian_diag_stencil(th, 1e-2, 4)


def test_negative_eigenvalue_message(capsys):
    from dalia_amd.inla import check_hessian
    ev = check_hessian(np.diag([2.0, -1.0]))
    assert ev.min() < 0 and "Negative eigenvalues detected" in capsys.readouterr().out
    check_hessian(np.eye(3))
    assert capsys.readouterr().out == ""


@pytest.mark.parametrize("manifold", ["plane", "sphere"])
def test_synthetic_field_is_an_exact_prior_draw(manifold):
    """The data generator draws x_st from N(0, Qp^-1): its linear map B satisfies B B^T = Qp^-1
    (dense inverse of the oracle's assembled prior) on a small model."""
    import dataclasses
    d = dataclasses.replace(synth.make_st_model(4, 3, 1, 0), manifold=manifold)
    th = np.array([-0.5, 0.3, 0.2, 1.0])
    N = d.ns * d.nt
    B = np.stack([synth.prior_transform(d.c0, d.g1, d.m0, d.m1, d.m2, th, e, manifold) for e in np.eye(N)], axis=1)
    C = np.linalg.inv(inla.qprior_st_sparse(d, th).toarray())
    assert np.linalg.norm(B @ B.T - C) / np.linalg.norm(C) < 1e-12
