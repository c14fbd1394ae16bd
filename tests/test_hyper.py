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
