"""The committed fixtures: the generator reproduces their inputs and the oracle their outputs."""

import os

import numpy as np
import pytest

from dalia_amd import synth
from oracle import bta, inla

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False))


@pytest.mark.parametrize("name", ["st_tiny_m5_nt4_nb2", "st_ns400_nt12_nb4"])
def test_generator_reproduces_inputs(name):
    g = load(name)
    d = synth.make_st_model(int(g["m"]), int(g["nt"]), int(g["nb"]), int(g["seed"]))
    np.testing.assert_allclose(d.y[:16], g["y_head"], rtol=1e-10, atol=1e-12)  # up to the host LAPACK's rounding
    assert d.y.sum() == pytest.approx(float(g["y_sum"]), rel=1e-10)


def test_oracle_reproduces_tiny_fixture():
    g = load("st_tiny_m5_nt4_nb2")
    d = synth.make_st_model(5, 4, 2, 0)
    for k, th in enumerate(g["thetas"]):
        f, parts = inla.f_theta(d, th, return_parts=True)
        assert f == pytest.approx(g["f"][k], rel=1e-12)
        assert np.allclose(parts["mu"], g["mu"][k], rtol=1e-10, atol=1e-12)
        _, grad = inla.adjoint_gradient(d, th)
        assert np.allclose(grad, g["grad"][k], rtol=1e-9)
    th = g["thetas"][0]
    D, L, R, T = inla.sparse_to_bta(inla.qcond_sparse(d, th), 4, 25, 2, "bta")
    bta.pobtaf(D, L, R, T)
    assert np.allclose(D, g["L_diag"], rtol=1e-12, atol=1e-14)
    bta.pobtasi(D, L, R, T)
    assert np.allclose(L, g["X_lower"], rtol=1e-10, atol=1e-14)


def test_fixture_gradients_agree():
    """Adjoint and central-difference gradients stored in the fixtures agree."""
    for name in ("st_tiny_m5_nt4_nb2", "st_ns400_nt12_nb4"):
        g = load(name)
        for k in range(len(g["f"])):
            a, b = g["grad"][k], g["grad_fd"][k]
            assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-6
