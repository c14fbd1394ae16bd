"""GPU: coregional f(theta) (nv = 2, 3) through the HIP path vs the oracle, and batching."""

import numpy as np
import pytest

from dalia_amd import synth
from oracle import coreg as ocoreg
from oracle import inla as oinla

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(4, 3, 2, 2, 1), (5, 4, 1, 3, 2), (6, 5, 2, 2, 3), (4, 6, 3, 3, 4)])
def test_coreg_f_matches_oracle(shape):
    from dalia_amd.coreg import CoregInla
    d = synth.make_coreg_model(*shape)
    print(f"case coreg nv={d.nv} ns={d.ns} nt={d.nt} nb={d.nb}", flush=True)
    eng = CoregInla(d)
    rng = np.random.default_rng(7)
    for k in range(3):
        th = d.theta + (0.0 if k == 0 else 0.1 * rng.standard_normal(d.theta.size))
        f = eng.f(th)
        ref = ocoreg.f_theta_coreg(d, th)
        assert abs(f - ref) <= 1e-10 * abs(ref), (k, f, ref)


def test_coreg_fd_gradient_matches_oracle():
    from dalia_amd.coreg import CoregInla
    d = synth.make_coreg_model(4, 3, 2, 2, 5)
    eng = CoregInla(d)
    f, g = eng.fd_gradient(d.theta, h=1e-4)
    ref = oinla.fd_gradient(d, d.theta, 1e-4, fun=lambda th: ocoreg.f_theta_coreg(d, th))
    assert abs(f - ref[0]) <= 1e-10 * abs(ref[0])
    # both sides difference their own f (agreeing to ~1e-12 relative) over 2h = 2e-4
    assert np.abs(g - ref[1]).max() <= 1e-4 * max(1.0, np.abs(ref[1]).max())


def test_coreg_and_univariate_batched_in_one_graph():
    from dalia_amd.batch import BatchedF
    from dalia_amd.coreg import CoregInla
    from dalia_amd.inla import STInla
    dc = synth.make_coreg_model(5, 4, 2, 2, 6)
    du = synth.make_st_model(6, 4, 3, 1)
    engines = [CoregInla(dc), STInla(du)]
    assignment = [0, 1, 0, 1, 0]
    bf = BatchedF(engines, assignment)
    rng = np.random.default_rng(3)
    for rep in range(2):
        thetas = [(dc.theta if a == 0 else du.theta) + 0.05 * rng.standard_normal(9 if a == 0 else 4)
                  for a in assignment]
        fv = bf(thetas)
        for e, a in enumerate(assignment):
            ref = ocoreg.f_theta_coreg(dc, thetas[e]) if a == 0 else oinla.f_theta(du, thetas[e])
            assert abs(fv[e] - ref) <= 1e-10 * abs(ref), (rep, e)


def _fd4(fun, theta, h):
    """Fourth-order central differences of a CPU function (the oracle)."""
    g = np.zeros(theta.size)
    for j in range(theta.size):
        e = np.zeros(theta.size)
        e[j] = h
        g[j] = (-fun(theta + 2 * e) + 8 * fun(theta + e) - 8 * fun(theta - e) + fun(theta - 2 * e)) / (12 * h)
    return g


@pytest.mark.parametrize("shape", [(4, 3, 2, 2, 11), (5, 4, 1, 3, 12)])
def test_coreg_adjoint_gradient_matches_oracle(shape):
    from dalia_amd.coreg import CoregInla
    d = synth.make_coreg_model(*shape)
    print(f"case coreg adjoint nv={d.nv} ns={d.ns} nt={d.nt} nb={d.nb}", flush=True)
    eng = CoregInla(d)
    rng = np.random.default_rng(5)
    th = d.theta + 0.1 * rng.standard_normal(d.theta.size)
    f, g = eng.value_and_grad(th)
    fun = lambda t: ocoreg.f_theta_coreg(d, t)
    f_ref = fun(th)
    g_ref = _fd4(fun, th, 1e-3)
    # the adjoint's prior term is each variate's closed form; the reference's is logdet Qp|BT
    assert abs(f - f_ref) <= 1e-10 * abs(f_ref), (f, f_ref)
    assert np.abs(g - g_ref).max() <= 1e-6 * max(1.0, np.abs(g_ref).max()), (g, g_ref)
    # the engine is reusable after the selected inversion destroyed the factor
    assert abs(eng.f(th) - f_ref) <= 1e-10 * abs(f_ref)


def test_coreg_autograd_node():
    import torch

    from dalia_amd.coreg import CoregInla
    from dalia_amd.inla import FTheta
    d = synth.make_coreg_model(4, 3, 2, 2, 13)
    eng = CoregInla(d)
    th = torch.tensor(d.theta, dtype=torch.float64, requires_grad=True)
    f = FTheta.apply(th, eng)
    f.backward()
    f_ref, g_ref = eng.value_and_grad(d.theta)
    assert abs(float(f) - f_ref) <= 1e-12 * abs(f_ref)
    assert np.abs(th.grad.numpy() - g_ref).max() <= 1e-10 * max(1.0, np.abs(g_ref).max())


def test_coreg_hessian_from_adjoint_gradients():
    from dalia_amd.coreg import CoregInla
    from dalia_amd.inla import hessian_adjoint
    d = synth.make_coreg_model(4, 3, 1, 2, 17)
    print("case hessian_adjoint coreg", flush=True)
    eng = CoregInla(d)
    H = hessian_adjoint(eng, d.theta, h=1e-3)
    H_ref = oinla.fd_hessian(d, d.theta, 1e-3, fun=lambda th: ocoreg.f_theta_coreg(d, th))
    assert np.abs(H - H_ref).max() <= 1e-4 * np.abs(H_ref).max(), (H, H_ref)


def test_coreg_twisted_autotuned(tmp_path):
    """The coregional engine leaves the one- or two-ended factorization to the autotuner (tuning key
    7): the choice is recorded and re-applied from the file, and f and the adjoint gradient agree with
    the oracle whichever it is, and across both settings to 1e-12 / 1e-10."""
    from dalia_amd import autotune
    from dalia_amd.coreg import CoregInla
    d = synth.make_coreg_model(6, 7, 2, 2, 9)  # b = 72 (two tile columns), nt = 7
    print("case coreg twisted autotune", flush=True)
    eng = CoregInla(d)
    path = str(tmp_path / "tune.json")
    c = autotune.tune(eng, path=path, reps=1)
    assert c.get("tw") in (0, 1)
    assert autotune.tune(eng, path=path)["tw"] == c["tw"]
    ref = ocoreg.f_theta_coreg(d, d.theta)
    out = {}
    for tw in (False, True):
        eng.solver.set_twisted(tw)
        f = eng.f(d.theta)
        assert abs(f - ref) <= 1e-10 * abs(ref), (tw, f, ref)
        out[tw] = eng.value_and_grad(d.theta)
    (f0, g0), (f1, g1) = out[False], out[True]
    assert abs(f1 - f0) <= 1e-12 * abs(f0)
    assert np.linalg.norm(g1 - g0) <= 1e-10 * np.linalg.norm(g0)
