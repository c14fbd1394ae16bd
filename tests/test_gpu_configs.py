"""GPU parity at the benchmark configurations themselves.

* ns=400, nt=12, nb=4 (the CPU-runnable case): f, mu, both log-determinants, the adjoint gradient
  and the marginal variances against the committed oracle fixture.
* ns=2025, nt=100, nb=6 (the plain drop-in and adjoint case), full size: the same quantities
  against a committed oracle fixture (scalars, gradient, a strided sample of mu, the diagonals of
  six selected-inverse blocks), and the BT log-determinant of the prior against its
  Kronecker-eigen closed form.
* ns=4096, nt=250, nb=6 (the large adjoint case, 67 GB of blocks): size-independent checks, the
  prior's BT log-determinant against the closed form, the adjoint gradient against GPU central
  differences on two theta components, the two-ended engine against the one-ended one and the
  checkpointed engine against the one-ended one.
* the bivariate coregional instance ns=576, nt=24 (block size 1152) against the oracle, and the
  trivariate ns=1600, nt=128 (block size 4800) adjoint gradient against GPU central differences.

Bars: f and log-determinants 1e-10 relative, mu and selected-inverse entries 1e-9 relative,
adjoint gradient 1e-8 relative against the oracle's adjoint and 1e-6 relative (to the gradient's
largest component) against central differences with h = 1e-4 (1e-4 at ns=4096, nt=250, where the
quotient is noise-limited; see that test): f is a difference of log-determinants ~10 |f|, so the
differences carry a rounding noise of ~noise(f) / 2h in every component, whatever its size.
"""

import os

import numpy as np
import pytest

from dalia_amd import synth
from oracle import coreg as ocoreg
from oracle import inla as oinla

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def _model_from_fixture(g):
    d = synth.make_st_model(int(g["m"]), int(g["nt"]), int(g["nb"]), int(g["seed"]))
    # the generator reproduces the fixture's inputs (up to the rounding of the host's LAPACK)
    np.testing.assert_allclose(d.y[:16], g["y_head"], rtol=1e-10, atol=1e-12)
    assert d.y.sum() == pytest.approx(float(g["y_sum"]), rel=1e-10)
    return d


def test_ns400_nt12_against_fixture():
    from dalia_amd.inla import STInla
    g = load("st_ns400_nt12_nb4")
    d = _model_from_fixture(g)
    print(f"case fixture ns={d.ns} nt={d.nt} nb={d.nb}", flush=True)
    eng = STInla(d)
    th = g["thetas"][0]
    f = eng.f(th)
    assert abs(f - g["f"][0]) <= 1e-10 * abs(g["f"][0])
    assert rel(eng.mu.cpu().numpy(), g["mu"][0]) < 1e-9
    s = eng.scal[:5].cpu().numpy()
    assert abs(s[0] - g["logdet_qc"][0]) <= 1e-10 * abs(g["logdet_qc"][0])
    assert abs(s[4] - g["logdet_qp"][0]) <= 1e-10 * abs(g["logdet_qp"][0])
    f2, grad = eng.value_and_grad(th)
    assert abs(f2 - g["f"][0]) <= 1e-10 * abs(g["f"][0])
    assert rel(grad, g["grad"][0]) < 1e-8
    assert rel(grad, g["grad_fd"][0]) < 1e-6
    mv = eng.marginal_variances_latent(th, np.zeros(d.n_latent))
    assert rel(mv, g["marginal_var"][0]) < 1e-9


@pytest.fixture(scope="module")
def full():
    from dalia_amd.inla import STInla
    g = load("st_ns2025_nt100_nb6")
    d = _model_from_fixture(g)
    print(f"case full-size ns={d.ns} nt={d.nt} nb={d.nb}", flush=True)
    return g, d, STInla(d)


def test_ns2025_nt100_f_mu_logdets(full):
    g, d, eng = full
    th = g["theta"]
    f = eng.f(th)
    print(f"f gpu {f:.15g} oracle {float(g['f']):.15g}", flush=True)
    assert abs(f - g["f"]) <= 1e-10 * abs(g["f"])
    s = eng.scal[:5].cpu().numpy()
    assert abs(s[0] - g["logdet_qc"]) <= 1e-10 * abs(g["logdet_qc"])
    assert abs(s[4] - g["logdet_qp"]) <= 1e-10 * abs(g["logdet_qp"])
    mu = eng.mu.cpu().numpy()
    assert rel(mu[::int(g["mu_stride"])], g["mu_sub"]) < 1e-9
    assert abs(mu.sum() - g["mu_sum"]) <= 1e-9 * np.abs(mu).sum()
    assert abs(np.linalg.norm(mu) - g["mu_norm"]) <= 1e-9 * g["mu_norm"]


def test_ns2025_nt100_prior_logdet_closed_form(full):
    """pobtf's log-determinant of Qp|BT at full size against the Kronecker-eigen closed form,
    evaluated on the host from an eigendecomposition of the FEM matrices."""
    g, d, eng = full
    th = g["theta"]
    eng.f(th)
    ld_bt = float(eng.scal[4].item())
    ld_cf = oinla.prior_logdet_closed_form(d, th)
    ld_cf = ld_cf[0] if isinstance(ld_cf, tuple) else ld_cf
    assert abs(ld_bt - ld_cf) <= 1e-10 * abs(ld_cf), (ld_bt, ld_cf)


def test_ns2025_nt100_adjoint_gradient(full):
    g, d, eng = full
    th = g["theta"]
    f, grad = eng.value_and_grad(th)
    print(f"grad gpu {grad} oracle {g['grad']}", flush=True)
    assert abs(f - g["f"]) <= 1e-10 * abs(g["f"])
    assert rel(grad, g["grad"]) < 1e-8


def test_ns2025_nt100_selected_inverse_blocks(full):
    g, d, eng = full
    th = g["theta"]
    mv = eng.marginal_variances_latent(th, np.zeros(d.n_latent))
    b, n, a = d.ns, d.nt, d.nb
    for k, i in enumerate(g["x_blocks"]):
        assert rel(mv[i * b:(i + 1) * b], g["x_diag"][k]) < 1e-9, i
    assert rel(mv[n * b:], np.diag(g["x_tip"])) < 1e-9
    # the sub-diagonal and arrow blocks of the same selected inverse (the factor was overwritten)
    Xl = eng.solver.A_lower_diagonal_blocks
    lows = [i for i in g["x_blocks"] if i < n - 1]
    for k, i in enumerate(lows):
        assert rel(np.diagonal(Xl[i].cpu().numpy()), g["x_lower_diag"][k]) < 1e-9, i
    ar = eng.solver.A_arrow_bottom_blocks.cpu().numpy().sum(axis=(1, 2))
    assert rel(ar, g["x_arrow_sum"]) < 1e-8
    assert rel(eng.solver.A_arrow_tip_block.cpu().numpy(), g["x_tip"]) < 1e-9


def _central_difference(fun, theta, j, h=1e-4):
    e = np.zeros(theta.size)
    e[j] = h
    return (fun(theta + e) - fun(theta - e)) / (2 * h)


@pytest.mark.timeout(900)
def test_ns4096_nt250_size_checks():
    """The large adjoint case (67 GB of blocks, 104 GB handle) on one GPU: the two-ended engine
    (the default) against the closed-form prior log-determinant and central differences, the
    one-ended engine against it, and the checkpointed engine against the one-ended one.

    Central differences are noise-limited here: f (|f| ~ 1e6, 1.0e6 unknowns) carries ~1e-6 of
    rounding noise, so the h = 1e-4 quotient is good to ~1e-2 (measured: the two engines' quotients
    differ by 1.3e-3, their adjoint gradients by 4.6e-10 relative); bar 1e-4 of the largest
    component.  Engines with different elimination orders agree to 1e-11 (f) and 2e-9 (gradient),
    the one-ended and the checkpointed one (same order, sums over segments in another order) to
    1e-12 and 5e-10 (measured 1.2e-10)."""
    import torch

    from dalia_amd.ckpt import STInlaCkpt
    from dalia_amd.inla import STInla
    d = synth.make_workload("st_ns4096_nt250_nb6")
    print(f"case large ns={d.ns} nt={d.nt} nb={d.nb}", flush=True)
    th = d.theta
    eng = STInla(d)
    assert eng.twisted
    f = eng.f(th)
    ld_bt = float(eng.scal[4].item())
    ld_cf = oinla.prior_logdet_closed_form(d, th)
    ld_cf = ld_cf[0] if isinstance(ld_cf, tuple) else ld_cf
    print(f"f {f:.15g} logdet Qp bt {ld_bt:.15g} closed form {ld_cf:.15g}", flush=True)
    assert abs(ld_bt - ld_cf) <= 1e-10 * abs(ld_cf)
    f2, grad = eng.value_and_grad(th)
    # f2 takes logdet Qp|BT from the closed form, f from pobtf: they differ by half the difference
    # of the two log-determinants (checked above to 1e-10 of logdet Qp, which is ~10 |f| here)
    assert abs(f2 - f) <= 0.5e-10 * abs(ld_cf) + 1e-12 * abs(f)
    for j in (0, 3):
        cd = _central_difference(eng.f, th, j)
        print(f"d/dtheta_{j}: adjoint {grad[j]:.12g} central difference {cd:.12g}", flush=True)
        assert abs(grad[j] - cd) <= 1e-4 * np.abs(grad).max()
    del eng
    torch.cuda.empty_cache()
    one = STInla(d, twisted=False)
    f_o = one.f(th)
    f2_o, grad_o = one.value_and_grad(th)
    print(f"one-ended: f rel. difference {abs(f_o - f) / abs(f):.2e}, gradient {rel(grad_o, grad):.2e}", flush=True)
    assert abs(f_o - f) <= 1e-11 * abs(f) and abs(f2_o - f2) <= 1e-11 * abs(f2)
    assert rel(grad_o, grad) < 2e-9
    del one
    torch.cuda.empty_cache()
    # the factor checkpointed every 16 time blocks (12.7 GB instead of 104 GB)
    ck = STInlaCkpt(d, K=16)
    f_c = ck.f(th)
    assert abs(f_c - f_o) <= 1e-12 * abs(f_o)
    f2_c, grad_c = ck.value_and_grad(th)
    print(f"checkpointed: f {f_c:.15g} gradient rel. difference {rel(grad_c, grad_o):.2e}, "
          f"{ck.memory_bytes() / 1e9:.1f} GB", flush=True)
    assert abs(f2_c - f2_o) <= 1e-12 * abs(f2_o) and rel(grad_c, grad_o) < 5e-10
    del ck
    torch.cuda.empty_cache()


def test_coreg2_ns576_nt24_against_oracle():
    """One instance of the bivariate batch: f against the oracle, the adjoint gradient against GPU
    central differences in all 9 components."""
    from dalia_amd.coreg import CoregInla
    d = synth.make_coreg_model(*synth.COREG_WORKLOADS["coreg2_ns576_nt24_nb4"])
    print(f"case coreg nv={d.nv} ns={d.ns} nt={d.nt} nb={d.nb}", flush=True)
    eng = CoregInla(d)
    rng = np.random.default_rng(11)
    for k in range(2):
        th = d.theta + (0.0 if k == 0 else 0.05 * rng.standard_normal(d.theta.size))
        f = eng.f(th)
        ref = ocoreg.f_theta_coreg(d, th)
        assert abs(f - ref) <= 1e-10 * abs(ref), (k, f, ref)
    th = d.theta
    f, grad = eng.value_and_grad(th)
    cd = np.array([_central_difference(eng.f, th, j) for j in range(th.size)])
    assert np.abs(grad - cd).max() <= 1e-6 * np.abs(cd).max(), (grad, cd)


@pytest.mark.timeout(900)
def test_coreg3_ns1600_nt128_adjoint_vs_central_differences():
    """The trivariate case (block size 4800, 47 GB of blocks): the adjoint gradient against GPU
    central differences on a range component and a mixing weight."""
    import torch

    from dalia_amd.coreg import CoregInla
    d = synth.make_coreg_model(*synth.COREG_WORKLOADS["coreg3_ns1600_nt128_nb3"])
    print(f"case coreg nv={d.nv} ns={d.ns} nt={d.nt} nb={d.nb} b={d.b}", flush=True)
    eng = CoregInla(d)
    th = d.theta
    f, grad = eng.value_and_grad(th)
    f_plain = eng.f(th)
    ld_p = float(eng.scal[4].item())
    # value_and_grad's prior log-determinant is each variate's closed form, f's is pobtf's
    assert abs(f_plain - f) <= 0.5e-10 * abs(ld_p) + 1e-12 * abs(f), (f_plain, f, ld_p)
    for j in (0, th.size - 1):
        cd = _central_difference(eng.f, th, j)
        print(f"d/dtheta_{j}: adjoint {grad[j]:.12g} central difference {cd:.12g}", flush=True)
        assert abs(grad[j] - cd) <= 1e-6 * np.abs(grad).max()  # the vector bar, see above
    # the 31-point diagonal Hessian stencil in one graph at this size: the centre value against the
    # eager f, two diagonal entries against central differences of the adjoint gradient
    from dalia_amd.batch import batched_hessian_diags
    (hd,), bf = batched_hessian_diags([eng], [th], h=5e-3, points=3)
    print(f"31-point batch done: {len(bf.assignment)} evaluations", flush=True)
    assert len(bf.assignment) == 31
    f_centre = eng._f_from_scalars(th, bf.slots[0].cpu().numpy())  # the replay's first evaluation
    assert abs(f_centre - f_plain) <= 1e-12 * abs(f_plain)
    for j in (0, th.size - 1):
        e = np.zeros(th.size)
        e[j] = 1e-3
        hj = (eng.value_and_grad(th + e)[1][j] - eng.value_and_grad(th - e)[1][j]) / 2e-3
        print(f"H_{j}{j}: stencil {hd[j]:.10g} adjoint differences {hj:.10g}", flush=True)
        assert abs(hd[j] - hj) <= 2e-3 * max(abs(hj), 1e-3 * np.abs(hd).max())
    del eng, bf
    torch.cuda.empty_cache()


def test_coreg2_hess57_batch_in_one_graph():
    """The 57-evaluation batch: the 19-point diagonal Hessian stencils of three bivariate instances of
    uneven size (ns 576, 1024, 1600; nt 24, 48, 96) in one graph replay, each instance's schedule
    chosen by the autotuner.  Every value against an eager evaluation of the same engine (1e-12),
    the smallest instance's values against the oracle (1e-10), and the diagonals against central
    differences of the adjoint gradient for the two smaller instances."""
    from dalia_amd.batch import batched_hessian_diags
    from dalia_amd.coreg import CoregInla
    from dalia_amd.inla import hessian_adjoint, hessian_diag_stencil
    datas = [synth.make_coreg_model(*sp) for sp in synth.COREG_BATCHES["coreg2_hess57"]]
    print("case batch57 instances", [(d.ns, d.nt, d.b) for d in datas], flush=True)
    engines = [CoregInla(d) for d in datas]
    thetas = [d.theta for d in datas]
    diags, bf = batched_hessian_diags(engines, thetas, h=5e-3, points=3, tune=True)
    assert len(bf.assignment) == 57
    pts = [p for t in thetas for p in hessian_diag_stencil(t, 5e-3, 3)]
    fv = bf(pts)
    for e, inst in enumerate(bf.assignment):
        fe = engines[inst].f(pts[e])
        assert abs(fv[e] - fe) <= 1e-12 * abs(fe), (e, inst)
    for e in range(19):
        ref = ocoreg.f_theta_coreg(datas[0], pts[e])
        assert abs(fv[e] - ref) <= 1e-10 * abs(ref), e
    print("batch57 values checked; Hessian diagonals against adjoint central differences", flush=True)
    for k in (0, 1):
        Ha = hessian_adjoint(engines[k], thetas[k], h=1e-3)
        assert np.abs(diags[k] - np.diag(Ha)).max() <= 2e-3 * np.abs(np.diag(Ha)).max(), (diags[k], np.diag(Ha))


def test_coreg3_hessian_diag_31_points_against_oracle():
    """The 31-point diagonal stencil (theta of 15 entries) of a trivariate model in one graph, every
    value against the oracle, at a size the oracle finishes in seconds (ns = 100, nt = 6)."""
    from dalia_amd.batch import batched_hessian_diags
    from dalia_amd.coreg import CoregInla
    from dalia_amd.inla import hessian_diag_stencil
    d = synth.make_coreg_model(10, 6, 3, 3, 21)
    print(f"case 31-point trivariate ns={d.ns} nt={d.nt}", flush=True)
    eng = CoregInla(d)
    diags, bf = batched_hessian_diags([eng], [d.theta], h=5e-3, points=3)
    pts = hessian_diag_stencil(d.theta, 5e-3, 3)
    assert len(pts) == 31
    fv = bf(list(pts))
    ref = np.array([ocoreg.f_theta_coreg(d, p) for p in pts])
    assert np.abs(fv - ref).max() <= 1e-10 * np.abs(ref).max()
    h = 5e-3
    ref_diag = (ref[1::2] - 2 * ref[0] + ref[2::2]) / h ** 2
    assert np.abs(diags[0] - ref_diag).max() <= 1e-5 * np.abs(ref_diag).max()
