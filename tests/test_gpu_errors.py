"""A readiness wait of the persistent factorization that gives up must surface as an error.

The failure is forced from the host (tuning key 5: the give-up flag is raised before the launch),
so every wait returns at once and every kernel exits normally; the factor is then invalid.  The
log-determinant becomes a NaN with the give-up payload, and every driver raises `DhError` instead
of returning a finite, wrong f(theta) — also under graph replay.  The reference raises on a NaN
log-determinant (`src/dalia/solvers/structured_solver.py:163-164`).
"""

import ctypes

import numpy as np
import pytest

from dalia_amd import _lib, synth
from tests.bta_cases import random_bta

pytestmark = pytest.mark.gpu


def _force(h, on: bool):
    _lib.check(_lib.load().dh_set_tuning(h, 5, 1 if on else 0), "dh_set_tuning")


def test_solver_logdet_raises_and_recovers():
    from dalia_amd.solver import HipBTASolver
    n, b, a = 3, 200, 4
    print(f"case forced give-up solver n={n} b={b} a={a}", flush=True)
    M, S = random_bta(n, b, a, seed=3)
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    _force(s.h, True)
    s.cholesky(S, "bta")
    info = ctypes.c_int(0)
    assert _lib.load().dh_get_info(s.h, ctypes.byref(info)) == -4
    with pytest.raises(_lib.DhError):
        s.logdet("bta")
    _force(s.h, False)
    s.cholesky(S, "bta")
    assert s.info() == 0
    ld = s.logdet("bta")
    assert abs(ld - np.linalg.slogdet(M)[1]) <= 1e-10 * abs(ld)


def test_stinla_f_and_gradient_raise():
    from dalia_amd.inla import STInla
    from oracle import inla as oinla
    d = synth.make_st_model(8, 4, 3, 0)
    print("case forced give-up STInla ns=64 nt=4", flush=True)
    eng = STInla(d)
    _force(eng.h, True)
    with pytest.raises(_lib.DhError):
        eng.f(d.theta)
    with pytest.raises(_lib.DhError):
        eng.value_and_grad(d.theta)
    _force(eng.h, False)
    f_ref = oinla.f_theta(d, d.theta)
    assert abs(eng.f(d.theta) - f_ref) <= 1e-10 * abs(f_ref)


def test_graph_replay_raises():
    from dalia_amd.batch import BatchedF
    from dalia_amd.inla import STInla
    d = synth.make_st_model(8, 4, 3, 0)
    print("case forced give-up BatchedF graph", flush=True)
    eng = STInla(d)
    _force(eng.h, True)  # read when the graph is captured
    bf = BatchedF([eng], [0, 0])
    with pytest.raises(_lib.DhError):
        bf([d.theta, d.theta + 0.01])


def test_coregional_f_raises():
    from dalia_amd.coreg import CoregInla
    d = synth.make_coreg_model(6, 4, 2, 2, 1)
    print("case forced give-up CoregInla nv=2", flush=True)
    eng = CoregInla(d)
    _force(eng.h, True)
    with pytest.raises(_lib.DhError):
        eng.f(d.theta)
