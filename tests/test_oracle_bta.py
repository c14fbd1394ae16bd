"""Pin the CPU oracle's BTA routines against dense LAPACK on the assembled matrix."""

import numpy as np
import pytest

from oracle import bta
from tests.bta_cases import random_bta


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("n,b,a", [(1, 5, 0), (1, 4, 2), (4, 6, 3), (5, 8, 0), (3, 1, 2), (6, 3, 1)])
def test_pobtaf_equals_dense_cholesky(n, b, a):
    M, _ = random_bta(n, b, a, seed=n + 10 * b + 100 * a)
    D, L, R, T = bta.dense_to_bta(M, n, b, a)
    if a == 0:
        R = T = None
    bta.pobtaf(D, L, R, T)
    Lfull = np.linalg.cholesky(M)
    Dref, Lref, Rref, Tref = bta.dense_to_bta(Lfull, n, b, a)
    assert rel(D, Dref) < 1e-12
    if n > 1:
        assert rel(L, Lref) < 1e-12
    if a:
        assert rel(R, Rref) < 1e-12
        assert rel(T, Tref) < 1e-12
    # invariant: L L^T = A
    Lm = bta.bta_to_dense(D, L, R, T, lower_only=True)
    assert rel(Lm @ Lm.T, M) < 1e-12
    assert abs(bta.logdet(D, T) - np.linalg.slogdet(M)[1]) < 1e-10 * max(1.0, abs(np.linalg.slogdet(M)[1]))


@pytest.mark.parametrize("n,b,a", [(1, 5, 2), (4, 6, 3), (5, 8, 0), (3, 1, 2)])
def test_pobtas_equals_dense_solve(n, b, a):
    M, _ = random_bta(n, b, a, seed=3 + n)
    D, L, R, T = bta.dense_to_bta(M, n, b, a)
    if a == 0:
        R = T = None
    bta.pobtaf(D, L, R, T)
    rhs = np.random.default_rng(1).standard_normal(n * b + a)
    y = bta.pobtas(D, L, R, T, rhs.copy(), "N")
    x = bta.pobtas(D, L, R, T, y, "C")
    assert rel(x, np.linalg.solve(M, rhs)) < 1e-12
    # forward sweep alone is L^-1 rhs
    Lm = bta.bta_to_dense(D, L, R, T, lower_only=True)
    y2 = bta.pobtas(D, L, R, T, rhs.copy(), "N")
    assert rel(y2, np.linalg.solve(Lm, rhs)) < 1e-12
    with pytest.raises(ValueError):
        bta.pobtas(D, L, R, T, rhs.copy(), "X")


@pytest.mark.parametrize("n,b,a", [(1, 5, 2), (4, 6, 3), (5, 8, 0), (3, 1, 2), (2, 7, 1)])
def test_pobtasi_equals_dense_inverse(n, b, a):
    M, _ = random_bta(n, b, a, seed=50 + n)
    D, L, R, T = bta.dense_to_bta(M, n, b, a)
    if a == 0:
        R = T = None
    bta.pobtaf(D, L, R, T)
    bta.pobtasi(D, L, R, T)
    Xd, Xl, Xr, Xt = bta.dense_to_bta(np.linalg.inv(M), n, b, a)
    assert rel(D, Xd) < 1e-11
    if n > 1:
        assert rel(L, Xl) < 1e-11
    if a:
        assert rel(R, Xr) < 1e-11
        assert rel(T, Xt) < 1e-11


def test_tridiagonal_closed_form_b1():
    """b = 1, a = 0: the BTA is a tridiagonal matrix whose LDL^T is the textbook recurrence."""
    n = 9
    rng = np.random.default_rng(0)
    d = 3.0 + rng.random(n)
    e = rng.random(n - 1) - 0.5
    M = np.diag(d) + np.diag(e, -1) + np.diag(e, 1)
    D, L, _, _ = bta.dense_to_bta(M, n, 1, 0)
    bta.pobtf(D, L)
    dd = np.zeros(n)
    dd[0] = d[0]
    for i in range(1, n):
        dd[i] = d[i] - e[i - 1] ** 2 / dd[i - 1]
    assert np.allclose(D[:, 0, 0] ** 2, dd, rtol=1e-13)
    assert abs(bta.logdet(D) - np.sum(np.log(dd))) < 1e-12


def test_logdet_nan_raises():
    D = np.array([[[np.nan]]])
    with pytest.raises(ValueError):
        bta.logdet(D)
