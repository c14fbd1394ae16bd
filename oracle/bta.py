"""CPU oracle for the block-tridiagonal-arrowhead (BTA) kernels.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (`dalia_amd/`) imports
this module; only `tests/`, `__graft_entry__.smoke()` and the CPU-baseline
leg of `bench.py` use it, as the checker.

The reference calls these routines from the external Serinv package
(`src/dalia/solvers/structured_solver.py:13`, calls at `:83-93`, `:111-140`,
`:175-186`); Serinv is not vendored and its pin is an unversioned VCS URL
(`pyproject.toml:40`), so this is a restatement from the mathematics of the
block Cholesky of a block-tridiagonal-with-arrowhead SPD matrix, in the same
buffer convention as `SerinvSolver.__init__` (`structured_solver.py:38-60`):

* ``D``  (n, b, b)   diagonal blocks            A_ii
* ``L``  (n-1, b, b) sub-diagonal blocks        A_{i+1,i}
* ``R``  (n, a, b)   arrowhead bottom blocks    A_{a,i}
* ``T``  (a, a)      arrowhead tip              A_aa

All routines work in place and in float64 (LAPACK through scipy).

Pinning: `tests/test_oracle_bta.py` checks every routine against dense
numpy on the assembled matrix (cholesky / slogdet / solve / inv) and against
the invariants L L^T = A and Q X = I on the pattern.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla


def _chol(a: np.ndarray) -> np.ndarray:
    return sla.cholesky(a, lower=True, check_finite=False)


def _rtrsm(x: np.ndarray, l: np.ndarray) -> np.ndarray:
    """x @ inv(l).T for lower-triangular l."""
    return sla.solve_triangular(l, x.T, lower=True, check_finite=False).T


def pobtaf(D, L, R, T):
    """In-place BTA Cholesky (the `pobtaf` call of `structured_solver.py:83-88`)."""
    n = D.shape[0]
    arrow = R is not None and R.shape[1] > 0
    for i in range(n):
        D[i] = _chol(D[i])
        if i < n - 1:
            L[i] = _rtrsm(L[i], D[i])
        if arrow:
            R[i] = _rtrsm(R[i], D[i])
        if i < n - 1:
            D[i + 1] -= L[i] @ L[i].T
            if arrow:
                R[i + 1] -= R[i] @ L[i].T
        if arrow:
            T -= R[i] @ R[i].T
    if arrow:
        T[:] = _chol(T)


def pobtf(D, L):
    """In-place block-tridiagonal Cholesky without arrowhead (`structured_solver.py:89-93`)."""
    pobtaf(D, L, None, None)


def pobtas(D, L, R, T, rhs, trans: str):
    """In-place forward ("N") or backward ("C") block triangular solve.

    `rhs` is the flat vector of length n*b + a, as `structured_solver.py:111-127`
    passes it.
    """
    n, b = D.shape[0], D.shape[1]
    arrow = R is not None and R.shape[1] > 0
    a = R.shape[1] if arrow else 0
    x = rhs[: n * b].reshape(n, b)
    xa = rhs[n * b: n * b + a]
    if trans == "N":
        for i in range(n):
            if i > 0:
                x[i] -= L[i - 1] @ x[i - 1]
            x[i] = sla.solve_triangular(D[i], x[i], lower=True, check_finite=False)
        if arrow:
            for i in range(n):
                xa -= R[i] @ x[i]
            xa[:] = sla.solve_triangular(T, xa, lower=True, check_finite=False)
    elif trans == "C":
        if arrow:
            xa[:] = sla.solve_triangular(T, xa, lower=True, trans="T", check_finite=False)
        for i in range(n - 1, -1, -1):
            if i < n - 1:
                x[i] -= L[i].T @ x[i + 1]
            if arrow:
                x[i] -= R[i].T @ xa
            x[i] = sla.solve_triangular(D[i], x[i], lower=True, trans="T", check_finite=False)
    else:
        raise ValueError(f"unknown trans {trans!r}")
    return rhs


def pobts(D, L, rhs, trans: str):
    return pobtas(D, L, None, None, rhs, trans)


def logdet(D, T=None) -> float:
    """2 * sum log diag of the factor (`structured_solver.py:151-166`)."""
    s = 0.0
    for i in range(D.shape[0]):
        s += float(np.sum(np.log(np.diagonal(D[i]))))
    if T is not None and T.shape[0] > 0:
        s += float(np.sum(np.log(np.diagonal(T))))
    if np.isnan(s):
        raise ValueError("Logdet is NaN. Check the input matrix.")
    return 2.0 * s


def pobtasi(D, L, R, T):
    """In-place selected inversion on the BTA pattern (Takahashi recurrences).

    On entry the buffers hold the factor from `pobtaf`; on exit they hold the
    blocks X_ii, X_{i+1,i}, X_{a,i}, X_aa of the inverse (diagonal blocks as
    full symmetric matrices).
    """
    n, b = D.shape[0], D.shape[1]
    arrow = R is not None and R.shape[1] > 0
    if arrow:
        Ti = sla.solve_triangular(T, np.eye(T.shape[0]), lower=True, check_finite=False)
        Xaa = Ti.T @ Ti
    for i in range(n - 1, -1, -1):
        Li = sla.solve_triangular(D[i], np.eye(b), lower=True, check_finite=False)
        if i < n - 1:
            W = L[i] @ Li  # L_{i+1,i} L_ii^-1
        if arrow:
            Wa = R[i] @ Li
        # X_{a,i}
        if arrow:
            Xai = -Xaa @ Wa
            if i < n - 1:
                Xai -= R[i + 1] @ W  # R[i+1] already holds X_{a,i+1}
        # X_{i+1,i}
        if i < n - 1:
            Xli = -D[i + 1] @ W  # D[i+1] holds X_{i+1,i+1}
            if arrow:
                Xli -= R[i + 1].T @ Wa
        Xii = Li.T @ Li
        if i < n - 1:
            Xii -= W.T @ Xli
        if arrow:
            Xii -= Wa.T @ Xai
        Xii = 0.5 * (Xii + Xii.T)
        if i < n - 1:
            L[i] = Xli
        if arrow:
            R[i] = Xai
        D[i] = Xii
    if arrow:
        T[:] = Xaa


def pobtsi(D, L):
    pobtasi(D, L, None, None)


# ---------------------------------------------------------------------------
# dense helpers for brute-force pinning
# ---------------------------------------------------------------------------

def bta_to_dense(D, L, R=None, T=None, lower_only: bool = False) -> np.ndarray:
    """Dense symmetric matrix from BTA blocks (diagonal blocks taken as given)."""
    n, b = D.shape[0], D.shape[1]
    a = 0 if R is None else R.shape[1]
    N = n * b + a
    M = np.zeros((N, N))
    for i in range(n):
        M[i * b:(i + 1) * b, i * b:(i + 1) * b] = D[i]
        if i < n - 1:
            M[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = L[i]
            if not lower_only:
                M[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = L[i].T
        if a:
            M[n * b:, i * b:(i + 1) * b] = R[i]
            if not lower_only:
                M[i * b:(i + 1) * b, n * b:] = R[i].T
    if a:
        M[n * b:, n * b:] = T
    return M


def dense_to_bta(M: np.ndarray, n: int, b: int, a: int):
    """Slice a dense matrix into BTA blocks (copies)."""
    D = np.stack([M[i * b:(i + 1) * b, i * b:(i + 1) * b] for i in range(n)]).copy()
    L = np.stack([M[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] for i in range(n - 1)]).copy() \
        if n > 1 else np.zeros((0, b, b))
    R = np.stack([M[n * b:n * b + a, i * b:(i + 1) * b] for i in range(n)]).copy() if a else None
    T = M[n * b:, n * b:].copy() if a else None
    return D, L, R, T
