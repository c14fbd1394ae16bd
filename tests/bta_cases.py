"""Seeded SPD block-tridiagonal-arrowhead test matrices (shared by the tests)."""

import numpy as np
import scipy.sparse as sps


def random_bta(n: int, b: int, a: int, seed: int, dominance: float = 1.0, first=None):
    """Dense SPD matrix with the BTA pattern, as (dense, scipy CSR).

    first: optional callable (block i, rng) -> array of b ints; row r of the sub-diagonal block
    (i + 1, i) is zero in the columns left of first[r] (b: the whole row), as a banded spatial
    ordering makes it."""
    rng = np.random.default_rng(seed)
    N = n * b + a
    M = np.zeros((N, N))
    for i in range(n):
        B = rng.standard_normal((b, b)) / np.sqrt(b)
        M[i * b:(i + 1) * b, i * b:(i + 1) * b] = B @ B.T
        if i < n - 1:
            C = rng.standard_normal((b, b)) / np.sqrt(b)
            if first is not None:
                C[np.arange(b)[None, :] < np.asarray(first(i, rng))[:, None]] = 0.0
            M[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = 0.5 * C
            M[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = 0.5 * C.T
        if a:
            R = rng.standard_normal((a, b)) / np.sqrt(b)
            M[n * b:, i * b:(i + 1) * b] = 0.3 * R
            M[i * b:(i + 1) * b, n * b:] = 0.3 * R.T
    if a:
        T = rng.standard_normal((a, a))
        M[n * b:, n * b:] = T @ T.T
    # make it SPD: shift by the most negative eigenvalue plus a margin
    ev = np.linalg.eigvalsh(M)
    M += (max(0.0, -ev[0]) + dominance) * np.eye(N)
    return M, sps.csr_matrix(M)


def banded_first(b: int, width: int):
    """first[r] = r - width: the sub-diagonal blocks of a spatial ordering of bandwidth `width`."""
    return lambda i, rng: np.maximum(0, np.arange(b) - width)


def ragged_first(b: int):
    """Uneven rows: a random first column per row (not monotone), and in every other block one
    64-row tile row and a few single rows with no entries at all."""
    def first(i, rng):
        f = rng.integers(0, b + 1, size=b)
        f = np.minimum(f, np.maximum(0, np.arange(b) - rng.integers(0, 3 * 64)))
        if i % 2 == 0 and b >= 3 * 64:
            t = int(rng.integers(1, b // 64))
            f[64 * t: 64 * t + 64] = b
        f[rng.integers(0, b, size=3)] = b
        return f
    return first
