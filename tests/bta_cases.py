"""Seeded SPD block-tridiagonal-arrowhead test matrices (shared by the tests)."""

import numpy as np
import scipy.sparse as sps


def random_bta(n: int, b: int, a: int, seed: int, dominance: float = 1.0):
    """Dense SPD matrix with the BTA pattern, as (dense, scipy CSR)."""
    rng = np.random.default_rng(seed)
    N = n * b + a
    M = np.zeros((N, N))
    for i in range(n):
        B = rng.standard_normal((b, b)) / np.sqrt(b)
        M[i * b:(i + 1) * b, i * b:(i + 1) * b] = B @ B.T
        if i < n - 1:
            C = rng.standard_normal((b, b)) / np.sqrt(b)
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
