"""Device-resident, theta-independent data of the spatio-temporal Gaussian model.

The reference rebuilds `Q_prior` with scipy `kron` and `Q_conditional` with a
sparse product `A^T D A` on every f(theta) evaluation
(`submodels/spatio_temporal.py:114-149`, `core/model.py:384-423`).  Here every
theta-independent piece is prepared once and kept in HBM:

* the union CSR pattern of c0, g1, g2, g3 with the four value arrays;
* the diagonals / sub-diagonals of the tridiagonal temporal matrices;
* ``A_st^T A_st`` (CSR, global columns), ``X^T A_st`` (dense) and ``X^T X``;
* ``A^T y`` and ``y^T y``;
* the eigenvalues of ``c0^-1/2 g1 c0^-1/2`` for the Kronecker-eigen prior
  log-determinant (valid when c0 is diagonal, g2 = g1 c0^-1 g1 and
  g3 = g1 c0^-1 g2, which `closed_form_ok` checks).

Per evaluation only the 12 Kronecker coefficients and exp(theta_obs) change;
`dh_assemble` writes the dense BTA blocks from these arrays.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import scipy.sparse as sps
import torch

from dalia_amd import _lib
from dalia_amd.hyper import kron_coefficients


def _csr(m) -> sps.csr_matrix:
    m = sps.csr_matrix(m)
    m.sum_duplicates()
    m.sort_indices()
    return m


def fem_union(c0, g1, g2, g3):
    """Union pattern (CSR) of the four FEM matrices and their values on it."""
    mats = [_csr(x) for x in (c0, g1, g2, g3)]
    ns = mats[0].shape[0]
    P = _csr(sum(abs(m) for m in mats))
    prow = np.repeat(np.arange(ns, dtype=np.int64), np.diff(P.indptr))
    keys = prow * ns + P.indices.astype(np.int64)
    vals = np.zeros((4, P.nnz))
    for k, m in enumerate(mats):
        mrow = np.repeat(np.arange(ns, dtype=np.int64), np.diff(m.indptr))
        mk = mrow * ns + m.indices.astype(np.int64)
        pos = np.searchsorted(keys, mk)
        vals[k, pos] = m.data
    return P.indptr.astype(np.int32), P.indices.astype(np.int32), vals


def tridiag_parts(m0, m1, m2, nt: int):
    diag = np.zeros((3, nt))
    off = np.zeros((3, max(nt - 1, 0)))
    for k, m in enumerate((m0, m1, m2)):
        d = sps.csr_matrix(m).toarray() if nt <= 4096 else None
        if d is None:
            raise ValueError("nt too large for the dense tridiagonal check")
        if np.any(np.triu(d, 2)) or np.any(np.tril(d, -2)):
            raise ValueError("temporal matrices must be tridiagonal")
        if not np.allclose(d, d.T):
            raise ValueError("temporal matrices must be symmetric")
        diag[k] = np.diagonal(d)
        off[k] = np.diagonal(d, -1)
    return diag, off


def closed_form_ok(c0, g1, g2, g3, rtol: float = 1e-10) -> bool:
    c0 = sps.csr_matrix(c0)
    if abs(c0 - sps.diags(c0.diagonal())).sum() != 0.0:
        return False
    ci = sps.diags(1.0 / c0.diagonal())
    g2c = g1 @ ci @ g1
    g3c = g1 @ ci @ g2
    ok2 = abs(g2c - g2).max() <= rtol * abs(g2).max()
    ok3 = abs(g3c - g3).max() <= rtol * abs(g3).max()
    return bool(ok2 and ok3)


class DeviceSTModel:
    """Uploads one `synth.STModelData` (or loaded DALIA inputs) to the device."""

    def __init__(self, data, device: str | torch.device = "cuda", eigen: bool = True):
        self.data = data
        self.device = torch.device(device)
        ns, nt, nb = data.ns, data.nt, data.nb
        self.ns, self.nt, self.nb = ns, nt, nb
        self.n_latent = ns * nt + nb
        self.n_obs = data.n_obs
        dev = self.device

        def up(x, dtype=torch.float64):
            return torch.as_tensor(np.ascontiguousarray(x)).to(dtype=dtype, device=dev)

        rp, col, val = fem_union(data.c0, data.g1, data.g2, data.g3)
        self.fem_rowptr = up(rp, torch.int32)
        self.fem_col = up(col, torch.int32)
        self.fem_val = up(val)
        self.fem_nnz = int(col.size)
        self.m_diag_h, self.m_off_h = tridiag_parts(data.m0, data.m1, data.m2, nt)
        self.m_diag = up(self.m_diag_h)
        self.m_off = up(self.m_off_h)
        A_st = _csr(data.a_st)
        X = np.ascontiguousarray(data.x_cov, dtype=np.float64)
        AtA = _csr(A_st.T @ A_st)
        self.ata_rowptr = up(AtA.indptr.astype(np.int64), torch.int64)
        self.ata_col = up(AtA.indices.astype(np.int32), torch.int32)
        self.ata_val = up(AtA.data)
        self.ata_nnz = int(AtA.nnz)
        self.ata_arrow = up(np.ascontiguousarray((A_st.T @ X).T))
        self.ata_tip = up(X.T @ X)
        aty = np.concatenate([A_st.T @ data.y, X.T @ data.y])
        self.aty = up(aty)
        self.yty = float(data.y @ data.y)
        self.fe_prec = float(data.fe_prec)
        self.logdet_c0 = float(np.sum(np.log(sps.csr_matrix(data.c0).diagonal())))
        self.closed_form = closed_form_ok(data.c0, data.g1, data.g2, data.g3) if eigen else False
        self.lam = None
        if self.closed_form:
            c = sps.csr_matrix(data.c0).diagonal()
            s = 1.0 / np.sqrt(c)
            S = (sps.diags(s) @ data.g1 @ sps.diags(s)).toarray()
            self.lam = up(np.linalg.eigvalsh(0.5 * (S + S.T)))
        self.struct = _lib.DhModel(
            ns=ns, nt=nt, nb=nb, fem_nnz=self.fem_nnz,
            fem_rowptr=self.fem_rowptr.data_ptr(), fem_col=self.fem_col.data_ptr(),
            fem_val=self.fem_val.data_ptr(),
            m_diag=self.m_diag.data_ptr(), m_off=self.m_off.data_ptr(),
            ata_nnz=self.ata_nnz, ata_rowptr=self.ata_rowptr.data_ptr(),
            ata_col=self.ata_col.data_ptr(), ata_val=self.ata_val.data_ptr(),
            ata_arrow=self.ata_arrow.data_ptr(), ata_tip=self.ata_tip.data_ptr(),
            fe_prec=self.fe_prec,
        )

    def ref(self):
        return ctypes.byref(self.struct)

    @staticmethod
    def coef_array(theta, manifold: str = "plane") -> np.ndarray:
        return np.ascontiguousarray(kron_coefficients(theta, manifold).ravel(), dtype=np.float64)

    def block_weights(self, dc: np.ndarray):
        """Per-block FEM weights w[q][m] for coefficient array dc[k][m].

        q < nt: diagonal block (t, t); q >= nt: sub-diagonal block (t+1, t).
        """
        wd = self.m_diag_h.T @ dc  # (nt, 4)
        wo = self.m_off_h.T @ dc   # (nt-1, 4)
        return np.concatenate([wd, wo], 0)

    def tau_scale(self, theta) -> float:
        return math.exp(float(theta[-1]))
