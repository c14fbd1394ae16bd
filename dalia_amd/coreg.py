"""f(theta) of the coregional (multivariate) Gaussian model on the HIP BTA path.

Mirrors `models/coregional_model.py` driven by `core/dalia.py:603-745`:
nv = 2 or 3 spatio-temporal variates on one mesh, each with its own
(r_s, r_t) prior field, fixed effects and Gaussian likelihood precision,
coupled by the coregional sigmas and lambdas.  The BTA has blocks of size
nv*ns (variate-major inside every time block) and an arrowhead of nv*nb.

Device work per evaluation (all on the handle's stream, no host sync until
the five scalars are read): coregional assembly of Qc (`dh_assemble_coreg`),
pobtaf, the information vector A^T D y (`dh_scale_indexed`), pobtas,
logdet, the quadratic form, then the BT prior (assembly without the
likelihood, pobtf, logdet).  The quadratic form is mu^T b: with Qc mu = b it
equals the reference's mu^T Qc mu (`core/dalia.py:1161`) up to the solve's
rounding.  Hyper-priors must be Gaussian (the coregional examples use
Gaussian priors throughout, `examples/gst_coreg2_small/run.py:80-200`).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import scipy.sparse as sps
import torch

from dalia_amd import _lib, hyper
from dalia_amd.inla import check_hessian, hessian_from_values, hessian_stencil
from dalia_amd.model import _csr, closed_form_ok, fem_union, tridiag_parts
from dalia_amd.solver import HipBTASolver


def mixing_weights(nv: int, theta) -> np.ndarray:
    """w[i, j, k]: q_ij = sum_k w[i,j,k] Qu_k (`coregional_model.py:304-497`).

    Complex theta is accepted (complex-step derivatives of the weights)."""
    theta = np.asarray(theta)
    dt = np.result_type(theta.dtype, np.float64)
    s2 = np.exp(2.0 * theta[3 * nv:4 * nv].astype(dt))
    lam = theta[4 * nv:].astype(dt)
    w = np.zeros((nv, nv, nv), dtype=dt)
    if nv == 2:
        l01 = lam[0]
        w[0, 0] = [1 / s2[0], l01 ** 2 / s2[1]]
        w[1, 0] = [0.0, -l01 / s2[1]]
        w[1, 1] = [0.0, 1 / s2[1]]
    elif nv == 3:
        l01, l02, l12 = lam
        w[0, 0] = [1 / s2[0], l01 ** 2 / s2[1], l12 ** 2 / s2[2]]
        w[1, 0] = [0.0, -l01 / s2[1], l02 * l12 / s2[2]]
        w[2, 0] = [0.0, 0.0, -l12 / s2[2]]
        w[1, 1] = [0.0, 1 / s2[1], l02 ** 2 / s2[2]]
        w[2, 1] = [0.0, 0.0, -l02 / s2[2]]
        w[2, 2] = [0.0, 0.0, 1 / s2[2]]
    else:
        raise ValueError("nv must be 2 or 3")
    for i in range(nv):
        for j in range(i + 1, nv):
            w[i, j] = w[j, i]
    return w


def coreg_param_row(nv: int, sigma_st: float, theta) -> np.ndarray:
    """Device parameters: C_ij[3][4] for every (i, j) (row-major), then exp(prec_o_i)."""
    theta = np.asarray(theta, dtype=np.float64)
    coef = [hyper.kron_coefficients(np.array([theta[3 * k], theta[3 * k + 1], sigma_st])) for k in range(nv)]
    w = mixing_weights(nv, theta)
    C = np.einsum("ijk,kab->ijab", w, np.stack(coef))
    taus = np.array([math.exp(theta[3 * i + 2]) for i in range(nv)])
    return np.concatenate([C.ravel(), taus])


class DeviceCoregModel:
    """θ-independent inputs of a coregional model, uploaded once."""

    def __init__(self, data, device="cuda"):
        self.data = data
        nv, ns, nt, nb = data.nv, data.ns, data.nt, data.nb
        self.nv, self.ns, self.nt, self.nb = nv, ns, nt, nb
        self.n_latent = data.n_latent
        self.n_obs_v = [int(y.size) for y in data.y]
        self.yty_v = [float(y @ y) for y in data.y]
        dev = torch.device(device)

        def up(x, dtype=torch.float64):
            return torch.as_tensor(np.ascontiguousarray(x)).to(dtype=dtype, device=dev)

        rp, col, val = fem_union(data.c0, data.g1, data.g2, data.g3)
        self.fem_rowptr, self.fem_col, self.fem_val = up(rp, torch.int32), up(col, torch.int32), up(val)
        md, mo = tridiag_parts(data.m0, data.m1, data.m2, nt)
        self.m_diag_h, self.m_off_h = md, mo
        self.m_diag, self.m_off = up(md), up(mo)
        # Kronecker-eigen closed form of each variate's prior log-determinant (the adjoint's prior term)
        self.logdet_c0 = float(np.sum(np.log(sps.csr_matrix(data.c0).diagonal())))
        self.closed_form = closed_form_ok(data.c0, data.g1, data.g2, data.g3)
        self.lam = None
        if self.closed_form:
            c = sps.csr_matrix(data.c0).diagonal()
            sc = 1.0 / np.sqrt(c)
            S = (sps.diags(sc) @ data.g1 @ sps.diags(sc)).toarray()
            self.lam = up(np.linalg.eigvalsh(0.5 * (S + S.T)))
        self.keep = []
        structs = []
        nst = ns * nt
        aty = np.zeros(self.n_latent)
        vidx = np.zeros(self.n_latent, dtype=np.int32)
        for i in range(nv):
            A = _csr(data.a_st[i])
            X = np.ascontiguousarray(data.x_cov[i], dtype=np.float64)
            AtA = _csr(A.T @ A)
            t = dict(rp=up(AtA.indptr.astype(np.int64), torch.int64), col=up(AtA.indices.astype(np.int32), torch.int32),
                     val=up(AtA.data), arrow=up(np.ascontiguousarray((A.T @ X).T)), tip=up(X.T @ X))
            self.keep.append(t)
            structs.append(_lib.DhModel(
                ns=ns, nt=nt, nb=nb, fem_nnz=int(col.size), fem_rowptr=self.fem_rowptr.data_ptr(),
                fem_col=self.fem_col.data_ptr(), fem_val=self.fem_val.data_ptr(), m_diag=self.m_diag.data_ptr(),
                m_off=self.m_off.data_ptr(), ata_nnz=int(AtA.nnz), ata_rowptr=t["rp"].data_ptr(),
                ata_col=t["col"].data_ptr(), ata_val=t["val"].data_ptr(), ata_arrow=t["arrow"].data_ptr(),
                ata_tip=t["tip"].data_ptr(), fe_prec=float(data.fe_prec)))
            # A_i^T y_i placed at the permuted latent positions (time-block t, variate i, node s)
            aty_st = (A.T @ data.y[i]).reshape(nt, ns)
            for tt in range(nt):
                aty[tt * nv * ns + i * ns: tt * nv * ns + (i + 1) * ns] = aty_st[tt]
                vidx[tt * nv * ns + i * ns: tt * nv * ns + (i + 1) * ns] = i
            aty[nv * nst + i * nb: nv * nst + (i + 1) * nb] = X.T @ data.y[i]
            vidx[nv * nst + i * nb: nv * nst + (i + 1) * nb] = i
        self.models = (_lib.DhModel * nv)(*structs)
        self.nst = nst
        self.aty = up(aty)
        self.vidx = up(vidx, torch.int32)


class CoregInla:
    """Plain f(theta) of the coregional model; the engine interface of `STInla`
    used by `BatchedF` (params, param_row, _enqueue_f, _f_from_scalars, set_stream)."""

    #: the autotuner chooses between the one- and two-ended factorization for this engine
    twisted_auto = True

    def __init__(self, data, device="cuda"):
        self.data = data
        self.dm = DeviceCoregModel(data, device=device)
        nv = data.nv
        self.solver = HipBTASolver(diagonal_blocksize=data.b, n_diag_blocks=data.nt, arrowhead_blocksize=data.a)
        self.lib = self.solver.lib
        self.h = self.solver.h
        N = data.n_latent
        self.b = torch.zeros(N, dtype=torch.float64, device="cuda")
        self.mu = torch.zeros(N, dtype=torch.float64, device="cuda")
        self.scal = torch.zeros(8, dtype=torch.float64, device="cuda")
        self.np = nv * nv * 12 + nv
        self.params = torch.zeros(self.np, dtype=torch.float64, device="cuda")
        self.params_h = torch.zeros(self.np, dtype=torch.float64).pin_memory()
        self.n_f_evals = 0

    def param_row(self, theta) -> np.ndarray:
        return coreg_param_row(self.data.nv, self.data.sigma_st, theta)

    def _set_params(self, theta):
        self.params_h.copy_(torch.from_numpy(self.param_row(theta)))
        self.params.copy_(self.params_h, non_blocking=True)

    def set_stream(self, stream) -> None:
        _lib.check(self.lib.dh_set_stream(self.h, ctypes.c_void_p(stream.cuda_stream)), "dh_set_stream")

    def _assemble(self, theta, sparsity: str, with_like: bool):
        """Write Qc ("bta", with_like) or the prior ("bt") into the blocks (the autotuner's hook)."""
        if theta is not None:
            self._set_params(theta)
        _lib.check(self.lib.dh_assemble_coreg(self.h, self.dm.models, self.data.nv, self.params.data_ptr(),
                                              _lib.SPARSITY[sparsity], int(with_like)), "dh_assemble_coreg")

    def _enqueue_f(self, theta, out=None):
        """Device work of one f evaluation; out[0] = logdet Qc, out[1] = mu^T b, out[4] = logdet Qp|BT."""
        if theta is not None:
            self._set_params(theta)
        out = self.scal if out is None else out
        self._fused_a(out)
        _lib.check(self.lib.dh_cholesky(self.h, 1), "dh_cholesky")
        self._fused_b(out)
        _lib.check(self.lib.dh_cholesky(self.h, 0), "dh_cholesky")
        self._fused_c(out)

    # the phases between the factorizations (see STInla._fused_a)
    def _fused_a(self, out):
        _lib.check(self.lib.dh_assemble_coreg(self.h, self.dm.models, self.data.nv, self.params.data_ptr(), 1, 1),
                   "dh_assemble_coreg")

    def _fused_b(self, out):
        L, h, nv = self.lib, self.h, self.data.nv
        N = self.data.n_latent
        taus = self.params.data_ptr() + 8 * nv * nv * 12
        _lib.check(L.dh_scale_indexed(h, self.dm.aty.data_ptr(), self.dm.vidx.data_ptr(), ctypes.c_void_p(taus),
                                      self.b.data_ptr(), N), "dh_scale_indexed")
        _lib.check(L.dh_scale(h, self.b.data_ptr(), 1.0, self.mu.data_ptr(), N), "dh_scale")
        _lib.check(L.dh_solve(h, self.mu.data_ptr(), 1), "dh_solve")
        _lib.check(L.dh_logdet(h, 1, out[0:1].data_ptr()), "dh_logdet")
        _lib.check(L.dh_dot(h, self.mu.data_ptr(), self.b.data_ptr(), N, out[1:2].data_ptr()), "dh_dot")
        _lib.check(L.dh_assemble_coreg(h, self.dm.models, nv, self.params.data_ptr(), 0, 0), "dh_assemble_coreg")

    def _fused_c(self, out):
        _lib.check(self.lib.dh_logdet(self.h, 0, out[4:5].data_ptr()), "dh_logdet")

    def _scalars(self, theta):
        theta = np.asarray(theta, dtype=np.float64)
        lp, _ = hyper.log_prior_and_grad(theta, self.data.priors)
        lik = sum(hyper.likelihood_eta0(float(theta[3 * i + 2]), self.dm.n_obs_v[i], self.dm.yty_v[i])
                  for i in range(self.data.nv))
        return lp, lik

    def _f_from_scalars(self, theta, s) -> float:
        ld_c, quad, ld_p = float(s[0]), float(s[1]), float(s[4])
        if np.isnan(ld_c) or np.isnan(ld_p):
            _lib.raise_nan_logdet(self.lib, self.h, ld_c, ld_p)
        lp, lik = self._scalars(theta)
        f = 0.5 * ld_c - 0.5 * quad - (lp + lik + 0.5 * ld_p)
        if np.isnan(f):
            raise ValueError("f(theta) is NaN")
        return float(f)

    def f(self, theta) -> float:
        theta = np.asarray(theta, dtype=np.float64)
        if any(p.kind != "gaussian" for p in self.data.priors):
            raise ValueError("coregional hyper-priors must be Gaussian")
        self._enqueue_f(theta)
        s = self.scal[:5].cpu().numpy()
        self.n_f_evals += 1
        return self._f_from_scalars(theta, s)

    def fd_gradient(self, theta, h: float = 1e-4):
        """Central differences, 2d+1 evaluations (`core/dalia.py:513-601`)."""
        theta = np.asarray(theta, dtype=np.float64)
        d = theta.size
        f0 = self.f(theta)
        g = np.zeros(d)
        for j in reversed(range(d)):
            e = np.zeros(d)
            e[j] = h
            g[j] = (self.f(theta + e) - self.f(theta - e)) / (2 * h)
        return f0, g

    def hessian(self, theta, h: float = 5e-3, offdiag_points: int = 4):
        pts = hessian_stencil(theta, h, offdiag_points)
        fv = np.array([self.f(p) for p in pts])
        H = hessian_from_values(fv, len(theta), h, offdiag_points)
        check_hessian(H)
        return H

    # ------------------------------------------------------------------
    # adjoint gradient
    def _coef_grads(self, theta):
        """dC[p][i, j] (3 x 4) = d C_ij / d theta_p for every theta entry, C_ij = sum_k w_ijk c_k."""
        nv, sst = self.data.nv, self.data.sigma_st
        theta = np.asarray(theta, dtype=np.float64)
        d = theta.size
        coef = np.stack([hyper.kron_coefficients(np.array([theta[3 * k], theta[3 * k + 1], sst])) for k in range(nv)])
        w = mixing_weights(nv, theta)
        dC = np.zeros((d, nv, nv, 3, 4))
        for k in range(nv):
            dc = hyper.kron_coefficients_grad(np.array([theta[3 * k], theta[3 * k + 1], sst]))
            for a in range(2):  # r_s_k, r_t_k
                dC[3 * k + a] = w[:, :, k, None, None] * dc[a][None, None]
        hcs = 1e-30
        for p in range(3 * nv, d):  # sigmas and lambdas: complex-step derivative of the weights
            tc = theta.astype(np.complex128)
            tc[p] += 1j * hcs
            dw = mixing_weights(nv, tc).imag / hcs
            dC[p] = np.einsum("ijk,kab->ijab", dw, coef)
        return dC

    def forward(self, theta):
        """Forward half of the adjoint: f(theta) with each variate's closed-form prior term."""
        theta = np.asarray(theta, dtype=np.float64)
        if any(p.kind != "gaussian" for p in self.data.priors):
            raise ValueError("coregional hyper-priors must be Gaussian")
        if not self.dm.closed_form:
            raise ValueError("the coregional adjoint needs g2 = g1 c0^-1 g1 and g3 = g1 c0^-1 g2 (lumped c0)")
        self._set_params(theta)
        L, h, nv, N = self.lib, self.h, self.data.nv, self.data.n_latent
        _lib.check(L.dh_assemble_coreg(h, self.dm.models, nv, self.params.data_ptr(), 1, 1), "dh_assemble_coreg")
        _lib.check(L.dh_cholesky(h, 1), "dh_cholesky")
        taus = self.params.data_ptr() + 8 * nv * nv * 12
        _lib.check(L.dh_scale_indexed(h, self.dm.aty.data_ptr(), self.dm.vidx.data_ptr(), ctypes.c_void_p(taus),
                                      self.b.data_ptr(), N), "dh_scale_indexed")
        _lib.check(L.dh_scale(h, self.b.data_ptr(), 1.0, self.mu.data_ptr(), N), "dh_scale")
        _lib.check(L.dh_solve(h, self.mu.data_ptr(), 1), "dh_solve")
        _lib.check(L.dh_logdet(h, 1, self.scal[0:1].data_ptr()), "dh_logdet")
        _lib.check(L.dh_dot(h, self.mu.data_ptr(), self.b.data_ptr(), N, self.scal[1:2].data_ptr()), "dh_dot")
        eig = torch.zeros((nv, 4), dtype=torch.float64, device="cuda")
        for k in range(nv):
            gam = np.array(hyper.interpretable2compute(float(theta[3 * k]), float(theta[3 * k + 1]),
                                                       float(self.data.sigma_st)), dtype=np.float64)
            _lib.check(L.dh_prior_eig(h, ctypes.byref(self.dm.models[k]), self.dm.lam.data_ptr(),
                                      gam.ctypes.data_as(_lib.c_double_p), eig[k].data_ptr()), "dh_prior_eig")
        s = self.scal[:2].cpu().numpy()
        e = eig.cpu().numpy()
        ld_c, quad = float(s[0]), float(s[1])
        if np.isnan(ld_c):
            _lib.raise_nan_logdet(self.lib, self.h, ld_c)
        nst = self.dm.nst
        ld_p = sum(self.dm.nt * self.dm.logdet_c0 + e[k, 0] for k in range(nv)) \
            - 2.0 * nst * float(np.sum(theta[3 * nv:4 * nv]))
        dld_p = np.zeros(theta.size)
        for k in range(nv):
            dg = hyper.DGAMMA.T @ e[k, 1:4]
            dld_p[3 * k], dld_p[3 * k + 1] = dg[0], dg[1]
            dld_p[3 * nv + k] = -2.0 * nst
        lp, glp = hyper.log_prior_and_grad(theta, self.data.priors)
        lik = sum(hyper.likelihood_eta0(float(theta[3 * i + 2]), self.dm.n_obs_v[i], self.dm.yty_v[i])
                  for i in range(nv))
        f = 0.5 * ld_c - 0.5 * quad - (lp + lik + 0.5 * ld_p)
        self._state = dict(theta=theta.copy(), f=float(f), dld_p=dld_p, glp=glp)
        return float(f)

    def backward(self):
        """Gradient of the last `forward`: selected inversion of Qc and its contractions."""
        st = getattr(self, "_state", None)
        if st is None:
            raise RuntimeError("backward() needs a preceding forward() on this engine")
        theta = st["theta"]
        L, h = self.lib, self.h
        nv, ns, nb, nt, b = self.data.nv, self.dm.ns, self.dm.nb, self.dm.nt, self.data.b
        nq = 2 * nt - 1
        m0 = ctypes.byref(self.dm.models[0])
        cmu = torch.zeros((nv, nv, nq * 4), dtype=torch.float64, device="cuda")
        cx = torch.zeros_like(cmu)
        sc = torch.zeros((2, nv), dtype=torch.float64, device="cuda")
        for i in range(nv):
            for j in range(nv):
                _lib.check(L.dh_contract_fem_vec_v(h, m0, self.mu.data_ptr(), b, i * ns, j * ns, cmu[i, j].data_ptr()),
                           "dh_contract_fem_vec_v")
        for k in range(nv):
            _lib.check(L.dh_ata_quad_v(h, ctypes.byref(self.dm.models[k]), self.mu.data_ptr(), b, k * ns, k * nb,
                                       sc[1, k:k + 1].data_ptr()), "dh_ata_quad_v")
        _lib.check(L.dh_selected_inversion(h, 1), "dh_selected_inversion")
        for i in range(nv):
            for j in range(nv):
                _lib.check(L.dh_contract_fem_v(h, m0, i * ns, j * ns, cx[i, j].data_ptr()), "dh_contract_fem_v")
        for k in range(nv):
            _lib.check(L.dh_contract_ata_v(h, ctypes.byref(self.dm.models[k]), k * ns, k * nb,
                                           sc[0, k:k + 1].data_ptr()), "dh_contract_ata_v")
        prod = self.mu * self.dm.aty
        bmu = torch.stack([prod[self.dm.vidx == k].sum() for k in range(nv)]).cpu().numpy()
        cxmu = (cx + cmu).cpu().numpy().reshape(nv, nv, nq, 4)
        tr_ata, quad = sc.cpu().numpy()
        mult = np.concatenate([np.ones(nt), 2.0 * np.ones(nt - 1)])[:, None]
        dC = self._coef_grads(theta)
        g = np.zeros(theta.size)
        for p in range(theta.size):
            if p < 3 * nv and p % 3 == 2:  # likelihood precision of variate k
                k = p // 3
                et = math.exp(float(theta[p]))
                g[p] = 0.5 * et * tr_ata[k] + 0.5 * et * quad[k] - et * bmu[k] \
                    - 0.5 * self.dm.n_obs_v[k] + 0.5 * et * self.dm.yty_v[k]
                continue
            acc = 0.0
            for i in range(nv):
                for j in range(nv):
                    if not np.any(dC[p, i, j]):
                        continue
                    w = np.concatenate([self.dm.m_diag_h.T @ dC[p, i, j], self.dm.m_off_h.T @ dC[p, i, j]], 0) * mult
                    acc += float(np.sum(w * cxmu[i, j]))
            g[p] = 0.5 * acc - 0.5 * st["dld_p"][p]
        g -= st["glp"]
        self._state = None
        return st["f"], g

    def value_and_grad(self, theta):
        self.forward(theta)
        return self.backward()

