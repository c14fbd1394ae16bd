"""One f(theta) evaluation of DALIA's Gaussian spatio-temporal model on the GPU,
its finite-difference gradient and its adjoint gradient.

`STInla.f` follows `DALIA._evaluate_f` (`src/dalia/core/dalia.py:603-745`,
Gaussian branch) term by term:

    f = [1/2 logdet Qc - 1/2 mu^T Qc mu]                      (:1110-1168)
        - [log pi(theta) + lik(eta=0) + 1/2 logdet Qp|BT]     (:679-683, :1074-1108)

with Qc = Qp + e^tau A^T A assembled on the device (`dh_assemble`),
mu = Qc^-1 (e^tau A^T y) by pobtaf + pobtas, the quadratic form by an SpMV
with Qc (`dh_qc_matvec`, as `:1161`), and the prior log-determinant by a BT
factorisation of Qp that drops the arrowhead (`:1100-1103`), reusing the same
buffers in the same order (Qc factor first, then Qp).

`STInla.fd_gradient` is the reference's 2d+1 central-difference stencil
(`core/dalia.py:545-581`).  `STInla.value_and_grad` replaces it by the adjoint
formula

    df/dtheta_j = 1/2 tr(Qc^-1 dQc) - db^T mu + 1/2 mu^T dQc mu
                  - 1/2 d logdet Qp|BT - d log pi - d lik,

where the trace is a contraction of the selected-inverse blocks (pobtasi)
against the sparse Kronecker blocks of dQ (`dh_contract_fem`) and against
A^T A (`dh_contract_ata`), and the prior log-determinant and its derivatives
come from the Kronecker-eigen closed form (`dh_prior_eig`) when the FEM
matrices satisfy it, or else from a BT factorisation plus selected inversion
of Qp.  `FTheta` wraps it as a `torch.autograd.Function`: forward = factor,
solve, log-determinants; backward = selected inversion + contractions.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from dalia_amd import _lib, hyper
from dalia_amd.model import DeviceSTModel
from dalia_amd.solver import HipBTASolver


class STInla:
    def __init__(self, data, device="cuda", closed_form_prior: bool = True, overlap: int = 0,
                 twisted: bool = True, pattern_selinv: bool = True, profile_from_assembly: bool = True,
                 inverse_beside_solve: bool = True, selinv_beside_solve: bool = True):
        """overlap: work run beside each factorization on a second stream (`HipBTASolver.set_overlap`),
        at most this mode: 0 none (default), 1 the inverse of the diagonal factors, 2 also the
        selected inversion's products of the factor.  Measured at ns = 2025, nt = 100: the adjoint
        gradient gains about 2 % with 2, as the factorization slows down beside that work.
        twisted: two-ended factorizations (`HipBTASolver.set_twisted`, default, for nt >= 3;
        evaluations with overlap > 0 stay one-ended).  Measured at ns = 2025, nt = 100: pobtaf
        77.9 -> 70.7 ms, the adjoint gradient 169.1 -> 162.3 ms.
        pattern_selinv: the selected inversion in congruence form with the couplings' blocks formed
        only on the matrix's tiles (`HipBTASolver.set_selected_inversion_pattern`, default): every
        consumer here reads the selected inverse on Qc's pattern (the adjoint contractions, the
        marginal variances).
        profile_from_assembly: the zero-tile profiles of Qc's couplings recorded by the assembly
        (`HipBTASolver.set_profile_from_assembly`, default) instead of scanned before each
        factorization (0.66 ms at ns = 2025, nt = 100); the same profiles.
        inverse_beside_solve: the inverse of the diagonal factors formed on a second stream beside
        the solve's forward sweep, which waits for each chunk of it before reading it
        (`HipBTASolver.set_inverse_beside_solve`, default); bitwise the same results.
        selinv_beside_solve: the selected inversion of the gradient on a second stream beside the
        solve's backward sweep, each recursion step after the sweep's step that reads its blocks
        (`HipBTASolver.set_selinv_beside_solve`, default); bitwise the same results."""
        self.data = data
        self.manifold = getattr(data, "manifold", "plane")
        if self.manifold not in hyper.MANIFOLDS:
            raise ValueError(f"Manifold not supported: {self.manifold}")
        self.dm = DeviceSTModel(data, device=device, eigen=closed_form_prior)
        self.solver = HipBTASolver(diagonal_blocksize=data.ns, n_diag_blocks=data.nt,
                                   arrowhead_blocksize=data.nb)
        self.lib = self.solver.lib
        self.h = self.solver.h
        self.twisted = bool(twisted) and data.nt >= 3
        if self.twisted:
            self.solver.set_twisted(True)
        self.pattern_selinv = bool(pattern_selinv)
        self.solver.set_selected_inversion_pattern(self.pattern_selinv)
        self.solver.set_profile_from_assembly(bool(profile_from_assembly))
        self.solver.set_inverse_beside_solve(bool(inverse_beside_solve))
        self.solver.set_selinv_beside_solve(bool(selinv_beside_solve))
        N = self.dm.n_latent
        self.mu = torch.zeros(N, dtype=torch.float64, device="cuda")
        self.yp = torch.zeros(N, dtype=torch.float64, device="cuda")
        self.yl = torch.zeros(N, dtype=torch.float64, device="cuda")
        self.scal = torch.zeros(16, dtype=torch.float64, device="cuda")
        nq = 2 * data.nt - 1
        self.cx = torch.zeros(nq * 4, dtype=torch.float64, device="cuda")
        self.cmu = torch.zeros(nq * 4, dtype=torch.float64, device="cuda")
        self.eig = torch.zeros(4, dtype=torch.float64, device="cuda")
        # theta-dependent kernel parameters: 12 Kronecker coefficients + exp(theta_obs)
        self.params = torch.zeros(13, dtype=torch.float64, device="cuda")
        self.params_h = torch.zeros(13, dtype=torch.float64).pin_memory()
        # (gamma_s, gamma_t, gamma_st) for the closed-form prior term of the captured gradient
        self.gam = torch.zeros(3, dtype=torch.float64, device="cuda")
        self.gam_h = torch.zeros(3, dtype=torch.float64).pin_memory()
        self.ones = torch.ones(1, dtype=torch.float64, device="cuda")
        self._state = None
        self._graph = None
        self._ggraph = None
        self._aux = None  # second stream of the adjoint evaluation (default priority)
        self.n_f_evals = 0
        self._theta_last = None  # last theta evaluated (the reference keeps it in model.theta)
        self._ov_max = int(overlap)
        self._ov = 0          # current overlap mode of the handle
        self._ov2_ok = None   # whether mode 2's per-block slots were reserved

    # ------------------------------------------------------------------
    def _set_params(self, theta):
        """Stage the theta-dependent scalars in pinned memory and copy them to the device."""
        self._theta_last = np.array(theta, dtype=np.float64)
        p = np.concatenate([DeviceSTModel.coef_array(theta, self.manifold), [self.dm.tau_scale(theta)]])
        self.params_h.copy_(torch.from_numpy(p))
        self.params.copy_(self.params_h, non_blocking=True)

    def _assemble(self, theta, sparsity: str, with_like: bool):
        if theta is not None:
            self._set_params(theta)
        _lib.check(self.lib.dh_assemble(self.h, self.dm.ref(), self.params.data_ptr(),
                                        _lib.SPARSITY[sparsity], int(with_like)), "dh_assemble")

    def _overlap(self, mode: int) -> None:
        """Work the next factorizations run beside themselves: 0 none, 1 the inverse of the
        diagonal factors (the solve), 2 also the selected inversion's products of the factor
        (falls back to 1 when their per-block slots do not fit); capped by the engine's option."""
        mode = min(mode, self._ov_max)
        if mode == 2:
            if self._ov2_ok is None:
                self._ov2_ok = self.solver.set_overlap(2)
                self._ov = 2 if self._ov2_ok else self._ov
            if not self._ov2_ok:
                mode = 1
        if mode != self._ov:
            self.solver.set_overlap(mode)
            self._ov = mode

    def _conditional(self, theta, out=None, grad: bool = False, terms: bool = True):
        """Assemble Qc, factor, solve for mu; queue logdet and the quadratic-form pieces.

        theta None: the parameters already on the device are used (graph replay).
        out: 5-element device tensor for (logdet Qc, mu^T Qp mu, mu^T tau A^T A mu,
        mu^T A^T y, logdet Qp) (default: the engine's own slots).
        grad: a selected inversion of Qc follows (its products of the factor run beside the
        factorization).
        terms: False leaves the quadratic-form pieces to the caller (`_mu_terms`).
        """
        out = self.scal if out is None else out
        self._assemble(theta, "bta", True)
        self._overlap(2 if grad else 1)
        _lib.check(self.lib.dh_cholesky(self.h, 1), "dh_cholesky")
        self._after_qc(out, terms)

    def _after_qc(self, out, terms: bool = True):
        """After the factorization of Qc: mu, logdet Qc and (terms) the quadratic-form pieces into out."""
        L = self.lib
        N = self.dm.n_latent
        # mu = e^tau A^T y: scale by params[12] on the device (a 1x1 "scale" kernel via dot-free copy)
        _lib.check(L.dh_scale_dev(self.h, self.dm.aty.data_ptr(), self.params[12:13].data_ptr(),
                                  self.mu.data_ptr(), N), "dh_scale_dev")
        # logdet first: it reads the factor, so after the solve it would keep a selected inversion
        # from running beside the solve's backward sweep
        _lib.check(L.dh_logdet(self.h, 1, out[0:1].data_ptr()), "dh_logdet")
        _lib.check(L.dh_solve(self.h, self.mu.data_ptr(), 1), "dh_solve")
        if terms:
            self._mu_terms(out)

    def _mu_terms(self, out):
        """mu^T Qp mu, mu^T tau A^T A mu and mu^T A^T y into out[1:4] (they read only mu and the
        model's data, not the factor)."""
        L = self.lib
        N = self.dm.n_latent
        _lib.check(L.dh_qc_matvec(self.h, self.dm.ref(), self.params.data_ptr(),
                                  self.mu.data_ptr(), self.yp.data_ptr(), self.yl.data_ptr()), "dh_qc_matvec")
        _lib.check(L.dh_dot(self.h, self.mu.data_ptr(), self.yp.data_ptr(), N, out[1:2].data_ptr()), "dh_dot")
        _lib.check(L.dh_dot(self.h, self.mu.data_ptr(), self.yl.data_ptr(), N, out[2:3].data_ptr()), "dh_dot")
        _lib.check(L.dh_dot(self.h, self.mu.data_ptr(), self.dm.aty.data_ptr(), N, out[3:4].data_ptr()), "dh_dot")

    def _scalars(self, theta):
        lp, glp = hyper.log_prior_and_grad(theta, self.data.priors)
        lik = hyper.likelihood_eta0(float(theta[-1]), self.dm.n_obs, self.dm.yty)
        return lp, glp, lik

    # ------------------------------------------------------------------
    def _enqueue_f(self, theta, out=None):
        """All device work of one plain f evaluation (no host synchronisation)."""
        out = self.scal if out is None else out
        self._conditional(theta, out)
        self._assemble(None, "bt", False)
        self._overlap(0)  # logdet only
        _lib.check(self.lib.dh_cholesky(self.h, 0), "dh_cholesky")
        _lib.check(self.lib.dh_logdet(self.h, 0, out[4:5].data_ptr()), "dh_logdet")

    def set_stream(self, stream) -> None:
        _lib.check(self.lib.dh_set_stream(self.h, ctypes.c_void_p(stream.cuda_stream)), "dh_set_stream")

    # the three phases of one plain f evaluation between the factorizations, for a batch whose
    # factorizations run as one launch over all instances (`dalia_amd.batch.BatchedF`)
    def _fused_a(self, out):
        self._assemble(None, "bta", True)

    def _fused_b(self, out):
        self._after_qc(out)
        self._assemble(None, "bt", False)

    def _fused_c(self, out):
        _lib.check(self.lib.dh_logdet(self.h, 0, out[4:5].data_ptr()), "dh_logdet")

    def param_row(self, theta) -> np.ndarray:
        """The 13 device parameters of theta (Kronecker coefficients, exp(theta_obs))."""
        return np.concatenate([DeviceSTModel.coef_array(theta, self.manifold), [self.dm.tau_scale(theta)]])

    def hessian(self, theta, h: float = 5e-3, offdiag_points: int = 4):
        """FD Hessian of f at theta (`core/dalia.py:767-910`), plain f per stencil point.

        offdiag_points=4: the reference's stencil, 1 + 2d + 4 d(d-1)/2 points;
        offdiag_points=2: 1 + 2d + 2 d(d-1)/2 points (+,+ and -,- corners).
        """
        pts = hessian_stencil(theta, h, offdiag_points)
        fv = np.array([self.f(p) for p in pts])
        H = hessian_from_values(fv, len(theta), h, offdiag_points)
        check_hessian(H)
        return H

    def _f_from_scalars(self, theta, s) -> float:
        ld_c, quad_p, quad_l, _, ld_p = s
        if np.isnan(ld_c) or np.isnan(ld_p):
            _lib.raise_nan_logdet(self.lib, self.h, ld_c, ld_p)
        lp, _, lik = self._scalars(theta)
        f = 0.5 * ld_c - 0.5 * (quad_p + quad_l) - (lp + lik + 0.5 * ld_p)
        if np.isnan(f):
            raise ValueError("f(theta) is NaN")
        return float(f)

    def capture_f(self):
        """Capture one plain f evaluation (parameter upload included) in a HIP graph."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):  # warm-up on the side stream before capture
            _lib.check(self.lib.dh_set_stream(self.h, ctypes.c_void_p(s.cuda_stream)), "dh_set_stream")
            self._set_params(self.data.theta)
            self._enqueue_f(None)
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            cs = torch.cuda.current_stream()
            _lib.check(self.lib.dh_set_stream(self.h, ctypes.c_void_p(cs.cuda_stream)), "dh_set_stream")
            self.params.copy_(self.params_h, non_blocking=True)
            self._enqueue_f(None)
        _lib.check(self.lib.dh_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                   "dh_set_stream")
        self._graph = g
        return g

    def f_graph(self, theta) -> float:
        """Plain f(theta) by replaying the captured graph."""
        if self._graph is None:
            self.capture_f()
        theta = np.asarray(theta, dtype=np.float64)
        p = np.concatenate([DeviceSTModel.coef_array(theta, self.manifold), [self.dm.tau_scale(theta)]])
        self.params_h.copy_(torch.from_numpy(p))
        self._graph.replay()
        s = self.scal[:5].cpu().numpy()
        self.n_f_evals += 1
        self._state = None
        return self._f_from_scalars(theta, s)

    def f(self, theta) -> float:
        """Plain drop-in f(theta): two block factorisations as in the reference."""
        theta = np.asarray(theta, dtype=np.float64)
        self._set_params(theta)
        self._enqueue_f(None)
        s = self.scal[:5].cpu().numpy()
        ld_c, quad_p, quad_l, _, ld_p = s
        if np.isnan(ld_c) or np.isnan(ld_p):
            _lib.raise_nan_logdet(self.lib, self.h, ld_c, ld_p)
        lp, _, lik = self._scalars(theta)
        f = 0.5 * ld_c - 0.5 * (quad_p + quad_l) - (lp + lik + 0.5 * ld_p)
        self.n_f_evals += 1
        self._state = None
        if np.isnan(f):
            raise ValueError("f(theta) is NaN")
        return float(f)

    def fd_gradient(self, theta, h: float = 1e-4):
        """Reference central differences: 2d+1 plain evaluations, reverse order."""
        theta = np.asarray(theta, dtype=np.float64)
        d = theta.size
        n_eval = 2 * d + 1
        tm = np.repeat(theta.reshape(-1, 1), n_eval, axis=1)
        tm[:, 1:1 + d] += h * np.eye(d)
        tm[:, d + 1:] -= h * np.eye(d)
        fv = np.zeros(n_eval)
        for k in range(n_eval - 1, -1, -1):
            fv[k] = self.f(tm[:, k])
        return fv[0], (fv[1:1 + d] - fv[d + 1:]) / (2 * h)

    # ------------------------------------------------------------------
    def _prior_logdet_terms(self, theta):
        """(logdet Qp|BT, d/dtheta_{0..2}) via the Kronecker-eigen closed form."""
        gs, gt, gst = hyper.interpretable2compute(float(theta[0]), float(theta[1]), float(theta[2]), self.manifold)
        gam = np.array([gs, gt, gst], dtype=np.float64)
        _lib.check(self.lib.dh_prior_eig(self.h, self.dm.ref(), self.dm.lam.data_ptr(),
                                         gam.ctypes.data_as(_lib.c_double_p), self.eig.data_ptr()), "dh_prior_eig")

    def forward(self, theta):
        """Forward half of the adjoint: f(theta) with the closed-form prior term."""
        theta = np.asarray(theta, dtype=np.float64)
        self._conditional(theta, grad=True)
        if self.dm.closed_form:
            self._prior_logdet_terms(theta)
            s = self.scal[:4].cpu().numpy()
            e = self.eig.cpu().numpy()
            ld_p = self.dm.nt * self.dm.logdet_c0 + e[0]
            dld_p = hyper.dgamma(theta, self.manifold).T @ e[1:4]
        else:
            s = self.scal[:4].cpu().numpy()
            ld_p, dld_p = None, None
        ld_c, quad_p, quad_l, bmu = s
        if np.isnan(ld_c):
            _lib.raise_nan_logdet(self.lib, self.h, ld_c)
        lp, glp, lik = self._scalars(theta)
        self._state = dict(theta=theta.copy(), ld_c=ld_c, quad_p=quad_p, quad_l=quad_l, bmu=bmu,
                           ld_p=ld_p, dld_p=dld_p, lp=lp, glp=glp, lik=lik)
        if ld_p is None:
            return None
        return float(0.5 * ld_c - 0.5 * (quad_p + quad_l) - (lp + lik + 0.5 * ld_p))

    def backward(self):
        """Gradient of the last `forward` (selected inversion + contractions)."""
        st = self._state
        if st is None:
            raise RuntimeError("backward() needs a preceding forward() on this engine")
        theta = st["theta"]
        L = self.lib
        et = self.dm.tau_scale(theta)
        # mu^T G_m mu blocks (before the buffers are overwritten)
        _lib.check(L.dh_contract_fem_vec(self.h, self.dm.ref(), self.mu.data_ptr(), self.cmu.data_ptr()),
                   "dh_contract_fem_vec")
        _lib.check(L.dh_selected_inversion(self.h, 1), "dh_selected_inversion")
        _lib.check(L.dh_contract_fem(self.h, self.dm.ref(), self.cx.data_ptr()), "dh_contract_fem")
        _lib.check(L.dh_contract_ata(self.h, self.dm.ref(), self.scal[5:6].data_ptr()), "dh_contract_ata")
        ld_p, dld_p = st["ld_p"], st["dld_p"]
        if ld_p is None:
            ld_p, dld_p = self._prior_by_selected_inversion(theta)
        return self._gradient(st, ld_p, dld_p)

    def _gradient(self, st, ld_p, dld_p):
        """f and the adjoint gradient from the state of `forward` and the contractions in
        self.cx (selected inverse against the FEM blocks), self.cmu (mu against them) and
        self.scal[5] (selected inverse against A^T A)."""
        theta = st["theta"]
        et = self.dm.tau_scale(theta)
        nq = 2 * self.dm.nt - 1
        cx = self.cx.cpu().numpy().reshape(nq, 4)
        cmu = self.cmu.cpu().numpy().reshape(nq, 4)
        tr_ata = float(self.scal[5].item())
        dc = hyper.kron_coefficients_grad(theta, self.manifold)
        nt = self.dm.nt
        mult = np.concatenate([np.ones(nt), 2.0 * np.ones(nt - 1)])[:, None]
        g = np.zeros(theta.size)
        for j in range(3):
            w = self.dm.block_weights(dc[j]) * mult
            g[j] = 0.5 * float(np.sum(w * cx)) + 0.5 * float(np.sum(w * cmu)) - 0.5 * dld_p[j]
        g[3] = 0.5 * et * tr_ata - et * st["bmu"] + 0.5 * st["quad_l"]
        g[3] -= 0.5 * self.dm.n_obs - 0.5 * et * self.dm.yty
        g -= st["glp"]
        f = 0.5 * st["ld_c"] - 0.5 * (st["quad_p"] + st["quad_l"]) - (st["lp"] + st["lik"] + 0.5 * ld_p)
        self._state = None
        return float(f), g

    def _prior_by_selected_inversion(self, theta):
        """logdet Qp|BT and its theta-gradient by pobtf + pobtsi (general FEM inputs)."""
        L = self.lib
        self._assemble(None, "bt", False)
        self._overlap(2)
        _lib.check(L.dh_cholesky(self.h, 0), "dh_cholesky")
        _lib.check(L.dh_logdet(self.h, 0, self.scal[6:7].data_ptr()), "dh_logdet")
        _lib.check(L.dh_selected_inversion(self.h, 0), "dh_selected_inversion")
        cp = torch.zeros_like(self.cx)
        _lib.check(L.dh_contract_fem(self.h, self.dm.ref(), cp.data_ptr()), "dh_contract_fem")
        nq = 2 * self.dm.nt - 1
        cp = cp.cpu().numpy().reshape(nq, 4)
        dc = hyper.kron_coefficients_grad(theta, self.manifold)
        mult = np.concatenate([np.ones(self.dm.nt), 2.0 * np.ones(self.dm.nt - 1)])[:, None]
        d = np.array([float(np.sum(self.dm.block_weights(dc[j]) * mult * cp)) for j in range(3)])
        ld_p = float(self.scal[6].item())
        if np.isnan(ld_p):
            _lib.raise_nan_logdet(self.lib, self.h, ld_p)
        return ld_p, d

    def value_and_grad(self, theta):
        """f(theta) and its adjoint gradient: the whole evaluation enqueued at once (as the captured
        graph runs it), then one read-back; `forward` + `backward` without the closed-form prior."""
        if not self.dm.closed_form:
            self.forward(theta)
            return self.backward()
        theta = np.asarray(theta, dtype=np.float64)
        self._stage_grad_params(theta)
        self.params.copy_(self.params_h, non_blocking=True)
        self.gam.copy_(self.gam_h, non_blocking=True)
        self._enqueue_grad()
        return self._grad_readback(theta)

    # ------------------------------------------------------------------
    # the whole adjoint evaluation as one HIP graph
    def _enqueue_grad(self):
        """Device work of one adjoint evaluation with the parameters already on the device: the
        forward half, the closed-form prior term, mu^T G_m mu, the selected inversion and its
        contractions (no host synchronisation in between)."""
        L, m = self.lib, self.dm.ref()
        main = torch.cuda.current_stream()
        self.set_stream(main)
        self._conditional(None, grad=True, terms=False)
        # The work that reads only mu and the model's data runs on a second stream beside the
        # selected inversion, which overwrites the factor (and so follows the solve and logdet on
        # the main stream).  Both use the reduction scratch of the handle: the contractions of the
        # selected inverse wait for the second stream.
        if self._aux is None:
            self._aux = torch.cuda.Stream()
        aux = self._aux
        aux.wait_stream(main)
        self.set_stream(aux)
        self._mu_terms(self.scal)
        _lib.check(L.dh_prior_eig_dev(self.h, m, self.dm.lam.data_ptr(), self.gam.data_ptr(), self.eig.data_ptr()),
                   "dh_prior_eig_dev")
        _lib.check(L.dh_contract_fem_vec(self.h, m, self.mu.data_ptr(), self.cmu.data_ptr()), "dh_contract_fem_vec")
        self.set_stream(main)
        _lib.check(L.dh_selected_inversion(self.h, 1), "dh_selected_inversion")
        main.wait_stream(aux)
        _lib.check(L.dh_contract_fem(self.h, m, self.cx.data_ptr()), "dh_contract_fem")
        _lib.check(L.dh_contract_ata(self.h, m, self.scal[5:6].data_ptr()), "dh_contract_ata")

    def _stage_grad_params(self, theta):
        theta = np.asarray(theta, dtype=np.float64)
        self._theta_last = theta.copy()
        self.params_h.copy_(torch.from_numpy(self.param_row(theta)))
        gam = hyper.interpretable2compute(float(theta[0]), float(theta[1]), float(theta[2]), self.manifold)
        self.gam_h.copy_(torch.tensor(gam, dtype=torch.float64))

    def capture_grad(self):
        """Capture one adjoint evaluation (parameter upload included) in a HIP graph.  Needs the
        closed-form prior term; the overlap option does not apply under capture."""
        if not self.dm.closed_form:
            raise ValueError("the captured gradient needs the closed-form prior term")
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):  # warm-up on the side stream before capture
            _lib.check(self.lib.dh_set_stream(self.h, ctypes.c_void_p(s.cuda_stream)), "dh_set_stream")
            self._stage_grad_params(self.data.theta)
            self.params.copy_(self.params_h, non_blocking=True)
            self.gam.copy_(self.gam_h, non_blocking=True)
            self._enqueue_grad()
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            cs = torch.cuda.current_stream()
            _lib.check(self.lib.dh_set_stream(self.h, ctypes.c_void_p(cs.cuda_stream)), "dh_set_stream")
            self.params.copy_(self.params_h, non_blocking=True)
            self.gam.copy_(self.gam_h, non_blocking=True)
            self._enqueue_grad()
        _lib.check(self.lib.dh_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                   "dh_set_stream")
        self._ggraph = g
        return g

    def value_and_grad_graph(self, theta):
        """`value_and_grad` by replaying the captured adjoint evaluation (captured on first use)."""
        if self._ggraph is None:
            self.capture_grad()
        theta = np.asarray(theta, dtype=np.float64)
        self._stage_grad_params(theta)
        self._ggraph.replay()
        return self._grad_readback(theta)

    def _grad_readback(self, theta):
        """f and the gradient from the device results of one enqueued adjoint evaluation."""
        s = self.scal[:4].cpu().numpy()
        e = self.eig.cpu().numpy()
        ld_c, quad_p, quad_l, bmu = s
        if np.isnan(ld_c):
            _lib.raise_nan_logdet(self.lib, self.h, ld_c)
        ld_p = self.dm.nt * self.dm.logdet_c0 + e[0]
        dld_p = hyper.dgamma(theta, self.manifold).T @ e[1:4]
        lp, glp, lik = self._scalars(theta)
        st = dict(theta=theta.copy(), ld_c=ld_c, quad_p=quad_p, quad_l=quad_l, bmu=bmu, ld_p=ld_p, dld_p=dld_p,
                  lp=lp, glp=glp, lik=lik)
        self._state = None
        return self._gradient(st, ld_p, dld_p)

    # ------------------------------------------------------------------
    # marginal variances (the selected inverse of Qc at theta)
    def _selected_inverse(self, theta, x_star):
        """Factor Qc(theta) and overwrite the factor with its selected inverse.

        Mirrors DALIA._compute_covariance_latent_parameters (core/dalia.py:912-937).
        For the Gaussian likelihood Qc does not depend on x, so x_star is only
        validated, as the reference requires both or neither.
        """
        if (theta is None) != (x_star is None):
            raise ValueError("BOTH or NEITHER theta and x_star must be provided to compute the marginal variances.")
        if theta is None:
            if self._theta_last is None:
                raise ValueError("no stored theta: call f / value_and_grad first or pass theta and x_star")
            theta = self._theta_last
        self._assemble(np.asarray(theta, dtype=np.float64), "bta", True)
        self._overlap(2)
        _lib.check(self.lib.dh_cholesky(self.h, 1), "dh_cholesky")
        _lib.check(self.lib.dh_selected_inversion(self.h, 1), "dh_selected_inversion")

    def marginal_variances_latent(self, theta=None, x_star=None) -> np.ndarray:
        """diag(Qc(theta)^-1) over all n_latent entries (spatio-temporal, then fixed effects).

        Mirrors DALIA.get_marginal_variances_latent_parameters (core/dalia.py:939-962).
        """
        self._selected_inverse(theta, x_star)
        out = torch.empty(self.dm.n_latent, dtype=torch.float64, device="cuda")
        _lib.check(self.lib.dh_marginal_variances(self.h, out.data_ptr()), "dh_marginal_variances")
        return out.cpu().numpy()

    def marginal_variances_observations(self, theta=None, x_star=None) -> np.ndarray:
        """diag(A Qc(theta)^-1 A^T), one entry per observation.

        Mirrors DALIA.get_marginal_variances_observations (core/dalia.py:964-1016).
        """
        if not hasattr(self, "_a_dev"):
            a = self.data.a_st.tocsr()
            self._a_dev = (torch.from_numpy(a.indptr.astype(np.int64)).cuda(),
                           torch.from_numpy(a.indices.astype(np.int32)).cuda(),
                           torch.from_numpy(a.data.astype(np.float64)).cuda(),
                           torch.from_numpy(np.ascontiguousarray(self.data.x_cov, dtype=np.float64)).cuda())
        self._selected_inverse(theta, x_star)
        rp, col, val, xc = self._a_dev
        out = torch.empty(self.dm.n_obs, dtype=torch.float64, device="cuda")
        _lib.check(self.lib.dh_obs_variances(self.h, self.dm.n_obs, rp.data_ptr(), col.data_ptr(), val.data_ptr(),
                                             xc.data_ptr(), out.data_ptr()), "dh_obs_variances")
        return out.cpu().numpy()


def hessian_stencil(theta, h: float, offdiag_points: int = 4) -> np.ndarray:
    """Stencil points of the FD Hessian (rows), f(theta) first, then +-e_i, then pairs."""
    theta = np.asarray(theta, dtype=np.float64)
    d = theta.size
    E = h * np.eye(d)
    pts = [theta.copy()]
    for i in range(d):
        pts += [theta + E[i], theta - E[i]]
    for i in range(d):
        for j in range(i + 1, d):
            if offdiag_points == 4:
                pts += [theta + E[i] + E[j], theta + E[i] - E[j], theta - E[i] + E[j], theta - E[i] - E[j]]
            elif offdiag_points == 2:
                pts += [theta + E[i] + E[j], theta - E[i] - E[j]]
            else:
                raise ValueError("offdiag_points must be 4 or 2")
    return np.stack(pts)


def hessian_from_values(fv: np.ndarray, d: int, h: float, offdiag_points: int = 4) -> np.ndarray:
    """Assemble H from the stencil values (`core/dalia.py:887-902` for the 4-point form)."""
    H = np.zeros((d, d))
    f0 = fv[0]
    fp, fm = fv[1:1 + 2 * d:2], fv[2:2 + 2 * d:2]
    for i in range(d):
        H[i, i] = (fp[i] - 2 * f0 + fm[i]) / h ** 2
    k = 1 + 2 * d
    for i in range(d):
        for j in range(i + 1, d):
            if offdiag_points == 4:
                fpp, fpm, fmp, fmm = fv[k:k + 4]
                k += 4
                H[i, j] = (fpp - fpm - fmp + fmm) / (4 * h * h)
            else:
                fpp, fmm = fv[k:k + 2]
                k += 2
                H[i, j] = (fpp + fmm - fp[i] - fm[i] - fp[j] - fm[j] + 2 * f0) / (2 * h * h)
            H[j, i] = H[i, j]
    return H


def hessian_diag_stencil(theta, h: float, points: int = 3) -> np.ndarray:
    """Stencil points (rows) of the diagonal of the FD Hessian: theta first, then +-h e_i (points=3,
    1 + 2d rows) and +-2h e_i (points=5, 1 + 4d rows).  The reference's
    `hessian_diag_finite_difference_3pt/_5pt` (`utils/finite_difference_stencils.py:81-171`)
    evaluate f(theta) once per component; it is evaluated once here."""
    theta = np.asarray(theta, dtype=np.float64)
    d = theta.size
    E = h * np.eye(d)
    pts = [theta.copy()]
    for i in range(d):
        pts += [theta + E[i], theta - E[i]]
    if points == 5:
        for i in range(d):
            pts += [theta + 2 * E[i], theta - 2 * E[i]]
    elif points != 3:
        raise ValueError("points must be 3 or 5")
    return np.stack(pts)


def hessian_diag_from_values(fv: np.ndarray, d: int, h: float, points: int = 3) -> np.ndarray:
    """Diagonal of the Hessian from the values at `hessian_diag_stencil` (same formulas as
    `utils/finite_difference_stencils.py:117-118` and `:163-170`)."""
    f0 = fv[0]
    fp, fm = fv[1:1 + 2 * d:2], fv[2:2 + 2 * d:2]
    if points == 3:
        return (fp - 2 * f0 + fm) / h ** 2
    fpp, fmm = fv[1 + 2 * d::2], fv[2 + 2 * d::2]
    return (-fpp + 16 * fp - 30 * f0 + 16 * fm - fmm) / (12 * h ** 2)


def check_hessian(H: np.ndarray) -> np.ndarray:
    """The reference's eigenvalue check of the FD Hessian (`core/dalia.py:904-908`): a message
    when an eigenvalue is negative.  Returns the eigenvalues."""
    ev = np.linalg.eigvalsh(H)
    if np.any(ev < 0):
        print(f"Negative eigenvalues detected: {ev}", flush=True)
    return ev


def compute_covariance_hp(engine, theta, h: float = 5e-3, offdiag_points: int = 4) -> np.ndarray:
    """Covariance of the hyperparameters, the inverse of the FD Hessian of f at theta
    (`DALIA.compute_covariance_hp`, `core/dalia.py:747-765`); `engine` is an `STInla` or a
    `dalia_amd.coreg.CoregInla`.  The Hessian's stencil points run in one HIP graph
    (`dalia_amd.batch.batched_hessians`), the reference's 4-point form by default."""
    from dalia_amd.batch import batched_hessians
    (H,), _ = batched_hessians([engine], [np.asarray(theta, dtype=np.float64)], h=h, offdiag_points=offdiag_points)
    return np.linalg.inv(H)


def hessian_adjoint(engine, theta, h: float = 1e-3) -> np.ndarray:
    """Hessian of f by central differences of the adjoint gradient: 2d gradient
    evaluations instead of the 1 + 2d + 2d(d-1) values of the f-based stencil
    (`core/dalia.py:767-910`); symmetrised.  `engine` is an `STInla` or a
    `dalia_amd.coreg.CoregInla`; the selected inversions destroy its factor."""
    theta = np.asarray(theta, dtype=np.float64)
    d = theta.size
    H = np.zeros((d, d))
    for i in range(d):
        e = np.zeros(d)
        e[i] = h
        _, gp = engine.value_and_grad(theta + e)
        _, gm = engine.value_and_grad(theta - e)
        H[i] = (gp - gm) / (2 * h)
    return 0.5 * (H + H.T)


class FTheta(torch.autograd.Function):
    """f(theta) as an autograd node; backward runs the adjoint on the device.

    Forward factors Qc, solves for mu and evaluates the log-determinants;
    backward overwrites the factor with the selected inverse and contracts it.
    The engine is an `STInla` or a `dalia_amd.coreg.CoregInla`.
    """

    @staticmethod
    def forward(ctx, theta: torch.Tensor, engine):
        th = theta.detach().cpu().double().numpy()
        f_val = engine.forward(th)
        if f_val is None:  # no closed form: value needs the BT factorisation of Qp
            f_val = engine.f(th)
            engine.forward(th)
        ctx.engine = engine
        ctx.theta = th
        return torch.tensor(f_val, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        st = ctx.engine._state
        if st is None or not np.array_equal(st["theta"], ctx.theta):
            ctx.engine.forward(ctx.theta)
        _, g = ctx.engine.backward()
        return grad_out * torch.as_tensor(g, dtype=torch.float64), None
