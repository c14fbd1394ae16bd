"""f(theta) and its adjoint gradient with the factor checkpointed instead of stored.

The reference keeps the whole BTA factor in the solver's buffers
(`src/dalia/solvers/structured_solver.py:38-60`), 3 n b^2 doubles with the inverse this package
adds; at ns = 4096, nt = 250 that is 104 GB.  `STInlaCkpt` keeps a ring of K + 2 block slots and,
per segment of K time blocks, only the boundary factor L_{j,j-1}, L_{a,j-1} that the next segment
needs (`dh_ckpt_*` in include/dalia_hip.h):

* forward pass: assemble, factor and sweep forward segment by segment, saving the boundary factor
  of every segment, then factor the tip;
* backward pass: for the segments in reverse, re-assemble the blocks from theta, restore the
  boundary factor, re-factor, finish the solve and (for the gradient) form the segment's selected
  inverse and contract it against dQ/dtheta at once.

So an evaluation factors Qc twice (the last segment once) and never holds more than K + 2 blocks
of the factor; the prior's log-determinant comes from the Kronecker-eigen closed form or, for
plain f, from a segment-by-segment BT factorization.  `make_engine` picks this engine when the
full handle would not fit a memory budget.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from dalia_amd import _lib, hyper
from dalia_amd.inla import STInla
from dalia_amd.model import DeviceSTModel


def full_handle_bytes(n: int, b: int, a: int, twisted: bool = False) -> int:
    """Device bytes of a full `dh_create` handle (factor, inverse, scratch) for n blocks of b, an
    upper bound: two sets of G = min(n, 8) selected-inversion slots (W, S, Wa) and the
    factorization's own; twisted: with the transposed operand copies of the two-ended
    factorization (2 n bp^2 + 64 n bp doubles)."""
    bp = (b + 63) // 64 * 64
    bb = bp * bp
    G = min(n, 8)
    extra = (2 * n * bb + 64 * n * bp) if twisted else 0
    return 8 * ((3 * n + 3 * G) * bb + 2 * G * bp * (bp + 8) + 2 * G * 64 * bp + (n + 2) * 64 * bp + extra)


def ckpt_bytes(n: int, b: int, a: int, K: int) -> int:
    """Device bytes of a checkpoint context with segments of K blocks."""
    bp = (b + 63) // 64 * 64
    S = (n + K - 1) // K
    return full_handle_bytes(K + 2, b, a) + 8 * (S * (bp * bp + 64 * bp) + n * bp * 4)


def make_engine(data, memory_budget: float | None = None, checkpoint: int | None = None, **kw):
    """An `STInla` when the full handle fits `memory_budget` bytes (or no budget is given; two-ended
    when its transposed operand copies fit too), else an `STInlaCkpt` with segments of 16 time
    blocks (or fewer, down to the K that needs the least memory, when 16 does not fit).
    checkpoint=K forces the checkpointed engine with segments of K."""
    n, b, a = data.nt, data.ns, data.nb
    if checkpoint is not None:
        return STInlaCkpt(data, K=int(checkpoint), **kw)
    if memory_budget is None:
        return STInla(data, **kw)
    twisted = kw.get("twisted", True) and n >= 3
    if full_handle_bytes(n, b, a, twisted) <= memory_budget:
        return STInla(data, **kw)
    if twisted and full_handle_bytes(n, b, a) <= memory_budget:  # one-ended: no transposed copies
        return STInla(data, **{**kw, "twisted": False})
    ks = [K for K in range(1, min(16, n) + 1) if ckpt_bytes(n, b, a, K) <= memory_budget]
    if not ks:
        raise MemoryError(f"no checkpoint segment length fits {memory_budget / 1e9:.1f} GB "
                          f"(least: {min(ckpt_bytes(n, b, a, K) for K in range(1, n + 1)) / 1e9:.1f} GB)")
    return STInlaCkpt(data, K=max(ks), **kw)


class STInlaCkpt(STInla):
    """`STInla` with the factor checkpointed every K time blocks and recomputed in the backward pass.

    Same results and interface for f, value_and_grad, fd_gradient, hessian, the autograd node and
    the latent marginal variances; not capturable in a HIP graph."""

    def __init__(self, data, K: int = 16, device="cuda", closed_form_prior: bool = True,
                 pattern_selinv: bool = True):
        """pattern_selinv: each segment's Takahashi recursion in congruence form with the coupling
        blocks formed only on the matrix's tiles (`dh_ckpt_set_tuning` key 10), as `STInla`."""
        self.data = data
        self.manifold = getattr(data, "manifold", "plane")
        if self.manifold not in hyper.MANIFOLDS:
            raise ValueError(f"Manifold not supported: {self.manifold}")
        self.dm = DeviceSTModel(data, device=device, eigen=closed_form_prior)
        self.lib = _lib.load()
        self.K = int(K)
        stream = torch.cuda.current_stream()
        ck = ctypes.c_void_p()
        _lib.check(self.lib.dh_ckpt_create(data.nt, data.ns, data.nb, self.K, ctypes.c_void_p(stream.cuda_stream),
                                           ctypes.byref(ck)), "dh_ckpt_create")
        self.ck = ck
        _lib.check(self.lib.dh_ckpt_set_tuning(ck, 10, 1 if pattern_selinv else 0), "dh_ckpt_set_tuning")
        self.h = ctypes.c_void_p(self.lib.dh_ckpt_vectors(ck))
        self.solver = None
        N = self.dm.n_latent
        self.mu = torch.zeros(N, dtype=torch.float64, device="cuda")
        self.yp = torch.zeros(N, dtype=torch.float64, device="cuda")
        self.yl = torch.zeros(N, dtype=torch.float64, device="cuda")
        self.scal = torch.zeros(16, dtype=torch.float64, device="cuda")
        nq = 2 * data.nt - 1
        self.cx = torch.zeros(nq * 4, dtype=torch.float64, device="cuda")
        self.cmu = torch.zeros(nq * 4, dtype=torch.float64, device="cuda")
        self.eig = torch.zeros(4, dtype=torch.float64, device="cuda")
        self.params = torch.zeros(13, dtype=torch.float64, device="cuda")
        self.params_h = torch.zeros(13, dtype=torch.float64).pin_memory()
        self._state = None
        self._graph = None
        self.n_f_evals = 0
        self._theta_last = None
        self._ov_max = 0
        self._ov = 0

    def close(self):
        if getattr(self, "ck", None) is not None and self.ck.value:
            self.lib.dh_ckpt_destroy(self.ck)
            self.ck = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def memory_bytes(self) -> int:
        return int(self.lib.dh_ckpt_memory_bytes(self.ck))

    def set_stream(self, stream) -> None:
        _lib.check(self.lib.dh_ckpt_set_stream(self.ck, ctypes.c_void_p(stream.cuda_stream)), "dh_ckpt_set_stream")

    def _overlap(self, mode: int) -> None:
        return None

    def capture_f(self):
        raise NotImplementedError("the checkpointed engine is not captured in a HIP graph")

    def _passes(self, theta, selinv: bool, var=None):
        """Qc: forward pass (logdet to scal[0]) and backward pass (mu; with selinv the
        contractions into cx and scal[5], and the marginal variances into var)."""
        L, m = self.lib, self.dm.ref()
        self._set_params(theta)
        N = self.dm.n_latent
        _lib.check(L.dh_scale_dev(self.h, self.dm.aty.data_ptr(), self.params[12:13].data_ptr(), self.mu.data_ptr(), N),
                   "dh_scale_dev")
        _lib.check(L.dh_ckpt_forward(self.ck, m, self.params.data_ptr(), 1, self.mu.data_ptr(),
                                     self.scal[0:1].data_ptr()), "dh_ckpt_forward")
        if selinv:
            _lib.check(L.dh_ckpt_backward(self.ck, m, self.params.data_ptr(), self.mu.data_ptr(), 1,
                                          self.cx.data_ptr(), self.scal[5:6].data_ptr(),
                                          None if var is None else var.data_ptr()), "dh_ckpt_backward")
        else:
            _lib.check(L.dh_ckpt_backward(self.ck, m, self.params.data_ptr(), self.mu.data_ptr(), 0, None, None, None),
                       "dh_ckpt_backward")
        _lib.check(L.dh_qc_matvec(self.h, m, self.params.data_ptr(), self.mu.data_ptr(), self.yp.data_ptr(),
                                  self.yl.data_ptr()), "dh_qc_matvec")
        _lib.check(L.dh_dot(self.h, self.mu.data_ptr(), self.yp.data_ptr(), N, self.scal[1:2].data_ptr()), "dh_dot")
        _lib.check(L.dh_dot(self.h, self.mu.data_ptr(), self.yl.data_ptr(), N, self.scal[2:3].data_ptr()), "dh_dot")
        _lib.check(L.dh_dot(self.h, self.mu.data_ptr(), self.dm.aty.data_ptr(), N, self.scal[3:4].data_ptr()),
                   "dh_dot")
        v = ctypes.c_int()
        if L.dh_ckpt_get_info(self.ck, ctypes.byref(v)) == -4:  # also a give-up of the recomputation
            raise _lib.DhError("dh_ckpt_forward / dh_ckpt_backward failed: a fused launch gave up waiting "
                               "for a factored tile")

    def _raise_if_nan(self, *values):
        if any(np.isnan(v) for v in values):
            v = ctypes.c_int()
            if self.lib.dh_ckpt_get_info(self.ck, ctypes.byref(v)) == -4 or any(_lib.is_giveup_nan(x) for x in values):
                raise _lib.DhError("dh_ckpt_forward failed: a fused launch gave up waiting for a factored tile")
            raise ValueError("Logdet is NaN. Check the input matrix.")

    def f(self, theta) -> float:
        """Plain f(theta): the checkpointed BTA factorization twice (forward, backward) and a
        segment-by-segment BT factorization for the prior's log-determinant, as the reference's
        f factors Qc and Qp (`core/dalia.py:603-745`)."""
        theta = np.asarray(theta, dtype=np.float64)
        self._passes(theta, selinv=False)
        _lib.check(self.lib.dh_ckpt_forward(self.ck, self.dm.ref(), self.params.data_ptr(), 0, None,
                                            self.scal[4:5].data_ptr()), "dh_ckpt_forward")
        s = self.scal[:5].cpu().numpy()
        self._raise_if_nan(s[0], s[4])
        self.n_f_evals += 1
        self._state = None
        return self._f_from_scalars(theta, s)

    def forward(self, theta):
        """Both passes with the selected inverse contracted on the fly; the value of f."""
        theta = np.asarray(theta, dtype=np.float64)
        if not self.dm.closed_form:
            raise ValueError("the checkpointed adjoint needs the closed-form prior term "
                             "(lumped c0, g2 = g1 c0^-1 g1, g3 = g1 c0^-1 g2)")
        self._passes(theta, selinv=True)
        _lib.check(self.lib.dh_contract_fem_vec(self.h, self.dm.ref(), self.mu.data_ptr(), self.cmu.data_ptr()),
                   "dh_contract_fem_vec")
        self._prior_logdet_terms(theta)
        s = self.scal[:4].cpu().numpy()
        e = self.eig.cpu().numpy()
        ld_c, quad_p, quad_l, bmu = s
        self._raise_if_nan(ld_c)
        ld_p = self.dm.nt * self.dm.logdet_c0 + e[0]
        dld_p = hyper.dgamma(theta, self.manifold).T @ e[1:4]
        lp, glp, lik = self._scalars(theta)
        self._state = dict(theta=theta.copy(), ld_c=ld_c, quad_p=quad_p, quad_l=quad_l, bmu=bmu, ld_p=ld_p,
                           dld_p=dld_p, lp=lp, glp=glp, lik=lik)
        return float(0.5 * ld_c - 0.5 * (quad_p + quad_l) - (lp + lik + 0.5 * ld_p))

    def value_and_grad(self, theta):
        self.forward(theta)
        return self.backward()

    def backward(self):
        """Gradient of the last `forward` (its contractions are already on the device)."""
        st = self._state
        if st is None:
            raise RuntimeError("backward() needs a preceding forward() on this engine")
        return self._gradient(st, st["ld_p"], st["dld_p"])

    def marginal_variances_latent(self, theta=None, x_star=None) -> np.ndarray:
        """diag(Qc(theta)^-1), formed segment by segment in the backward pass
        (`core/dalia.py:939-962`)."""
        if (theta is None) != (x_star is None):
            raise ValueError("BOTH or NEITHER theta and x_star must be provided to compute the marginal variances.")
        if theta is None:
            if self._theta_last is None:
                raise ValueError("no stored theta: call f / value_and_grad first or pass theta and x_star")
            theta = self._theta_last
        out = torch.empty(self.dm.n_latent, dtype=torch.float64, device="cuda")
        self._passes(np.asarray(theta, dtype=np.float64), selinv=True, var=out)
        self._raise_if_nan(float(self.scal[0].item()))
        return out.cpu().numpy()

    def marginal_variances_observations(self, theta=None, x_star=None):
        raise NotImplementedError("observation variances need the whole selected inverse at once")
