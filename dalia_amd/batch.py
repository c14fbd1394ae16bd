"""Many f(theta) evaluations over model instances of uneven size in one HIP graph.

The evaluations left after the adjoint gradient (line-search points, the FD
Hessian stencil at the mode: `core/dalia.py:767-910`) are independent.  The
reference spreads them over MPI ranks (F-level groups, `core/dalia.py:541-573`);
on one GPU this module captures all of them in one graph:

* every model instance (its own ns, nt, nb) has its own engine, buffers and
  stream;
* the evaluations of one instance run back to back on its stream (they share
  its buffers); instances run concurrently (fork/join inside the capture);
* each evaluation first copies its theta-dependent parameters (13 for the
  univariate model, nv*nv*12 + nv for the coregional one) from a row of a
  pinned host table (a captured host-to-device copy), then runs the plain
  f(theta) kernels and writes its 5 scalars to its own device slot.

The graph is captured once for a given assignment of evaluations to
instances and replayed for every new set of theta values; the host only
fills the pinned table and reads the scalar slots back.
"""

from __future__ import annotations

import numpy as np
import torch

from dalia_amd.inla import (STInla, check_hessian, hessian_diag_from_values, hessian_diag_stencil,
                            hessian_from_values, hessian_stencil)


class BatchedF:
    def __init__(self, engines: list[STInla], assignment: list[int], tune: bool = False, fused: bool = False):
        """engines: one STInla (or CoregInla) per model instance; assignment[e] = instance of
        evaluation e.  tune: before the capture, let the autotuner pick each instance's pobtaf
        schedule on that instance's own matrix (cached per shape).  fused: run the evaluations in
        rounds (round r: the r-th evaluation of every instance) on one stream, the factorizations of
        a round as one persistent launch over all its instances (`dh_multi_cholesky`), instead of
        one stream per instance."""
        if tune:
            from dalia_amd import autotune
            self.choices = [autotune.tune(eng, workgroups=len(engines) > 1 and not fused) for eng in engines]
        self.engines = engines
        self.fused = bool(fused)
        self._multi = {}
        self.assignment = list(assignment)
        E = len(self.assignment)
        width = max(int(eng.params.numel()) for eng in engines)
        self.table = torch.zeros((E, width), dtype=torch.float64).pin_memory()
        self.slots = torch.zeros((E, 5), dtype=torch.float64, device="cuda")
        self.streams = [torch.cuda.Stream() for _ in engines]
        self.graph = None
        self.thetas = [None] * E

    def _enqueue_all(self, capture_stream, first_only: bool = False):
        """Every evaluation on its instance's stream (first_only: one per instance, the warm-up
        that loads the kernels before the capture; the evaluations of an instance all run the same
        launches on the same buffers)."""
        for k, eng in enumerate(self.engines):
            s = self.streams[k]
            s.wait_stream(capture_stream)
            with torch.cuda.stream(s):
                eng.set_stream(s)
                for e, inst in enumerate(self.assignment):
                    if inst != k:
                        continue
                    eng.params.copy_(self.table[e, :eng.params.numel()], non_blocking=True)
                    eng._enqueue_f(None, out=self.slots[e])
                    if first_only:
                        break
            capture_stream.wait_stream(s)

    # -- fused rounds -------------------------------------------------------
    def _rounds(self):
        per = [[e for e, inst in enumerate(self.assignment) if inst == k] for k in range(len(self.engines))]
        R = max(len(p) for p in per)
        return [[(k, per[k][r]) for k in range(len(self.engines)) if r < len(per[k])] for r in range(R)]

    def _multi_for(self, insts):
        """The launch context of the factorizations of these instances (created outside capture)."""
        key = tuple(insts)
        if key not in self._multi:
            from dalia_amd import _lib
            import ctypes
            arr = (ctypes.c_void_p * len(insts))(*[self.engines[k].h.value for k in insts])
            mc = ctypes.c_void_p()
            lib = self.engines[insts[0]].lib
            _lib.check(lib.dh_multi_create(arr, len(insts), ctypes.byref(mc)), "dh_multi_create")
            self._multi[key] = mc
        return self._multi[key]

    def _enqueue_fused(self, stream, first_only: bool = False):
        from dalia_amd import _lib
        import ctypes
        for eng in self.engines:
            eng.set_stream(stream)
        sp = ctypes.c_void_p(stream.cuda_stream)
        with torch.cuda.stream(stream):
            for rnd in self._rounds()[:1] if first_only else self._rounds():
                insts = [k for k, _ in rnd]
                mc = self._multi_for(insts)
                lib = self.engines[insts[0]].lib
                for k, e in rnd:
                    eng = self.engines[k]
                    eng.params.copy_(self.table[e, :eng.params.numel()], non_blocking=True)
                    eng._fused_a(self.slots[e])
                _lib.check(lib.dh_multi_cholesky(mc, 1, sp), "dh_multi_cholesky")
                for k, e in rnd:
                    self.engines[k]._fused_b(self.slots[e])
                _lib.check(lib.dh_multi_cholesky(mc, 0, sp), "dh_multi_cholesky")
                for k, e in rnd:
                    self.engines[k]._fused_c(self.slots[e])

    def close(self):
        for mc in self._multi.values():
            self.engines[0].lib.dh_multi_destroy(mc)
        self._multi = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def capture(self):
        for e, inst in enumerate(self.assignment):  # valid parameters for the warm-up pass
            row = self.engines[inst].param_row(self.engines[inst].data.theta)
            self.table[e, :row.size].copy_(torch.from_numpy(row))
        cur = torch.cuda.current_stream()
        if self.fused:
            for rnd in self._rounds():  # every launch context before the capture
                self._multi_for([k for k, _ in rnd])
            self._enqueue_fused(cur, first_only=True)
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._enqueue_fused(torch.cuda.current_stream())
            for eng in self.engines:
                eng.set_stream(torch.cuda.current_stream())
            self.graph = g
            return g
        self._enqueue_all(cur, first_only=True)  # eager warm-up (allocator, kernels loaded)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._enqueue_all(torch.cuda.current_stream())
        for eng in self.engines:
            eng.set_stream(torch.cuda.current_stream())
        self.graph = g
        return g

    def __call__(self, thetas) -> np.ndarray:
        """f at thetas[e] on instance assignment[e]; one graph launch for all."""
        if self.graph is None:
            self.capture()
        assert len(thetas) == len(self.assignment)
        for e, inst in enumerate(self.assignment):
            row = self.engines[inst].param_row(thetas[e])
            self.table[e, :row.size].copy_(torch.from_numpy(row))
        self.graph.replay()
        s = self.slots.cpu().numpy()
        return np.array([self.engines[inst]._f_from_scalars(np.asarray(thetas[e]), s[e])
                         for e, inst in enumerate(self.assignment)])


def batched_hessians(engines: list[STInla], thetas, h: float = 5e-3, offdiag_points: int = 2, tune: bool = False,
                     fused: bool = False):
    """FD Hessians of every instance at its own theta, all stencil points in one graph launch."""
    stencils = [hessian_stencil(t, h, offdiag_points) for t in thetas]
    assignment = [k for k, st in enumerate(stencils) for _ in range(len(st))]
    bf = BatchedF(engines, assignment, tune=tune, fused=fused)
    fv = bf([p for st in stencils for p in st])
    out, k0 = [], 0
    for st, t in zip(stencils, thetas):
        H = hessian_from_values(fv[k0:k0 + len(st)], len(t), h, offdiag_points)
        check_hessian(H)
        out.append(H)
        k0 += len(st)
    return out, bf


def batched_hessian_diags(engines, thetas, h: float = 5e-3, points: int = 3, tune: bool = False,
                          fused: bool = False):
    """Diagonals of the FD Hessians of every instance at its own theta (1 + 2d stencil points
    each for points=3, the reference's `hessian_diag_finite_difference_3pt`), all in one graph
    launch.  Returns (list of diagonals, the BatchedF)."""
    stencils = [hessian_diag_stencil(t, h, points) for t in thetas]
    assignment = [k for k, st in enumerate(stencils) for _ in range(len(st))]
    bf = BatchedF(engines, assignment, tune=tune, fused=fused)
    fv = bf([p for st in stencils for p in st])
    out, k0 = [], 0
    for st, t in zip(stencils, thetas):
        out.append(hessian_diag_from_values(fv[k0:k0 + len(st)], len(t), h, points))
        k0 += len(st)
    return out, bf
