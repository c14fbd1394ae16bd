"""`HipBTASolver`: the drop-in for DALIA's `SerinvSolver` on one MI355X.

Mirrors `dalia.core.solver.Solver` (`src/dalia/core/solver.py:9-63`) and
`SerinvSolver` (`src/dalia/solvers/structured_solver.py:19-549`): the same
constructor arguments, the same `cholesky / solve / logdet /
selected_inversion / _structured_to_spmatrix / get_solver_memory` methods with
the same `sparsity` argument ("bt" or "bta"), the same errors (`ValueError`
on an unknown sparsity and on a NaN log-determinant) and the same public
timers `t_cholesky` and `t_solve`.  The arithmetic runs in
`libdalia_hip.so`; the dense blocks live in HBM and are exposed as strided
torch views `A_diagonal_blocks` (n, b, b), `A_lower_diagonal_blocks`
(n-1, b, b), `A_arrow_bottom_blocks` (n, a, b) and `A_arrow_tip_block` (a, a).
"""

from __future__ import annotations

import ctypes
import time

import numpy as np
import scipy.sparse as sps
import torch

from dalia_amd import _lib


def _check_sparsity(sparsity: str) -> int:
    if sparsity not in _lib.SPARSITY:
        raise ValueError(f"Unknown sparsity pattern: {sparsity}. Use 'bt' or 'bta'.")
    return _lib.SPARSITY[sparsity]


class _DevArray:
    """Minimal `__cuda_array_interface__` wrapper around a device pointer."""

    def __init__(self, ptr: int, shape, strides_elems):
        self.__cuda_array_interface__ = {
            "data": (int(ptr), False),
            "shape": tuple(int(s) for s in shape),
            "strides": tuple(int(s) * 8 for s in strides_elems),
            "typestr": "<f8",
            "version": 2,
        }


def device_view(ptr: int, shape, strides) -> torch.Tensor:
    """Strided float64 torch view of library-owned device memory (no copy)."""
    return torch.as_tensor(_DevArray(ptr, shape, strides), device="cuda")


class HipBTASolver:
    """Block-tridiagonal-arrowhead Cholesky solver on the GPU (FP64)."""

    def __init__(self, config=None, diagonal_blocksize: int = None, n_diag_blocks: int = None,
                 arrowhead_blocksize: int = 0, stream: torch.cuda.Stream | None = None, **kwargs):
        self.config = config
        self.diagonal_blocksize = int(diagonal_blocksize)
        self.n_diag_blocks = int(n_diag_blocks)
        self.arrowhead_blocksize = int(arrowhead_blocksize)
        self.lib = _lib.load()
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        _lib.check(self.lib.dh_create(self.n_diag_blocks, self.diagonal_blocksize, self.arrowhead_blocksize,
                                      ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h)), "dh_create")
        self.h = h
        lay = _lib.DhLayout()
        _lib.check(self.lib.dh_get_layout(self.h, ctypes.byref(lay)), "dh_get_layout")
        self.layout = lay
        n, b, a, bp, ap = lay.n, lay.b, lay.a, lay.bp, lay.ap
        self.A_diagonal_blocks = device_view(lay.diag, (n, b, b), (bp * bp, bp, 1))
        self.A_lower_diagonal_blocks = (device_view(lay.lower, (n - 1, b, b), (bp * bp, bp, 1))
                                        if n > 1 else torch.empty((0, b, b), dtype=torch.float64, device="cuda"))
        self.A_arrow_bottom_blocks = device_view(lay.arrow, (n, a, b), (ap * bp, bp, 1)) if a else None
        self.A_arrow_tip_block = device_view(lay.tip, (a, a), (ap, 1)) if a else None
        self._scal = torch.zeros(8, dtype=torch.float64, device="cuda")
        self.total_bytes = 0
        self.t_cholesky = 0.0
        self.t_solve = 0.0

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.dh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- Solver interface ------------------------------------------------------
    def _spmatrix_to_structured(self, A, sparsity: str) -> None:
        """Scatter a host sparse matrix into the device blocks."""
        sp = _check_sparsity(sparsity)
        coo = sps.coo_matrix(A)
        coo.sum_duplicates()
        rows = torch.as_tensor(coo.row.astype(np.int64), device="cuda")
        cols = torch.as_tensor(coo.col.astype(np.int64), device="cuda")
        vals = torch.as_tensor(coo.data.astype(np.float64), device="cuda")
        _lib.check(self.lib.dh_scatter_coo(self.h, int(coo.nnz), rows.data_ptr(), cols.data_ptr(),
                                           vals.data_ptr(), sp), "dh_scatter_coo")
        self._keep = (rows, cols, vals)

    def cholesky(self, A=None, sparsity: str = "bta") -> None:
        """Factor in place; `A` (scipy sparse) is scattered first, `None` keeps the blocks."""
        sp = _check_sparsity(sparsity)
        if A is not None:
            self._spmatrix_to_structured(A, sparsity)
        tic = time.perf_counter()
        _lib.check(self.lib.dh_cholesky(self.h, sp), "dh_cholesky")
        self.stream.synchronize()
        self.t_cholesky += time.perf_counter() - tic

    def cholesky_async(self, sparsity: str = "bta") -> None:
        _lib.check(self.lib.dh_cholesky(self.h, _check_sparsity(sparsity)), "dh_cholesky")

    def solve(self, rhs, sparsity: str = "bta"):
        """Solve in place (device float64 tensor of length n*b (+a)); returns rhs."""
        sp = _check_sparsity(sparsity)
        tic = time.perf_counter()
        host = isinstance(rhs, np.ndarray)
        x = torch.as_tensor(rhs, device="cuda").contiguous() if host else rhs
        if x.dtype != torch.float64 or not x.is_cuda or not x.is_contiguous():
            raise ValueError("rhs must be a contiguous float64 CUDA tensor or a numpy array")
        _lib.check(self.lib.dh_solve(self.h, x.data_ptr(), sp), "dh_solve")
        self.stream.synchronize()
        self.t_solve += time.perf_counter() - tic
        if host:
            rhs[:] = x.cpu().numpy()
        return rhs

    def solve_async(self, x: torch.Tensor, sparsity: str = "bta") -> None:
        _lib.check(self.lib.dh_solve(self.h, x.data_ptr(), _check_sparsity(sparsity)), "dh_solve")

    def logdet_async(self, sparsity: str, out: torch.Tensor) -> None:
        _lib.check(self.lib.dh_logdet(self.h, _check_sparsity(sparsity), out.data_ptr()), "dh_logdet")

    def logdet(self, sparsity: str = "bta") -> float:
        self.logdet_async(sparsity, self._scal[:1])
        v = float(self._scal[0].item())
        if np.isnan(v):
            raise ValueError("Logdet is NaN. Check the input matrix.")
        return v

    def set_overlap(self, mode: int) -> bool:
        """Work that later calls need, run beside the factorization on a second stream (`dh_set_tuning`
        key 4): 0 none, 1 the inverse of the diagonal factors (`solve`), 2 also the products of the
        factor alone (`selected_inversion`).  Mode 2 reserves 2 n bp^2 doubles; returns False, and
        keeps the previous mode, when they do not fit."""
        rc = self.lib.dh_set_tuning(self.h, 4, int(mode))
        if rc == -2 and int(mode) == 2:
            return False
        _lib.check(rc, "dh_set_tuning")
        return True

    def selected_inversion(self, sparsity: str = "bta") -> None:
        _lib.check(self.lib.dh_selected_inversion(self.h, _check_sparsity(sparsity)), "dh_selected_inversion")

    def info(self) -> int:
        v = ctypes.c_int()
        _lib.check(self.lib.dh_get_info(self.h, ctypes.byref(v)), "dh_get_info")
        return v.value

    def _structured_to_spmatrix(self, A, sparsity: str):
        """Blocks read on the lower-triangular pattern of A, symmetrised (`structured_solver.py:470-538`)."""
        _check_sparsity(sparsity)
        n, b, a = self.n_diag_blocks, self.diagonal_blocksize, self.arrowhead_blocksize
        B = sps.coo_matrix(sps.tril(sps.csc_matrix(A)))
        r, c = B.row.astype(np.int64), B.col.astype(np.int64)
        D = self.A_diagonal_blocks.cpu().numpy()
        Lw = self.A_lower_diagonal_blocks.cpu().numpy()
        nb = n * b
        br = np.where(r < nb, r // b, n)
        bc = np.where(c < nb, c // b, n)
        lr = np.where(r < nb, r % b, r - nb)
        lc = np.where(c < nb, c % b, c - nb)
        data = np.zeros(r.size)
        m = (br == bc) & (br < n)
        data[m] = D[br[m], lr[m], lc[m]]
        m2 = (br == bc + 1) & (br < n)
        data[m2] = Lw[bc[m2], lr[m2], lc[m2]]
        keep = m | m2
        if sparsity == "bta" and a:
            R = self.A_arrow_bottom_blocks.cpu().numpy()
            T = self.A_arrow_tip_block.cpu().numpy()
            m3 = (br == n) & (bc < n)
            data[m3] = R[bc[m3], lr[m3], lc[m3]]
            m4 = (br == n) & (bc == n)
            data[m4] = T[lr[m4], lc[m4]]
            keep = keep | m3 | m4
        out = sps.coo_matrix((data[keep], (r[keep], c[keep])), shape=B.shape).tocsc()
        return out + sps.tril(out, k=-1).T

    def get_solver_memory(self) -> int:
        self.total_bytes += int(self.lib.dh_memory_bytes(self.h))
        return self.total_bytes
