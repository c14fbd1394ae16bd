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
    def _order(self, sparsity: str) -> int:
        """Order of the matrix (and length of the right-hand side) for a sparsity."""
        n, b, a = self.n_diag_blocks, self.diagonal_blocksize, self.arrowhead_blocksize
        return n * b + (a if sparsity == "bta" else 0)

    def _spmatrix_to_structured(self, A, sparsity: str) -> None:
        """Scatter a sparse matrix into the device blocks.

        A: a scipy sparse matrix (DALIA's NumPy backend), or a device sparse matrix with
        `tocoo()` whose row / col / data arrays expose `__cuda_array_interface__` (cupyx, DALIA's
        CuPy backend: `src/dalia/__init__.py:20-54`), read in place on the device.  The order must
        be n*b + a (for "bt" also n*b: the arrowhead rows are ignored)."""
        sp = _check_sparsity(sparsity)
        N = self._order("bta")
        shape = tuple(int(x) for x in A.shape)
        ok = {(N, N)} if sparsity == "bta" else {(N, N), (self._order("bt"),) * 2}
        if shape not in ok:
            raise ValueError(f"matrix of shape {shape} does not match the solver's blocks "
                             f"(n*b + a = {N}, sparsity {sparsity!r})")
        if sps.issparse(A):
            coo = sps.coo_matrix(A)
            coo.sum_duplicates()
            rows = torch.as_tensor(coo.row.astype(np.int64), device="cuda")
            cols = torch.as_tensor(coo.col.astype(np.int64), device="cuda")
            vals = torch.as_tensor(coo.data.astype(np.float64), device="cuda")
        elif hasattr(A, "tocoo"):
            coo = A.tocoo()
            if hasattr(coo, "sum_duplicates"):
                coo.sum_duplicates()
            rows = torch.as_tensor(coo.row, device="cuda").to(torch.int64).contiguous()
            cols = torch.as_tensor(coo.col, device="cuda").to(torch.int64).contiguous()
            vals = torch.as_tensor(coo.data, device="cuda").to(torch.float64).contiguous()
        else:
            raise TypeError("A must be a scipy sparse matrix or a device sparse matrix with tocoo()")
        nnz = int(vals.numel())
        if rows.numel() != nnz or cols.numel() != nnz:
            raise ValueError("COO row, column and value arrays differ in length")
        _lib.check(self.lib.dh_scatter_coo(self.h, nnz, rows.data_ptr(), cols.data_ptr(),
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

    def _rhs_view(self, rhs, sparsity: str):
        """(device view, whether rhs is a host array) of a right-hand side, checked."""
        host = isinstance(rhs, np.ndarray)
        if host or isinstance(rhs, torch.Tensor):
            length = int(rhs.size if host else rhs.numel())
        elif hasattr(rhs, "__cuda_array_interface__"):  # CuPy (DALIA's GPU backend)
            length = int(np.prod(rhs.__cuda_array_interface__["shape"]))
        else:
            raise ValueError("rhs must be a numpy array, a CUDA tensor or a CUDA array")
        if length != self._order(sparsity):  # checked before anything touches the device
            raise ValueError(f"rhs of length {length} does not match the solver's blocks "
                             f"({self._order(sparsity)} for sparsity {sparsity!r})")
        if host:
            if rhs.dtype != np.float64:
                raise ValueError("rhs must be float64")
            x = torch.as_tensor(rhs, device="cuda").contiguous()
        else:
            x = rhs if isinstance(rhs, torch.Tensor) else torch.as_tensor(rhs, device="cuda")  # zero-copy view
        if x.dtype != torch.float64 or not x.is_cuda or not x.is_contiguous():
            raise ValueError("rhs must be a contiguous float64 CUDA tensor or a numpy array")
        return x, host

    def solve(self, rhs, sparsity: str = "bta"):
        """Solve in place (numpy array, CUDA tensor or CUDA array of length n*b (+a)); returns rhs."""
        sp = _check_sparsity(sparsity)
        tic = time.perf_counter()
        x, host = self._rhs_view(rhs, sparsity)
        _lib.check(self.lib.dh_solve(self.h, x.data_ptr(), sp), "dh_solve")
        self.stream.synchronize()
        self.t_solve += time.perf_counter() - tic
        if host:
            rhs[:] = x.cpu().numpy()
        return rhs

    def solve_async(self, x: torch.Tensor, sparsity: str = "bta") -> None:
        sp = _check_sparsity(sparsity)
        if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float64 and x.is_contiguous()
                and x.numel() == self._order(sparsity)):
            raise ValueError(f"solve_async needs a contiguous float64 CUDA tensor of length {self._order(sparsity)}")
        _lib.check(self.lib.dh_solve(self.h, x.data_ptr(), sp), "dh_solve")

    def logdet_async(self, sparsity: str, out: torch.Tensor) -> None:
        _lib.check(self.lib.dh_logdet(self.h, _check_sparsity(sparsity), out.data_ptr()), "dh_logdet")

    def logdet(self, sparsity: str = "bta") -> float:
        self.logdet_async(sparsity, self._scal[:1])
        v = float(self._scal[0].item())
        if np.isnan(v):
            _lib.raise_nan_logdet(self.lib, self.h, v)
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

    def set_twisted(self, on: bool = True) -> None:
        """Two-ended factorization for the next `cholesky` calls (`dh_set_tuning` key 7): blocks
        0 .. m-1 eliminated downward and n-1 .. m+1 upward in one persistent launch, meeting at block
        m = (n - 1) // 2, which halves the chain of diagonal tiles.  logdet, solve and the selected
        inverse are those of the one-ended factor; the factor blocks themselves differ (the buffers
        then hold a factor of the matrix with the blocks after m in reverse order).  Needs n >= 3."""
        _lib.check(self.lib.dh_set_tuning(self.h, 7, 1 if on else 0), "dh_set_tuning")

    def set_tile_profile(self, on: bool = True) -> None:
        """Zero-tile profiles in the persistent factorization (`dh_set_tuning` key 9, on by default):
        the sub-diagonal blocks are scanned for the first nonzero 64x64 tile of each tile row before
        each launch, and the products with the tiles left of it, which stay zero, are skipped."""
        _lib.check(self.lib.dh_set_tuning(self.h, 9, 1 if on else 0), "dh_set_tuning")

    def set_selected_inversion_pattern(self, on: bool = True) -> None:
        """The selected inversion in congruence form (`dh_set_tuning` key 10, off by default):
        X_ii = inv(L_ii)^T (I + M) inv(L_ii) with M = L^T (X_qq L + X_aq^T L_a) + L_a^T Ya, and the
        sub-diagonal blocks X_{i+1,i} formed only on the 64x64 tiles where the matrix's coupling block
        has entries (other tiles keep stale values).  For callers that read the selected inverse on
        the matrix's pattern only; the reference's selected_inversion returns whole blocks, which
        stays the default.  Needs the zero-tile profiles (`set_tile_profile`), else the dense form
        runs."""
        _lib.check(self.lib.dh_set_tuning(self.h, 10, 1 if on else 0), "dh_set_tuning")

    def set_profile_from_assembly(self, on: bool = True) -> None:
        """The zero-tile profiles recorded by the model assembly (`dh_assemble`) as it writes the
        coupling blocks, instead of a scan of the blocks at the start of the next `cholesky`
        (`dh_set_tuning` key 11, off by default).  Same profiles (the same test on the same values);
        valid for the one factorization that follows the assembly, and dropped by any other write
        through the library (`cholesky(A)`, a layout change).  Writing the blocks through the tensor
        views between the assembly and the factorization is not seen: leave it off then."""
        _lib.check(self.lib.dh_set_tuning(self.h, 11, 1 if on else 0), "dh_set_tuning")

    def set_inverse_beside_solve(self, on: bool = True) -> None:
        """The first `solve` after a factorization forms the inverse of the diagonal factors on a
        second stream, in chunks of blocks ordered as its forward sweep reaches them, and the sweep
        waits for each chunk before reading it (`dh_set_tuning` key 12, off by default; not under
        graph capture, where the inverse runs first on the capturing stream).  The same kernels on
        the same blocks: bitwise the same results."""
        _lib.check(self.lib.dh_set_tuning(self.h, 12, 1 if on else 0), "dh_set_tuning")

    def set_selinv_beside_solve(self, on: bool = True) -> None:
        """A `selected_inversion` after a `solve` runs on a second stream beside the solve's backward
        sweep, each recursion step after the sweep's step that reads the blocks it overwrites
        (`dh_set_tuning` key 13, off by default; congruence form only, not under graph capture).
        Work queued on the stream between the two calls runs beside the recursion and must not read
        the factor: `logdet` and the other library calls that do cancel the arrangement (the
        recursion then follows them), reads of the block views in between are not seen."""
        _lib.check(self.lib.dh_set_tuning(self.h, 13, 1 if on else 0), "dh_set_tuning")

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
