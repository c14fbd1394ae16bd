"""dh_logdet_noalloc, the entry point that allocates nothing, against dh_logdet through the C ABI.

Two calls in a row on one GPU share the library's fixed scratch array; the second has fewer
blocks, so stale partial sums of the first would show.  Bar: the log-determinant of the handle's
own entry point to 1e-10 relative (both run the same kernels in the same order, so in practice
the values are identical), and to dense LAPACK.
"""

import ctypes

import numpy as np
import pytest
import torch

from dalia_amd import _lib
from tests.bta_cases import random_bta

pytestmark = pytest.mark.gpu


def _factor(n, b, a, sparsity, seed):
    from dalia_amd.solver import HipBTASolver
    M, S = random_bta(n, b, a, seed=seed)
    s = HipBTASolver(diagonal_blocksize=b, n_diag_blocks=n, arrowhead_blocksize=a)
    s.cholesky(S, sparsity)
    assert s.info() == 0
    return s, M


def _noalloc(lib, s, sparsity, out):
    lay = s.layout
    a = lay.a if sparsity == "bta" else 0
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return lib.dh_logdet_noalloc(ctypes.c_void_p(lay.diag), lay.n, lay.b, lay.bp, lay.bp * lay.bp,
                                 ctypes.c_void_p(lay.tip if a else None), a, lay.ap, out.data_ptr(), stream)


@pytest.mark.parametrize("sparsity", ["bta", "bt"])
def test_noalloc_logdet_matches_handle_logdet_twice(sparsity):
    lib = _lib.load()
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    for n, b, a, seed in [(9, 300, 5, 1), (3, 130, 2, 2)]:  # the second call is smaller
        print(f"case noalloc {sparsity} n={n} b={b} a={a}", flush=True)
        s, M = _factor(n, b, a, sparsity, seed)
        ref = s.logdet(sparsity)
        assert _noalloc(lib, s, sparsity, out) == 0
        got = float(out.item())
        assert abs(got - ref) <= 1e-10 * abs(ref), (got, ref)
        Md = M if sparsity == "bta" else M[: n * b, : n * b]
        assert abs(got - np.linalg.slogdet(Md)[1]) <= 1e-10 * abs(ref)


def test_noalloc_rejects_bad_sizes_without_launching():
    lib = _lib.load()
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    dummy = torch.zeros(64 * 64, dtype=torch.float64, device="cuda")
    p = ctypes.c_void_p(dummy.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # more blocks than the fixed scratch array holds
    assert lib.dh_logdet_noalloc(p, 65537, 1, 1, 1, None, 0, 0, out.data_ptr(), st) == -1
    # a leading dimension smaller than the block, a stride that overlaps blocks, a missing tip
    assert lib.dh_logdet_noalloc(p, 1, 64, 32, 64 * 64, None, 0, 0, out.data_ptr(), st) == -1
    assert lib.dh_logdet_noalloc(p, 2, 64, 64, 64, None, 0, 0, out.data_ptr(), st) == -1
    assert lib.dh_logdet_noalloc(p, 1, 8, 64, 64 * 64, None, 2, 64, out.data_ptr(), st) == -1
    torch.cuda.synchronize()
