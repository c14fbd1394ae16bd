"""Write the golden fixtures under tests/golden/ from the CPU oracle.

The reference ships no usable golden vectors (its example outputs are git-lfs
stubs and pin only an optimiser end point), so these fixtures come from the
oracle in `oracle/`, itself pinned by dense LAPACK, the Gaussian marginal
likelihood identity and central differences (tests/test_oracle_*.py).

Run: python tests/golden/make_golden.py          (small fixtures)
     python tests/golden/make_golden.py full     (the ns=2025, nt=100, nb=6 fixture; ~30 GB of
                                                 host memory and a few minutes)
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from dalia_amd import synth  # noqa: E402
from oracle import bta, inla  # noqa: E402


def model_fixture(name, m, nt, nb, seed, thetas, full_blocks=False):
    d = synth.make_st_model(m, nt, nb, seed)
    out = {"m": m, "nt": nt, "nb": nb, "seed": seed, "thetas": np.array(thetas),
           "y_head": d.y[:16].copy(), "y_sum": d.y.sum()}
    fs, gs, gfds, mus, ldc, ldp, quads, mv = [], [], [], [], [], [], [], []
    for th in thetas:
        th = np.asarray(th, dtype=np.float64)
        f, parts = inla.f_theta(d, th, return_parts=True)
        f2, g = inla.adjoint_gradient(d, th)
        _, gfd = inla.fd_gradient(d, th, 1e-4)
        fs.append(f); gs.append(g); gfds.append(gfd); mus.append(parts["mu"])
        ldc.append(parts["logdet_qc"]); ldp.append(parts["logdet_qp"]); quads.append(parts["quad"])
        Qc = inla.qcond_sparse(d, th)
        D, L, R, T = inla.sparse_to_bta(Qc, d.nt, d.ns, d.nb, "bta")
        bta.pobtaf(D, L, R, T)
        if full_blocks and len(fs) == 1:
            out.update(L_diag=D.copy(), L_lower=L.copy(), L_arrow=R.copy(), L_tip=T.copy())
        bta.pobtasi(D, L, R, T)
        if full_blocks and len(fs) == 1:
            out.update(X_diag=D.copy(), X_lower=L.copy(), X_arrow=R.copy(), X_tip=T.copy())
        mv.append(np.concatenate([np.concatenate([np.diagonal(D[i]) for i in range(d.nt)]), np.diagonal(T)]))
    out.update(f=np.array(fs), grad=np.array(gs), grad_fd=np.array(gfds), mu=np.array(mus),
               logdet_qc=np.array(ldc), logdet_qp=np.array(ldp), quad=np.array(quads),
               marginal_var=np.array(mv))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "f =", fs)


#: blocks whose selected-inverse diagonals the full-size fixture stores
FULL_BLOCKS = (0, 1, 37, 50, 98, 99)
#: stride of the subsample of mu stored by the full-size fixture
MU_STRIDE = 61


def full_size_fixture(name, m, nt, nb, seed):
    """A fixture at a full benchmark shape: scalars, the adjoint gradient, a strided sample and
    checksums of mu, and the diagonals of a few selected-inverse blocks (the arrays themselves are
    too large to commit)."""
    import time
    t0 = time.perf_counter()
    d = synth.make_st_model(m, nt, nb, seed)
    th = np.asarray(d.theta, dtype=np.float64)
    f, parts = inla.f_theta(d, th, return_parts=True)
    print(name, "f", f, f"{time.perf_counter() - t0:.0f} s", flush=True)
    f2, g = inla.adjoint_gradient(d, th)
    print(name, "grad", g, f"{time.perf_counter() - t0:.0f} s", flush=True)
    D, L, R, T = inla.selected_inverse_blocks(d, th)
    mu = parts["mu"]
    out = {"m": m, "nt": nt, "nb": nb, "seed": seed, "theta": th,
           "y_head": d.y[:16].copy(), "y_sum": d.y.sum(),
           "f": f, "grad": g, "logdet_qc": parts["logdet_qc"], "logdet_qp": parts["logdet_qp"],
           "quad": parts["quad"], "mu_stride": MU_STRIDE, "mu_sub": mu[::MU_STRIDE].copy(),
           "mu_sum": mu.sum(), "mu_norm": np.linalg.norm(mu),
           "x_blocks": np.array(FULL_BLOCKS),
           "x_diag": np.stack([np.diagonal(D[i]).copy() for i in FULL_BLOCKS]),
           "x_lower_diag": np.stack([np.diagonal(L[i]).copy() for i in FULL_BLOCKS if i < nt - 1]),
           "x_tip": T.copy(), "x_arrow_sum": R.sum(axis=(1, 2))}
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "written", f"{time.perf_counter() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "full":
        full_size_fixture("st_ns2025_nt100_nb6", 45, 100, 6, 1)
        sys.exit(0)
    th0 = synth.DEFAULT_THETA
    model_fixture("st_tiny_m5_nt4_nb2", 5, 4, 2, 0, [th0, th0 + np.array([0.2, -0.1, 0.3, -0.4])],
                  full_blocks=True)
    model_fixture("st_ns400_nt12_nb4", 20, 12, 4, 0, [th0])
