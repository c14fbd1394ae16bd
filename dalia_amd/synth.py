"""Seeded synthetic inputs for the spatio-temporal Gaussian model.

The reference ships its example inputs as git-lfs stubs, so every workload
here is generated.  The construction follows the reference's model family
(spatio-temporal SPDE submodel + regression submodel + Gaussian likelihood,
`core/model.py:40-304`, `submodels/spatio_temporal.py:20-60`,
`submodels/regression.py:14-26`):

* spatial mesh: an m x m regular grid on the unit square, every cell split
  into two right triangles along the (0,0)-(1,1) diagonal; node index
  ``s = iy * m + ix``;
* P1 FEM: lumped mass ``c0`` (diagonal, area/3 per incident triangle),
  stiffness ``g1``, ``g2 = g1 c0^-1 g1``, ``g3 = g1 c0^-1 g2`` (the reference
  loads all four, `spatio_temporal.py:27-30`);
* temporal matrices for nt steps with dt = 1: ``m0`` the 1-D P1 mass,
  ``m1 = diag(1/2, 0, ..., 0, 1/2)`` (boundary term), ``m2`` the 1-D
  stiffness;
* observations at uniform points in [0,1]^2 x [0, nt); the spatio-temporal
  design row holds barycentric weights times linear-in-time weights (6
  entries), the regression design block holds N(0,1) covariates;
* ``y = A_st x_st + X beta + eps`` with ``x_st`` an exact draw from the
  spatio-temporal prior N(0, Qp^-1) at theta (`prior_transform`: the
  Kronecker-eigen factorisation of Qp, one eigendecomposition of the FEM
  matrices and one tridiagonal factorisation per spatial eigenvalue),
  ``beta ~ N(0,1)`` and ``eps ~ N(0, exp(-theta_obs))``.

Everything here is host-side numpy/scipy set-up data, built once per model.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np
import scipy.sparse as sps


@dataclasses.dataclass
class PriorSpec:
    """Hyper-prior of one hyperparameter (`prior_hyperparameters/*.py`)."""

    kind: str  # "penalized_complexity" or "gaussian"
    alpha: float = 0.0
    u: float = 0.0
    mean: float = 0.0
    precision: float = 0.5


@dataclasses.dataclass
class STModelData:
    """One spatio-temporal Gaussian model in DALIA's latent ordering.

    Latent vector: ``x = [x_st (time-major, index t*ns + s), beta (nb)]``,
    i.e. the spatio-temporal submodel first and the fixed effects last, as
    `core/model.py:49-55` orders them so that the fixed effects form the
    arrowhead.
    """

    ns: int
    nt: int
    nb: int
    c0: sps.csr_matrix
    g1: sps.csr_matrix
    g2: sps.csr_matrix
    g3: sps.csr_matrix
    m0: sps.csr_matrix
    m1: sps.csr_matrix
    m2: sps.csr_matrix
    a_st: sps.csr_matrix  # n_obs x (ns*nt)
    x_cov: np.ndarray  # n_obs x nb
    y: np.ndarray  # n_obs
    theta: np.ndarray  # (r_s, r_t, sigma_st, prec_o)
    priors: list  # PriorSpec per theta entry, same order
    fe_prec: float = 1e-3
    manifold: str = "plane"

    @property
    def n_obs(self) -> int:
        return int(self.y.shape[0])

    @property
    def n_latent(self) -> int:
        return self.ns * self.nt + self.nb

    def a_full(self) -> sps.csr_matrix:
        """Design matrix ``[A_st | X]`` (`core/model.py:212-233`)."""
        return sps.hstack([self.a_st, sps.csr_matrix(self.x_cov)], format="csr")


# --------------------------------------------------------------------------
# mesh and FEM
# --------------------------------------------------------------------------

def grid_triangles(m: int) -> tuple[np.ndarray, np.ndarray]:
    """Node coordinates (m*m, 2) and triangles (2*(m-1)^2, 3) of the grid."""
    h = 1.0 / (m - 1)
    ix, iy = np.meshgrid(np.arange(m), np.arange(m), indexing="xy")
    pts = np.stack([ix.ravel() * h, iy.ravel() * h], axis=1)
    cx, cy = np.meshgrid(np.arange(m - 1), np.arange(m - 1), indexing="xy")
    cx = cx.ravel()
    cy = cy.ravel()
    a = cy * m + cx
    b = a + 1
    c = a + m
    d = c + 1
    tri = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)], 0)
    return pts, tri


def p1_fem(pts: np.ndarray, tri: np.ndarray):
    """Lumped mass c0 and stiffness g1 of P1 elements (CSR, symmetric)."""
    n = pts.shape[0]
    p0, p1, p2 = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    e0 = p2 - p1  # edge opposite vertex 0
    e1 = p0 - p2
    e2 = p1 - p0
    area = 0.5 * np.abs(e2[:, 0] * (-e1[:, 1]) - e2[:, 1] * (-e1[:, 0]))
    edges = [e0, e1, e2]
    rows, cols, vals = [], [], []
    for i in range(3):
        for j in range(3):
            kij = (edges[i][:, 0] * edges[j][:, 0] + edges[i][:, 1] * edges[j][:, 1]) / (4.0 * area)
            rows.append(tri[:, i])
            cols.append(tri[:, j])
            vals.append(kij)
    g1 = sps.coo_matrix(
        (np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)
    ).tocsr()
    g1.sum_duplicates()
    lump = np.zeros(n)
    for i in range(3):
        np.add.at(lump, tri[:, i], area / 3.0)
    c0 = sps.diags(lump, format="csr")
    g1 = 0.5 * (g1 + g1.T)
    return c0.tocsr(), g1.tocsr()


def temporal_matrices(nt: int, dt: float = 1.0):
    """1-D P1 mass m0, boundary m1, stiffness m2 on nt nodes (CSR)."""
    main0 = np.full(nt, 2.0 * dt / 3.0)
    main0[0] = main0[-1] = dt / 3.0
    off0 = np.full(nt - 1, dt / 6.0)
    m0 = sps.diags([off0, main0, off0], [-1, 0, 1], format="csr")
    d1 = np.zeros(nt)
    d1[0] = d1[-1] = 0.5
    m1 = sps.diags(d1, format="csr")
    main2 = np.full(nt, 2.0 / dt)
    main2[0] = main2[-1] = 1.0 / dt
    off2 = np.full(nt - 1, -1.0 / dt)
    m2 = sps.diags([off2, main2, off2], [-1, 0, 1], format="csr")
    return m0, m1, m2


# --------------------------------------------------------------------------
# observations
# --------------------------------------------------------------------------

def design_matrix(m: int, nt: int, s_xy: np.ndarray, t: np.ndarray) -> sps.csr_matrix:
    """Barycentric-in-space times linear-in-time interpolation rows."""
    ns = m * m
    n_obs = s_xy.shape[0]
    fx = np.clip(s_xy[:, 0] * (m - 1), 0.0, m - 1 - 1e-12)
    fy = np.clip(s_xy[:, 1] * (m - 1), 0.0, m - 1 - 1e-12)
    cx = np.minimum(np.floor(fx).astype(np.int64), m - 2)
    cy = np.minimum(np.floor(fy).astype(np.int64), m - 2)
    u = fx - cx
    v = fy - cy
    a = cy * m + cx
    b = a + 1
    c = a + m
    d = c + 1
    lower = u >= v
    n1 = np.where(lower, b, c)
    w0 = np.where(lower, 1.0 - u, 1.0 - v)
    w1 = np.where(lower, u - v, v - u)
    w2 = np.where(lower, v, u)
    nodes = np.stack([a, n1, d], 1)
    ws = np.stack([w0, w1, w2], 1)
    t0 = np.minimum(np.floor(t).astype(np.int64), nt - 2)
    ft = np.clip(t - t0, 0.0, 1.0)
    wt = np.stack([1.0 - ft, ft], 1)
    tt = np.stack([t0, t0 + 1], 1)
    cols = (tt[:, :, None] * ns + nodes[:, None, :]).reshape(n_obs, 6)
    vals = (wt[:, :, None] * ws[:, None, :]).reshape(n_obs, 6)
    rows = np.repeat(np.arange(n_obs), 6)
    A = sps.csr_matrix((vals.ravel(), (rows, cols.ravel())), shape=(n_obs, ns * nt))
    A.sum_duplicates()
    return A


DEFAULT_THETA = np.array([-1.0, 0.7, 0.0, 1.5])

#: PC hyper-priors with the values of the reference's spatio-temporal example
#: (`examples/gst_small/run.py:33-35`, `:56-60`).
DEFAULT_PRIORS = [
    PriorSpec("penalized_complexity", alpha=0.01, u=0.5),
    PriorSpec("penalized_complexity", alpha=0.01, u=5.0),
    PriorSpec("penalized_complexity", alpha=0.01, u=3.0),
    PriorSpec("penalized_complexity", alpha=0.01, u=4.0),
]


def prior_transform(c0, g1, m0, m1, m2, theta, z: np.ndarray, manifold: str = "plane") -> np.ndarray:
    """x = B z with B B^T = Qp^-1 for the spatio-temporal prior at theta (time-major x).

    With c0 diagonal, g2 = g1 c0^-1 g1 and g3 = g1 c0^-1 g2, Qp = (I (x) V) blkdiag_j(T_j) (I (x) V^T)
    after ordering by spatial eigenvector, V = c0^1/2 U, U diag(lam) U^T = c0^-1/2 g1 c0^-1/2 and
    T_j = sum_k p_k(kappa^2 + lam_j) M_k tridiagonal; so x = (I (x) c0^-1/2 U) w with
    w_j = L_j^-T (U^T z)_j, T_j = L_j L_j^T.  z: (nt * ns,) standard normal."""
    from dalia_amd.hyper import kron_coefficients
    ns, nt = c0.shape[0], m0.shape[0]
    ci = 1.0 / np.sqrt(sps.csr_matrix(c0).diagonal())
    S = (sps.diags(ci) @ g1 @ sps.diags(ci)).toarray()
    lam, U = np.linalg.eigh(0.5 * (S + S.T))
    c = kron_coefficients(theta, manifold)  # rows M0, M1, M2; columns c0, g1, g2, g3
    k2 = c[2, 0] / c[2, 1]  # kappa^2
    lt = k2 + lam
    p = np.stack([c[0, 3] * lt ** 3, c[1, 2] * lt ** 2, c[2, 1] * lt])  # p_k(lt_j) (c[k, m] = p_k coefficient)
    Md = np.stack([M.diagonal() for M in (m0, m1, m2)])      # (3, nt)
    Mo = np.stack([M.diagonal(-1) for M in (m0, m1, m2)])    # (3, nt - 1)
    d = p.T @ Md   # (ns, nt): diagonals of T_j
    e = p.T @ Mo   # (ns, nt - 1): sub-diagonals
    # Cholesky of every T_j at once: l_t = sqrt(d_t - o_{t-1}^2), o_t = e_t / l_t
    L = np.zeros((ns, nt))
    O = np.zeros((ns, max(nt - 1, 0)))
    for t in range(nt):
        v = d[:, t] - (O[:, t - 1] ** 2 if t > 0 else 0.0)
        L[:, t] = np.sqrt(v)
        if t < nt - 1:
            O[:, t] = e[:, t] / L[:, t]
    # the noise is mapped into eigen coordinates (Z = U^T z) rather than drawn there: x = c0^-1/2 U
    # L(lam)^-T U^T z then depends on the eigenvalues alone, not on the basis LAPACK returns for a
    # degenerate eigenspace (a regular grid has many) or on the eigenvectors' signs
    Z = U.T @ np.asarray(z, dtype=np.float64).reshape(nt, ns).T  # (ns, nt)
    W = np.zeros_like(Z)
    for t in range(nt - 1, -1, -1):  # L_j^T w = z
        r = Z[:, t] - (O[:, t] * W[:, t + 1] if t < nt - 1 else 0.0)
        W[:, t] = r / L[:, t]
    X = (ci[:, None] * U) @ W  # (ns, nt)
    return np.ascontiguousarray(X.T).ravel()


def make_st_model(m: int, nt: int, nb: int, seed: int, obs_per_node: int = 4,
                  theta=None, priors=None) -> STModelData:
    """Build the seeded spatio-temporal model on an m x m grid."""
    rng = np.random.default_rng(seed)
    theta = np.array(DEFAULT_THETA if theta is None else theta, dtype=np.float64)
    pts, tri = grid_triangles(m)
    c0, g1 = p1_fem(pts, tri)
    c0inv = sps.diags(1.0 / c0.diagonal(), format="csr")
    g2 = (g1 @ c0inv @ g1).tocsr()
    g3 = (g1 @ c0inv @ g2).tocsr()
    g2 = (0.5 * (g2 + g2.T)).tocsr()
    g3 = (0.5 * (g3 + g3.T)).tocsr()
    m0, m1, m2 = temporal_matrices(nt)
    ns = m * m
    n_obs = obs_per_node * ns * nt
    s_xy = rng.uniform(0.0, 1.0, size=(n_obs, 2))
    t = rng.uniform(0.0, nt, size=n_obs)
    a_st = design_matrix(m, nt, s_xy, t)
    x_cov = rng.standard_normal((n_obs, nb))
    beta = rng.standard_normal(nb)
    # the spatio-temporal field: a draw from its prior at theta
    x_st = prior_transform(c0, g1, m0, m1, m2, theta, rng.standard_normal(ns * nt))
    eps = rng.standard_normal(n_obs) * math.exp(-0.5 * theta[3])
    y = a_st @ x_st + x_cov @ beta + eps
    return STModelData(
        ns=ns, nt=nt, nb=nb, c0=c0, g1=g1.tocsr(), g2=g2, g3=g3,
        m0=m0, m1=m1, m2=m2, a_st=a_st, x_cov=x_cov, y=y, theta=theta,
        priors=list(DEFAULT_PRIORS if priors is None else priors),
    )


#: The workloads, by name: (grid side m, nt, nb, seed).
WORKLOADS = {
    # CPU-runnable case: 20x20 grid (ns=400), nt=12, nb=4, seed 0
    "st_ns400_nt12_nb4": (20, 12, 4, 0),
    # plain drop-in and adjoint case: 45x45 grid (ns=2025), nt=100, nb=6, seed 1
    "st_ns2025_nt100_nb6": (45, 100, 6, 1),
    # large adjoint case: 64x64 grid (ns=4096), nt=250, nb=6, seed 1 (67 GB of BTA blocks)
    "st_ns4096_nt250_nb6": (64, 250, 6, 1),
}


def make_workload(name: str) -> STModelData:
    m, nt, nb, seed = WORKLOADS[name]
    return make_st_model(m, nt, nb, seed)


# --------------------------------------------------------------------------
# coregional (multivariate) model
# --------------------------------------------------------------------------

@dataclasses.dataclass
class CoregModelData:
    """nv spatio-temporal Gaussian models on one mesh, coupled by coregionalisation.

    Mirrors `models/coregional_model.py`: every variate i has its own
    spatio-temporal field (hyperparameters r_s_i, r_t_i; sigma_st fixed to
    ``sigma_st``, the per-model sigma being replaced by the coregional
    sigma_i), ``nb`` fixed effects and a Gaussian likelihood with precision
    prec_o_i.  theta = [r_s_0, r_t_0, prec_o_0, r_s_1, ..., sigma_0, ...,
    sigma_{nv-1}, lambda_0_1, (lambda_0_2, lambda_1_2)]
    (`coregional_model.py:101-150`, `configs/models_config.py:62-71`).
    The latent vector is permuted so that every time block holds all
    variates ([variate 0 nodes, variate 1 nodes, ...]) and the fixed effects
    of all variates form the arrowhead (`coregional_model.py:216-231`,
    `:849-915`).
    """

    nv: int
    ns: int
    nt: int
    nb: int  # fixed effects per variate
    c0: sps.csr_matrix
    g1: sps.csr_matrix
    g2: sps.csr_matrix
    g3: sps.csr_matrix
    m0: sps.csr_matrix
    m1: sps.csr_matrix
    m2: sps.csr_matrix
    a_st: list  # per variate, n_obs_i x (ns*nt)
    x_cov: list  # per variate, n_obs_i x nb
    y: list  # per variate
    theta: np.ndarray
    priors: list
    sigma_st: float = 0.0
    fe_prec: float = 1e-3

    @property
    def n_obs(self) -> int:
        return int(sum(v.shape[0] for v in self.y))

    @property
    def b(self) -> int:
        return self.nv * self.ns

    @property
    def a(self) -> int:
        return self.nv * self.nb

    @property
    def n_latent(self) -> int:
        return self.nv * (self.ns * self.nt + self.nb)

    def theta_keys(self) -> list:
        keys = []
        for i in range(self.nv):
            keys += [f"r_s_{i}", f"r_t_{i}", f"prec_o_{i}"]
        keys += [f"sigma_{i}" for i in range(self.nv)]
        keys += [f"lambda_{i}_{j}" for i in range(self.nv) for j in range(i + 1, self.nv)]
        return keys

    def latent_permutation(self) -> np.ndarray:
        """Columns of bdiag(A_i) in the permuted order (`coregional_model.py:849-915`)."""
        nst = self.ns * self.nt
        per = nst + self.nb
        blocks = [np.arange(nst).reshape(self.nt, self.ns) + i * per for i in range(self.nv)]
        st = np.hstack(blocks).ravel()
        fe = [i * per + nst + np.arange(self.nb) for i in range(self.nv)]
        return np.concatenate([st] + fe)

    def a_full(self) -> sps.csr_matrix:
        """Permuted block-diagonal design matrix bdiag([A_st_i | X_i])[:, perm]."""
        blocks = [sps.hstack([self.a_st[i], sps.csr_matrix(self.x_cov[i])], format="csr") for i in range(self.nv)]
        A = sps.block_diag(blocks, format="csr")
        return A[:, self.latent_permutation()].tocsr()

    def obs_precision_index(self) -> np.ndarray:
        """Variate of every observation (rows of a_full)."""
        return np.concatenate([np.full(self.y[i].shape[0], i) for i in range(self.nv)])


#: theta of the bivariate / trivariate workloads: per variate (r_s, r_t, prec_o),
#: then log sigmas, then lambdas.
COREG_THETA = {
    2: np.array([-1.0, 0.7, 1.5, -0.9, 0.6, 1.2, 0.0, 0.2, 0.5]),
    3: np.array([-1.0, 0.7, 1.5, -0.9, 0.6, 1.2, -1.1, 0.8, 1.0, 0.0, 0.2, -0.1, 0.5, -0.3, 0.4]),
}


def make_coreg_model(m: int, nt: int, nb: int, nv: int, seed: int, obs_per_node: int = 2,
                     theta=None) -> CoregModelData:
    """Seeded nv-variate coregional model on an m x m grid (shared mesh and time axis).

    Hyper-priors are Gaussian with precision 0.5 around the generating theta
    (lambdas around 0), as in `examples/gst_coreg2_small/run.py:80-200`.
    """
    if nv not in (2, 3):
        raise ValueError("nv must be 2 or 3")
    rng = np.random.default_rng(seed)
    theta = np.array(COREG_THETA[nv] if theta is None else theta, dtype=np.float64)
    base = make_st_model(m, nt, 1, seed, obs_per_node=1)
    ns = m * m
    a_st, x_cov, y = [], [], []
    pts, _ = grid_triangles(m)
    fields = []
    for i in range(nv):
        kx = rng.integers(1, 4, 6)
        ky = rng.integers(1, 4, 6)
        kt = rng.uniform(0.0, 6.0 * math.pi / max(nt, 2), 6)
        ph = rng.uniform(0.0, 2.0 * math.pi, 6)
        amp = rng.standard_normal(6) / math.sqrt(6)
        f = np.zeros((nt, ns))
        tn = np.arange(nt, dtype=np.float64)[:, None]
        for k in range(6):
            f += amp[k] * np.cos(math.pi * (kx[k] * pts[:, 0][None, :] + ky[k] * pts[:, 1][None, :]) + kt[k] * tn + ph[k])
        fields.append(f.ravel())
    for i in range(nv):
        n_obs = obs_per_node * ns * nt
        s_xy = rng.uniform(0.0, 1.0, size=(n_obs, 2))
        t = rng.uniform(0.0, nt, size=n_obs)
        A = design_matrix(m, nt, s_xy, t)
        X = rng.standard_normal((n_obs, nb))
        beta = rng.standard_normal(nb)
        x = fields[i] + (0.5 * fields[0] if i > 0 else 0.0)
        prec = theta[3 * i + 2]
        yi = A @ x + X @ beta + rng.standard_normal(n_obs) * math.exp(-0.5 * prec)
        a_st.append(A)
        x_cov.append(X)
        y.append(yi)
    priors = [PriorSpec("gaussian", mean=float(v), precision=0.5) for v in theta]
    n_lam = nv * (nv - 1) // 2
    for k in range(n_lam):
        priors[-1 - k] = PriorSpec("gaussian", mean=0.0, precision=0.5)
    return CoregModelData(nv=nv, ns=ns, nt=nt, nb=nb, c0=base.c0, g1=base.g1, g2=base.g2, g3=base.g3,
                          m0=base.m0, m1=base.m1, m2=base.m2, a_st=a_st, x_cov=x_cov, y=y, theta=theta,
                          priors=priors)


#: Coregional workloads: (grid side m, nt, nb per variate, nv, seed).
COREG_WORKLOADS = {
    "coreg2_ns576_nt24_nb4": (24, 24, 4, 2, 3),
    "coreg3_ns1600_nt128_nb3": (40, 128, 3, 3, 4),
}

#: Batches of bivariate instances of uneven size, (m, nt, nb, nv, seed) each.  coreg2_hess57: the
#: 3-point diagonal FD Hessian stencils (theta of 9 entries: 1 + 2 * 9 = 19 points) of three
#: instances, ns = 576, 1024, 1600 with nt = 24, 48, 96: 57 evaluations.
COREG_BATCHES = {
    "coreg2_hess57": [(24, 24, 4, 2, 3), (32, 48, 4, 2, 5), (40, 96, 4, 2, 7)],
}
