"""Read (and write) model inputs in DALIA's on-disk layout.

The reference loads each submodel from its own input directory
(`core/submodel.py:25-40`: `a.npz`, optional `x.npy`;
`submodels/spatio_temporal.py:27-35`: `c0, g1, g2, g3, m0, m1, m2` as
`scipy.sparse.save_npz` files) and the observations from `y.npy` in the
model directory, the parent of the submodel directories
(`core/model.py:239-252`).  The example trees use
``<model>/inputs_spatio_temporal/`` and ``<model>/inputs_regression/``, and
``model_1/``, ``model_2/``, ... for the variates of a coregional model.

These functions build the set-up data of this package (`STModelData`,
`CoregModelData`) from such a tree; everything on the device is then built
from them as for the synthetic workloads.
"""

from __future__ import annotations

import os

import numpy as np
import scipy.sparse as sps

from dalia_amd import synth

_ST_FILES = ("c0", "g1", "g2", "g3", "m0", "m1", "m2")


def _npz(path: str) -> sps.csr_matrix:
    return sps.csr_matrix(sps.load_npz(path))


def _read_variate(model_dir: str, st_dir: str, reg_dir: str):
    st = os.path.join(model_dir, st_dir)
    mats = {k: _npz(os.path.join(st, f"{k}.npz")) for k in _ST_FILES}
    a_st = _npz(os.path.join(st, "a.npz"))
    reg = os.path.join(model_dir, reg_dir)
    if os.path.isdir(reg):
        x_cov = np.asarray(_npz(os.path.join(reg, "a.npz")).toarray(), dtype=np.float64)
    else:
        x_cov = np.zeros((a_st.shape[0], 0))
    y = np.load(os.path.join(model_dir, "y.npy"), allow_pickle=False).astype(np.float64).ravel()
    ns, nt = mats["c0"].shape[0], mats["m0"].shape[0]
    if a_st.shape != (y.size, ns * nt):
        raise ValueError(f"{st}/a.npz has shape {a_st.shape}, expected ({y.size}, {ns * nt})")
    if x_cov.shape[0] != y.size:
        raise ValueError(f"{reg}/a.npz has {x_cov.shape[0]} rows, expected {y.size}")
    return mats, a_st, x_cov, y


def load_st_model(model_dir: str, theta, priors=None, fe_prec: float = 1e-3,
                  st_dir: str = "inputs_spatio_temporal", reg_dir: str = "inputs_regression") -> synth.STModelData:
    """Spatio-temporal Gaussian model (spatio-temporal submodel + regression) from disk.

    theta: (r_s, r_t, sigma_st, prec_o), the reference's order; priors: one
    `synth.PriorSpec` per entry (default: the PC priors of the workloads).
    """
    mats, a_st, x_cov, y = _read_variate(model_dir, st_dir, reg_dir)
    return synth.STModelData(ns=mats["c0"].shape[0], nt=mats["m0"].shape[0], nb=x_cov.shape[1], a_st=a_st,
                             x_cov=x_cov, y=y, theta=np.asarray(theta, dtype=np.float64),
                             priors=list(synth.DEFAULT_PRIORS if priors is None else priors), fe_prec=fe_prec,
                             **mats)


def load_coreg_model(root: str, nv: int, theta, priors, sigma_st: float = 0.0, fe_prec: float = 1e-3,
                     variate_dirs=None, st_dir: str = "inputs_spatio_temporal",
                     reg_dir: str = "inputs_regression") -> synth.CoregModelData:
    """Coregional model from ``root/model_1 .. model_nv`` (one model tree per variate).

    The variates must share the mesh and the time axis (as the reference
    checks, `coregional_model.py:31-99`); the FEM and temporal matrices of the
    first variate are used and the others are checked against them.
    """
    dirs = variate_dirs or [os.path.join(root, f"model_{i + 1}") for i in range(nv)]
    if len(dirs) != nv:
        raise ValueError("one directory per variate is required")
    parts = [_read_variate(d, st_dir, reg_dir) for d in dirs]
    mats0 = parts[0][0]
    for k, (mats, _, x_cov, _) in enumerate(parts):
        for name in _ST_FILES:
            if mats[name].shape != mats0[name].shape or abs(mats[name] - mats0[name]).max() > 1e-12 * max(
                    1.0, abs(mats0[name]).max()):
                raise ValueError(f"variate {k}: {name} differs from variate 0 (the variates must share the mesh)")
        if x_cov.shape[1] != parts[0][2].shape[1]:
            raise ValueError("every variate needs the same number of fixed effects")
    return synth.CoregModelData(nv=nv, ns=mats0["c0"].shape[0], nt=mats0["m0"].shape[0], nb=parts[0][2].shape[1],
                                a_st=[p[1] for p in parts], x_cov=[p[2] for p in parts], y=[p[3] for p in parts],
                                theta=np.asarray(theta, dtype=np.float64), priors=list(priors), sigma_st=sigma_st,
                                fe_prec=fe_prec, **mats0)


def _write_variate(model_dir: str, mats: dict, a_st, x_cov, y, st_dir: str, reg_dir: str) -> None:
    st = os.path.join(model_dir, st_dir)
    os.makedirs(st, exist_ok=True)
    for k, m in mats.items():
        sps.save_npz(os.path.join(st, f"{k}.npz"), sps.csc_matrix(m))
    sps.save_npz(os.path.join(st, "a.npz"), sps.csc_matrix(a_st))
    if x_cov.shape[1] > 0:
        reg = os.path.join(model_dir, reg_dir)
        os.makedirs(reg, exist_ok=True)
        sps.save_npz(os.path.join(reg, "a.npz"), sps.csc_matrix(x_cov))
    np.save(os.path.join(model_dir, "y.npy"), np.asarray(y))


def save_st_model(data: synth.STModelData, model_dir: str, st_dir: str = "inputs_spatio_temporal",
                  reg_dir: str = "inputs_regression") -> None:
    """Write a model in the same layout (e.g. a synthetic workload for the reference)."""
    mats = {k: getattr(data, k) for k in _ST_FILES}
    _write_variate(model_dir, mats, data.a_st, data.x_cov, data.y, st_dir, reg_dir)


def save_coreg_model(data: synth.CoregModelData, root: str, st_dir: str = "inputs_spatio_temporal",
                     reg_dir: str = "inputs_regression") -> None:
    mats = {k: getattr(data, k) for k in _ST_FILES}
    for i in range(data.nv):
        _write_variate(os.path.join(root, f"model_{i + 1}"), mats, data.a_st[i], data.x_cov[i], data.y[i], st_dir,
                       reg_dir)
