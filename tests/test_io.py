"""CPU: DALIA's on-disk input layout round-trips into the same model (no GPU)."""

import numpy as np
import pytest

from dalia_amd import io, synth
from oracle import coreg as ocoreg
from oracle import inla as oinla


def test_st_model_round_trip(tmp_path):
    d = synth.make_st_model(5, 4, 2, 3)
    io.save_st_model(d, str(tmp_path))
    for f in ("c0", "g1", "g2", "g3", "m0", "m1", "m2", "a"):
        assert (tmp_path / "inputs_spatio_temporal" / f"{f}.npz").exists()
    assert (tmp_path / "inputs_regression" / "a.npz").exists() and (tmp_path / "y.npy").exists()
    e = io.load_st_model(str(tmp_path), d.theta, d.priors)
    assert (e.ns, e.nt, e.nb, e.n_obs) == (d.ns, d.nt, d.nb, d.n_obs)
    assert abs(e.a_st - d.a_st).max() == 0 and np.array_equal(e.x_cov, d.x_cov) and np.array_equal(e.y, d.y)
    assert oinla.f_theta(e, d.theta) == oinla.f_theta(d, d.theta)


def test_coreg_model_round_trip_and_mesh_check(tmp_path):
    d = synth.make_coreg_model(4, 3, 2, 2, 1)
    io.save_coreg_model(d, str(tmp_path))
    e = io.load_coreg_model(str(tmp_path), 2, d.theta, d.priors)
    assert ocoreg.f_theta_coreg(e, d.theta) == ocoreg.f_theta_coreg(d, d.theta)
    # a variate on another mesh is refused
    import scipy.sparse as sps
    bad = 2.0 * sps.load_npz(str(tmp_path / "model_2" / "inputs_spatio_temporal" / "g1.npz"))
    sps.save_npz(str(tmp_path / "model_2" / "inputs_spatio_temporal" / "g1.npz"), bad)
    with pytest.raises(ValueError):
        io.load_coreg_model(str(tmp_path), 2, d.theta, d.priors)


def test_shape_mismatch_is_an_error(tmp_path):
    d = synth.make_st_model(4, 3, 1, 0)
    io.save_st_model(d, str(tmp_path))
    np.save(str(tmp_path / "y.npy"), d.y[:-1])
    with pytest.raises(ValueError):
        io.load_st_model(str(tmp_path), d.theta)
