"""GPU: the checkpointed engine (factor kept as segment boundaries and recomputed in the backward
pass) against the engine that stores the whole factor, and against the oracle.

Segment lengths cover one segment, segments of 2, 3 and 5 blocks, a last segment of one block,
and a block size with 15 tile columns (every task kind of the persistent factorization inside a
segment).  Bar: f to 1e-12 relative of the stored-factor engine (the same kernels on the same
data; only sums over segments are ordered differently), the gradient, mu and the marginal
variances to 1e-10; f against the oracle to 1e-10.
"""

import numpy as np
import pytest

from dalia_amd import _lib, synth
from oracle import inla as oinla

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def model17():
    return synth.make_st_model(8, 17, 3, 6)


@pytest.fixture(scope="module")
def reference17(model17):
    from dalia_amd.inla import STInla
    d = model17
    eng = STInla(d)
    th = d.theta + np.array([0.1, -0.1, 0.05, 0.2])
    f = eng.f(th)
    mu = eng.mu.cpu().numpy()
    fg, g = eng.value_and_grad(th)
    mv = eng.marginal_variances_latent(th, np.zeros(d.n_latent))
    return th, f, mu, fg, g, mv


@pytest.mark.parametrize("pattern", [True, False])
@pytest.mark.parametrize("K", [2, 3, 5, 16, 17])
def test_checkpointed_matches_stored_factor(model17, reference17, K, pattern):
    from dalia_amd.ckpt import STInlaCkpt
    d = model17
    th, f_ref, mu_ref, fg_ref, g_ref, mv_ref = reference17
    print(f"case ckpt n={d.nt} ns={d.ns} K={K} pattern={pattern}", flush=True)
    eng = STInlaCkpt(d, K=K, pattern_selinv=pattern)
    f = eng.f(th)
    assert abs(f - f_ref) <= 1e-12 * abs(f_ref), (f, f_ref)
    assert rel(eng.mu.cpu().numpy(), mu_ref) < 1e-10
    fg, g = eng.value_and_grad(th)
    assert abs(fg - fg_ref) <= 1e-12 * abs(fg_ref)
    assert rel(g, g_ref) < 1e-10, (g, g_ref)
    mv = eng.marginal_variances_latent(th, np.zeros(d.n_latent))
    assert rel(mv, mv_ref) < 1e-10
    # the engine is reusable, and agrees with the oracle
    f_or = oinla.f_theta(d, d.theta)
    assert abs(eng.f(d.theta) - f_or) <= 1e-10 * abs(f_or)
    assert eng.memory_bytes() > 0


@pytest.mark.parametrize("pattern", [True, False])
def test_checkpointed_large_tiles(pattern):
    """ns = 900 (15 tile columns per block, a banded profile), nt = 7 in segments of 3: f, the
    gradient and the latent marginal variances, with the segments' recursions in congruence form on
    the coupling tiles or in the W form, against the stored-factor engine with whole blocks."""
    from dalia_amd.ckpt import STInlaCkpt
    from dalia_amd.inla import STInla
    d = synth.make_st_model(30, 7, 4, 8)
    print(f"case ckpt large tiles ns={d.ns} nt={d.nt} pattern={pattern}", flush=True)
    ref = STInla(d, pattern_selinv=False)
    f_ref = ref.f(d.theta)
    _, g_ref = ref.value_and_grad(d.theta)
    mv_ref = ref.marginal_variances_latent(d.theta, np.zeros(d.n_latent))
    eng = STInlaCkpt(d, K=3, pattern_selinv=pattern)
    assert abs(eng.f(d.theta) - f_ref) <= 1e-12 * abs(f_ref)
    _, g = eng.value_and_grad(d.theta)
    assert rel(g, g_ref) < 1e-10
    assert rel(eng.marginal_variances_latent(d.theta, np.zeros(d.n_latent)), mv_ref) < 1e-10


def test_memory_budget_picks_the_engine(model17):
    from dalia_amd.ckpt import STInlaCkpt, ckpt_bytes, full_handle_bytes, make_engine
    from dalia_amd.inla import STInla
    d = model17
    assert type(make_engine(d)) is STInla
    full = full_handle_bytes(d.nt, d.ns, d.nb)
    eng = make_engine(d, memory_budget=0.7 * full)
    assert isinstance(eng, STInlaCkpt) and ckpt_bytes(d.nt, d.ns, d.nb, eng.K) <= 0.7 * full
    print(f"budget {0.7 * full / 1e6:.1f} MB -> K = {eng.K}, {eng.memory_bytes() / 1e6:.1f} MB", flush=True)
    one = make_engine(d, memory_budget=full)  # the one-ended handle fits, the two-ended one's copies do not
    assert type(one) is STInla and not one.twisted
    assert full_handle_bytes(d.nt, d.ns, d.nb, twisted=True) > full
    forced = make_engine(d, checkpoint=2)
    assert isinstance(forced, STInlaCkpt) and forced.K == 2
    f_ref = oinla.f_theta(d, d.theta)
    assert abs(forced.f(d.theta) - f_ref) <= 1e-10 * abs(f_ref)
    with pytest.raises(MemoryError):
        make_engine(d, memory_budget=1e3)


def test_checkpointed_giveup_raises(model17):
    """A give-up of a segment's persistent factorization (forced from the host: every wait returns
    at once, every kernel exits normally) raises DhError; the next evaluation is correct again."""
    from dalia_amd.ckpt import STInlaCkpt
    lib = _lib.load()
    eng = STInlaCkpt(model17, K=4)
    _lib.check(lib.dh_ckpt_set_tuning(eng.ck, 5, 1), "dh_ckpt_set_tuning")
    with pytest.raises(_lib.DhError):
        eng.f(model17.theta)
    with pytest.raises(_lib.DhError):
        eng.value_and_grad(model17.theta)
    _lib.check(lib.dh_ckpt_set_tuning(eng.ck, 5, 0), "dh_ckpt_set_tuning")
    assert lib.dh_ckpt_set_tuning(eng.ck, 4, 1) == -1
    f_ref = oinla.f_theta(model17, model17.theta)
    assert abs(eng.f(model17.theta) - f_ref) <= 1e-10 * abs(f_ref)
