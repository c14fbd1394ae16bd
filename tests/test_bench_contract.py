"""bench.py's command-line contract (no GPU): more than one GPU is refused with a message."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr
    assert r.stdout == ""


def test_world_size_above_one_is_refused():
    env = dict(os.environ, WORLD_SIZE="2")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr


def _brute_pobtaf_flops(n, b, a, prof):
    """Tile-by-tile count of the factorization's flops with the tiles (R, K < prof[s][R]) of each
    sub-diagonal block left out: loops over every tile product, written independently of the
    vectorised form in bench.py."""
    import numpy as np
    nbt = (b + 63) // 64
    w = [min(64, b - 64 * t) for t in range(nbt)]
    F = 0.0
    for i in range(n):
        F += b ** 3 / 3 + a * b * b + a * a * b  # diagonal block, its arrow TRSM, the tip update
    for s in range(n - 1):
        f = prof[s]
        for R in range(nbt):
            for c in range(nbt):
                if c < f[R]:
                    continue  # a zero tile of the coupling: no work at all
                for k in range(f[R], c):
                    F += 2 * w[R] * w[c] * w[k]
                F += w[R] * w[c] * w[c]
        for r in range(nbt):
            for c in range(r + 1):
                for K in range(max(f[r], f[c]), nbt):
                    F += (2 if r > c else 1) * w[r] * w[c] * w[K]
        for c in range(nbt):
            for K in range(f[c], nbt):
                F += 2 * a * w[c] * w[K]
    return F + a ** 3 / 3


def test_profiled_flop_count():
    import numpy as np
    sys.path.insert(0, ROOT)
    import bench
    for n, b, a in [(3, 200, 4), (4, 2025, 6), (2, 64, 0)]:
        nbt = (b + 63) // 64
        zero = np.zeros((n - 1, nbt), dtype=int)
        assert abs(bench.f_pobtaf_profile(n, b, a, zero) - bench.f_pobtaf(n, b, a)) <= 1e-9 * bench.f_pobtaf(n, b, a)
        rng = np.random.default_rng(n + b)
        for prof in (np.maximum(0, np.arange(nbt) - 3)[None, :].repeat(n - 1, 0), rng.integers(0, nbt + 1, (n - 1, nbt))):
            ref = _brute_pobtaf_flops(n, b, a, prof)
            assert abs(bench.f_pobtaf_profile(n, b, a, prof) - ref) <= 1e-9 * ref
