"""Per-shape choice of kernel variants, timed on the device and persisted.

Knobs (`dh_set_tuning`): key 2 = the pobtaf schedule (one persistent launch
of left-looking tile tasks, or one launch per 64-column panel), key 0 = the
Schur-update kernel of the per-panel schedule (64x64 tiles, or 128x128 tiles
with the k range split in two), key 6 = the number of workgroups of the persistent
launch (for the instances of a batch), and, for engines that leave it open
(`twisted_auto`), key 7 = the two-ended factorization, taken when it is at least
3 % faster than the one-ended one (a factorization bound by its products in bulk
gains nothing from halving its chain, and the two-ended recursion of the W form
is slower there).  `tune(engine)` times every
value of every knob on the engine's own matrix with HIP events, applies the
fastest and records it in a JSON file keyed by (padded block size, number of
blocks, arrowhead size, device name); later calls read the file.
"""

from __future__ import annotations

import json
import os

import numpy as np
import torch

from dalia_amd import _lib

DEFAULT_PATH = os.environ.get("DALIA_AMD_TUNING",
                              os.path.join(os.path.expanduser("~"), ".cache", "dalia_amd", "tuning.json"))


def _key(engine) -> str:
    lay = engine.solver.layout
    return f"v5_bp{lay.bp}_n{lay.n}_a{lay.a}_{torch.cuda.get_device_name(0)}"


def _load(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def apply(engine, choice: dict) -> None:
    lib, h = engine.lib, engine.h
    _lib.check(lib.dh_set_tuning(h, 2, int(choice["alg"])), "dh_set_tuning")
    _lib.check(lib.dh_set_tuning(h, 0, int(choice["syrk"])), "dh_set_tuning")
    _lib.check(lib.dh_set_tuning(h, 6, int(choice.get("wgs", 0))), "dh_set_tuning")
    if getattr(engine, "twisted_auto", False) and "tw" in choice:
        engine.solver.set_twisted(bool(choice["tw"]))


def _time(fn, reps: int) -> float:
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ts = []
    for _ in range(reps):
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


def tune(engine, path: str = DEFAULT_PATH, reps: int = 2, force: bool = False, workgroups: bool = False) -> dict:
    """workgroups: also choose the number of persistent workgroups (key 6) — for an instance of a
    batch, whose factorization runs beside those of the other instances."""
    cache = _load(path)
    key = _key(engine) + ("_w" if workgroups else "")
    if key in cache and not force:
        apply(engine, cache[key])
        return cache[key]
    lib, h = engine.lib, engine.h
    theta = engine.data.theta
    engine._set_params(theta)
    times = {}

    def factor():
        engine._assemble(None, "bta", True)
        _lib.check(lib.dh_cholesky(h, 1), "dh_cholesky")

    cands = {"dag": (1, 0), "syrk0": (0, 0), "syrk1": (0, 1)}
    for name, (alg, v) in cands.items():
        _lib.check(lib.dh_set_tuning(h, 2, alg), "dh_set_tuning")
        _lib.check(lib.dh_set_tuning(h, 0, v), "dh_set_tuning")
        factor()
        times[name] = _time(factor, reps)
    best = min(cands, key=lambda k: times[k])
    best_alg, best_syrk = cands[best]
    _lib.check(lib.dh_set_tuning(h, 2, best_alg), "dh_set_tuning")
    _lib.check(lib.dh_set_tuning(h, 0, best_syrk), "dh_set_tuning")
    twist = None
    if getattr(engine, "twisted_auto", False):
        twist = 0
        if best_alg == 1 and engine.data.nt >= 3:
            engine.solver.set_twisted(True)
            factor()
            times["dag_twisted"] = _time(factor, reps)
            twist = 1 if times["dag_twisted"] <= 0.97 * times[best] else 0
        engine.solver.set_twisted(bool(twist))
    wgs = 0
    if best_alg == 1 and workgroups:
        # the fewest persistent workgroups (of the default, a half, a quarter, an eighth) that keep the
        # factorization within 3 % of its fastest: a launch bound by its chain of diagonal tiles then
        # leaves CUs to the other instances of a batch
        full = torch.cuda.get_device_properties(0).multi_processor_count
        tw = {0: times["dag"]}
        for cap in (full // 2, full // 4, full // 8):
            _lib.check(lib.dh_set_tuning(h, 6, cap), "dh_set_tuning")
            factor()
            tw[cap] = _time(factor, reps)
        fastest = min(tw.values())
        ok = [c for c, t in tw.items() if c > 0 and t <= 1.03 * fastest]
        wgs = min(ok) if ok else 0
        times["wgs"] = {str(k): v for k, v in tw.items()}
        _lib.check(lib.dh_set_tuning(h, 6, wgs), "dh_set_tuning")

    choice = {"alg": best_alg, "syrk": best_syrk, "wgs": wgs, "ms": times}
    if twist is not None:
        choice["tw"] = twist
    apply(engine, choice)
    cache[key] = choice
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as fh:
            json.dump(cache, fh, indent=1)
    except OSError:
        pass
    return choice
