"""Benchmark: adjoint gradient evaluations of f(theta) per second on one MI355X.

Workload (default): the spatio-temporal Gaussian model on a 45x45 grid
(ns = 2025), nt = 100, nb = 6, 4*ns*nt observations, seed 1, FP64.  One
"step" is one adjoint gradient evaluation: assembly of Qc, pobtaf, pobtas,
logdet, quadratic form, Kronecker-eigen prior term, pobtasi and the
selected-inverse contractions.  Inputs are resident in HBM before timing.

Also reported on the same line: the plain GPU central-difference gradient
(2d+1 = 9 plain f(theta) evaluations, each with a BTA and a BT
factorisation), pobtaf TFLOP/s, the roofline of the dominant kernel (HIP
events on the library's stream around its launches inside the timed steps)
and a CPU baseline (the oracle on the host
cores, on a bounded sample, scaled to the full workload).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

def metric_name(ns: int, nt: int, nb: int) -> str:
    return f"∇f(θ) evals/s adjoint vs central-diff, ns={ns} nt={nt} nb={nb} FP64; pobtaf TFLOP/s"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def f_pobtaf(n, b, a):
    return n * b ** 3 / 3 + 2 * (n - 1) * b ** 3 + n * a * b ** 2 + 2 * (n - 1) * a * b ** 2 + n * a * a * b + a ** 3 / 3


def solve_launches(nt: int, nb: int, twisted: bool) -> int:
    """Kernel launches of one dh_solve: per block step the row GEMV with the coupling (or a copy),
    the column partials and their sum with the packed inverse, and the arrow partials (forward);
    the column partials with the coupling and arrow rows, their sum and the packed-inverse row GEMV
    (backward); plus packing, the arrow sums and the tip.  Two-ended: the two sides' steps share
    their launches, and the junction block runs alone."""
    a = 1 if nb else 0
    if not twisted:
        return 2 + nt * (3 + a) + a + 2 * a + 3 * nt - (0 if a else 1)
    m = (nt - 1) // 2
    steps = max(m, nt - 1 - m)
    fwd = steps * (3 + a) + 2 * a + (4 + a) + a
    bwd = (3 if a else 2) + 3 * steps
    return 2 + fwd + 2 * a + bwd


def tile_widths(b):
    return np.array([min(64, b - 64 * t) for t in range((b + 63) // 64)], dtype=float)


def f_pobtaf_profile(n, b, a, prof):
    """Algorithmic flops of pobtaf when tile (R, K) of sub-diagonal block s is zero for K < prof[s][R]
    (the zero-tile profile the factorization used; those tiles stay zero in the factor, and no
    product with them is counted).  Equals f_pobtaf for an all-zero profile."""
    w = tile_widths(b)
    nbt = len(w)
    suffix = np.concatenate([np.cumsum(w[::-1])[::-1], [0.0]])  # suffix[k] = sum_{K >= k} w[K]
    F = n * b ** 3 / 3 + n * a * b * b + n * a * a * b + a ** 3 / 3
    for s in range(n - 1):
        f = np.minimum(np.asarray(prof[s], dtype=int), nbt)
        for R in range(nbt):  # the coupling's TRSM: tiles (R, c >= f[R]), GEMM terms k in [f[R], c)
            fr = f[R]
            if fr >= nbt:
                continue
            cs = np.arange(fr, nbt)
            pre = np.concatenate([[0.0], np.cumsum(w[fr:])])
            F += float(np.sum(2 * w[R] * w[cs] * pre[cs - fr] + w[R] * w[cs] ** 2))
        K = suffix[np.maximum.outer(f, f)]  # the next diagonal block's update: K >= max(f[r], f[c])
        WW = np.outer(w, w)
        F += float(np.sum(np.tril(2 * WW * K, -1)) + np.sum(np.diag(WW) * np.diag(K)))
        F += float(2 * a * np.sum(w * suffix[f]))  # the arrow row against the coupling
    return F


def structural_profile(d):
    """The zero-tile profile of the sub-diagonal blocks of a spatio-temporal model's Qc from its
    structure (the union of the FEM patterns and the coupling A_st^T A_st makes between consecutive
    time steps): first nonzero 64-column tile per 64-row tile row, as the library computes it."""
    import scipy.sparse as sps
    ns = d.ns
    nbt = (ns + 63) // 64
    pat = (abs(d.c0) + abs(d.g1) + abs(d.g2) + abs(d.g3)).tocsr()
    nt = d.nt
    a = d.a_st.tocsc()
    cols = a[:, ns:2 * ns] if nt > 1 else None
    off = (abs(a[:, 0:ns]).T @ abs(cols)).T if cols is not None else sps.csr_matrix((ns, ns))
    P = (pat + off).tocoo()
    f = np.full(nbt, nbt)
    np.minimum.at(f, P.row // 64, P.col // 64)
    return f


def tile_profile(eng):
    """The zero-tile profiles of the engine's last factorization, (n - 1) x nbt."""
    from dalia_amd import _lib
    lay = eng.solver.layout
    nbt = lay.bp // 64
    buf = np.zeros((lay.n - 1) * nbt, dtype=np.int32)
    _lib.check(eng.lib.dh_get_tile_profile(eng.h, buf.ctypes.data, buf.size), "dh_get_tile_profile")
    return buf.reshape(lay.n - 1, nbt)


def f_pobtasi(n, b, a):
    return (n - 1) * (13.0 / 3.0) * b ** 3 + (2.0 / 3.0) * b ** 3 + (n - 1) * (6 * a * b * b + 4 * a * a * b)


def syrk_flops_per_launch(n, b, a):
    """Algorithmic flops of the n launches of k_syrk (SYRK counted once)."""
    per = [(b ** 3 + 2 * a * b * b + a * a * b) if i < n - 1 else a * a * b for i in range(n)]
    return float(np.mean(per))


def update_flops_per_launch(n, b, a):
    """Algorithmic flops of the k_update launches (right-looking rank-64 updates)."""
    tot, cnt = 0.0, 0
    nbt = (b + 63) // 64
    for i in range(n):
        lower = b if i < n - 1 else 0
        for p in range(nbt - 1):
            kp = min(64, b - 64 * p)
            m = b - 64 * (p + 1)
            if m <= 0:
                continue
            tot += 2.0 * kp * (m * (m + 1) / 2 + lower * m + a * m)
            cnt += 1
    return tot / max(cnt, 1), cnt


def dgemm_ceiling(n: int = 8192, reps: int = 5) -> float:
    """A measured FP64 GEMM ceiling on this GPU: the vendor library's DGEMM (torch.mm) at n^3, TFLOP/s."""
    import torch

    a = torch.randn(n, n, dtype=torch.float64, device="cuda")
    b = torch.randn(n, n, dtype=torch.float64, device="cuda")
    c = torch.mm(a, b)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        torch.mm(a, b, out=c)
    ev1.record()
    torch.cuda.synchronize()
    t = ev0.elapsed_time(ev1) * 1e-3 / reps
    del a, b, c
    torch.cuda.empty_cache()
    return 2.0 * n ** 3 / t / 1e12


def committed_traffic(kname: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary of this kernel and workload
    (FETCH_SIZE and WRITE_SIZE, each scaled by a copy-kernel calibration), or None."""
    tag = {"st_ns2025_nt100_nb6": "ns2025_nt100"}.get(workload)
    path = os.path.join(ROOT, "profiles", f"{kname}_{tag}_traffic.md")
    if tag is None or not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _median_time(fn, reps: int) -> float:
    fn()  # warm-up
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline(workload_args, n_full: int, threads: int, reps: int = 5, nt_s: int = 4):
    """The oracle (numpy/scipy restatement of the reference's CPU path) on the host cores.

    Measured, each a median of `reps` runs after one warm-up:
    * one f(theta) (BTA + BT factorisations, solve, logdets) on a bounded sample of the workload:
      the same ns and nb, nt_s = 4 time steps; scaled linearly in nt to the full workload (every
      cost of f is linear in nt at fixed ns) and by 9 evaluations per central-difference
      gradient (the reference's method) -> `value`;
    * pobtaf alone on the same sample (GFLOP/s);
    * f(theta) plus the selected inversion at the CPU-runnable ns=400, nt=12, nb=4 case, in full;
    * f on the sample with one BLAS thread (a single run after the warm-up).
    """
    from threadpoolctl import threadpool_info, threadpool_limits

    from dalia_amd import synth
    from oracle import bta as obta
    from oracle import inla as oinla

    m, nt, nb, seed = workload_args
    d = synth.make_st_model(m, nt_s, nb, seed)
    d0 = synth.make_st_model(*synth.WORKLOADS["st_ns400_nt12_nb4"])
    Qc = oinla.qcond_sparse(d, d.theta)

    def factor():
        D, L, R, T = oinla.sparse_to_bta(Qc, d.nt, d.ns, d.nb, "bta")
        t0 = time.perf_counter()
        obta.pobtaf(D, L, R, T)
        return time.perf_counter() - t0

    with threadpool_limits(limits=threads):
        blas = sorted({f"{i.get('internal_api')} {i.get('version')}" for i in threadpool_info()})
        t_f = _median_time(lambda: oinla.f_theta(d, d.theta), reps)
        log(f"CPU baseline: f on the sample {t_f:.2f} s")
        factor()
        t_pobtaf = float(np.median([factor() for _ in range(reps)]))
        log(f"CPU baseline: pobtaf on the sample {t_pobtaf:.2f} s")
        t_c0 = _median_time(lambda: (oinla.f_theta(d0, d0.theta), oinla.selected_inverse_blocks(d0, d0.theta)),
                            reps)
        log(f"CPU baseline: ns=400 nt=12 f + selected inversion {t_c0:.2f} s")
    with threadpool_limits(limits=1):
        t_f1 = _median_time(lambda: oinla.f_theta(d, d.theta), 1)
        log(f"CPU baseline: f on the sample with one thread {t_f1:.2f} s")
    scale = n_full / nt_s
    t_full = t_f * scale
    f_pob = f_pobtaf(nt_s, m * m, nb)
    return {
        "value": 1.0 / (9.0 * t_full),
        "unit": "grad_evals/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle f(theta) (scipy/LAPACK, BTA + BT factorisations) on ns={m * m}, nt={nt_s}, nb={nb}: "
                   f"median of {reps} after 1 warm-up = {t_f:.3f} s on {threads} threads, scaled linearly in nt "
                   f"to nt={n_full} and by 9 evaluations per central-difference gradient"),
        "cpu_model": _cpu_model(),
        "blas": blas,
        "f_sample_s": t_f,
        "f_full_s_scaled": t_full,
        "f_sample_s_1thread": t_f1,
        "pobtaf_sample_s": t_pobtaf,
        "pobtaf_gflops": f_pob / t_pobtaf / 1e9,
        "pobtaf_full_s_scaled": t_pobtaf * scale,
        "ns400_nt12_f_plus_selinv_s": t_c0,
    }


def _f_pobtf(n, b):
    return n * b ** 3 / 3 + 2 * (n - 1) * b ** 3


def _eager_pobtaf(eng, reps: int = 2):
    """(ms, TFLOP/s) of the persistent factorization of the engine's Qc, HIP events on its stream;
    the flops are those of the factorization as run (zero tiles of the profile excluded)."""
    import torch

    from dalia_amd import _lib
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ts = []
    for _ in range(reps + 1):
        eng._assemble(eng.data.theta, "bta", True)
        ev0.record()
        _lib.check(eng.lib.dh_cholesky(eng.h, 1), "dh_cholesky")
        ev1.record()
        torch.cuda.synchronize()
        ts.append(ev0.elapsed_time(ev1) * 1e-3)
    t = float(np.median(ts[1:]))
    lay = eng.solver.layout
    F = f_pobtaf_profile(lay.n, lay.b, lay.a, tile_profile(eng)) if lay.n > 1 else f_pobtaf(lay.n, lay.b, lay.a)
    return t * 1e3, F / t / 1e12


def _coreg_cpu_baseline(d, flops_total: float, reps: int = 3):
    """Oracle coregional f(theta) on one instance (median of reps after a warm-up), scaled by the
    factorization flops of the whole workload."""
    from oracle import coreg as ocoreg
    threads = min(len(os.sched_getaffinity(0)), 16)
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=threads):
        t = _median_time(lambda: ocoreg.f_theta_coreg(d, d.theta), reps)
    f_one = f_pobtaf(d.nt, d.b, d.a) + _f_pobtf(d.nt, d.b)
    return t, threads, t * flops_total / f_one


def bench_coreg_batch(args):
    """The batch of bivariate instances of uneven size: every 3-point diagonal FD Hessian stencil
    point (theta of 9 entries, 19 points per instance) of every instance in one graph replay."""
    import torch

    from dalia_amd import synth
    from dalia_amd.batch import BatchedF, batched_hessian_diags
    from dalia_amd.coreg import CoregInla
    from dalia_amd.inla import hessian_diag_stencil

    specs = synth.COREG_BATCHES[args.workload]
    t0 = time.perf_counter()
    datas = [synth.make_coreg_model(*sp) for sp in specs]
    engines = [CoregInla(d) for d in datas]
    torch.cuda.synchronize()
    log(f"set-up {time.perf_counter() - t0:.1f} s: instances {[(d.ns, d.nt, d.b) for d in datas]}")
    h = 5e-3
    thetas = [d.theta for d in datas]
    stencils = [hessian_diag_stencil(t, h, 3) for t in thetas]
    pts = [p for st in stencils for p in st]
    E = len(pts)
    diags, bf = batched_hessian_diags(engines, thetas, h=h, points=3, tune=True, fused=args.fused)  # + 1st replay
    log(f"autotuned per instance: {[(c['alg'], c.get('wgs', 0), c.get('tw', 0)) for c in bf.choices]} "
        f"(schedule, workgroups, two-ended); {E} evaluations in one graph")
    for _ in range(args.warmup):
        fv = bf(pts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fv = bf(pts)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / args.steps
    log(f"batch: {t_batch * 1e3:.1f} ms per replay of {E} evaluations ({E / t_batch:.1f} f evals/s)")
    # graph values against eager evaluations of the same engines (first, middle and last point)
    k0, worst = 0, 0.0
    for eng, st in zip(engines, stencils):
        for j in (0, len(st) // 2, len(st) - 1):
            fe = eng.f(st[j])
            worst = max(worst, abs(fe - fv[k0 + j]) / abs(fe))
        k0 += len(st)
    # each instance alone (its own graph of its 19 evaluations): the batch hides the smaller ones
    alone = []
    k0 = 0
    for eng, st in zip(engines, stencils):
        bk = BatchedF([eng], [0] * len(st))
        bk(list(st))
        torch.cuda.synchronize()
        ta = time.perf_counter()
        bk(list(st))
        torch.cuda.synchronize()
        alone.append((time.perf_counter() - ta) * 1e3)
    pob = [_eager_pobtaf(eng) for eng in engines]
    flops = sum(len(st) * (f_pobtaf(d.nt, d.b, d.a) + _f_pobtf(d.nt, d.b)) for st, d in zip(stencils, datas))
    log(f"alone: {[round(x, 1) for x in alone]} ms (sum {sum(alone):.1f}); pobtaf per instance "
        f"{[(round(m, 2), round(tf, 1)) for m, tf in pob]} (ms, TFLOP/s); graph vs eager {worst:.1e}")
    cpu = None
    if not args.no_cpu:
        t1, thr, t_cpu = _coreg_cpu_baseline(datas[0], flops)
        cpu = {"value": E / t_cpu, "unit": "f_evals/s", "cores": thr, "kind": "port",
               "sample": (f"oracle coregional f(theta) at ns={datas[0].ns}, nt={datas[0].nt} (b={datas[0].b}): "
                          f"median of 3 after 1 warm-up = {t1:.3f} s, scaled by the factorization flops of the "
                          f"{E} evaluations ({flops / 1e12:.1f} TFLOP)")}
    out = {
        "metric": "f(θ) evals/s, 57 batched diagonal FD Hessian evaluations of 3 uneven bivariate instances, FP64",
        "value": E / t_batch, "unit": "f_evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "instances": [{"ns": d.ns, "nt": d.nt, "b": d.b, "a": d.a,
                                                             "evaluations": len(st)}
                                                            for d, st in zip(datas, stencils)],
                   "evaluations": E, "stencil": "3-point diagonal, h=5e-3", "theta_dim": 9,
                   "fused_rounds": bool(args.fused),
                   "pobtaf_schedule_per_instance": [c["alg"] for c in bf.choices],
                   "persistent_workgroups_per_instance": [c.get("wgs", 0) for c in bf.choices]},
        "tflop_per_step": flops / 1e12, "tflops": flops / t_batch / 1e12,
        "alone_ms": alone, "sum_alone_ms": sum(alone),
        "pobtaf_per_instance": [{"ms": m, "tflops": tf} for m, tf in pob],
        "graph_vs_eager_max_rel": worst,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_coreg(args):
    """The trivariate model: adjoint gradient evaluations, then the 31-point diagonal FD Hessian
    stencil (theta of 15 entries) in one graph replay."""
    import torch

    from dalia_amd import autotune, synth
    from dalia_amd.batch import BatchedF
    from dalia_amd.coreg import CoregInla
    from dalia_amd.inla import hessian_diag_from_values, hessian_diag_stencil

    d = synth.make_coreg_model(*synth.COREG_WORKLOADS[args.workload])
    t0 = time.perf_counter()
    eng = CoregInla(d)
    choice = autotune.tune(eng)
    torch.cuda.synchronize()
    log(f"set-up {time.perf_counter() - t0:.1f} s (nv={d.nv}, ns={d.ns}, nt={d.nt}, b={d.b}, a={d.a}); "
        f"handle {eng.lib.dh_memory_bytes(eng.h) / 1e9:.1f} GB")
    th = d.theta
    for _ in range(args.warmup):
        eng.value_and_grad(th)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f, g = eng.value_and_grad(th)
    torch.cuda.synchronize()
    t_adj = (time.perf_counter() - t0) / args.steps
    log(f"adjoint gradient {t_adj * 1e3:.1f} ms, f={f:.12g}")
    h = 5e-3
    pts = hessian_diag_stencil(th, h, 3)
    bf = BatchedF([eng], [0] * len(pts))
    bf.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fv = bf(list(pts))
    torch.cuda.synchronize()
    t_hess = time.perf_counter() - t0
    hd = hessian_diag_from_values(fv, th.size, h, 3)
    f_eager = eng.f(th)
    log(f"31-point diagonal Hessian stencil: {t_hess:.2f} s in one replay; f(theta) graph {fv[0]:.12g} "
        f"eager {f_eager:.12g}")
    ms, tf = _eager_pobtaf(eng)
    flops = f_pobtaf(d.nt, d.b, d.a)  # dense count: the CPU baseline (oracle) factors dense blocks
    cpu = None
    if not args.no_cpu:
        small = synth.make_coreg_model(20, 16, d.nb, d.nv, 4)
        t1, thr, t_cpu = _coreg_cpu_baseline(small, 31 * (flops + _f_pobtf(d.nt, d.b)))
        cpu = {"value": len(pts) / t_cpu, "unit": "f_evals/s", "cores": thr, "kind": "port",
               "sample": (f"oracle trivariate f(theta) at ns=400, nt=16 (b=1200): median of 3 = {t1:.3f} s, "
                          f"scaled by the factorization flops of the 31 evaluations at b={d.b}, nt={d.nt}")}
    out = {
        "metric": f"∇f(θ) evals/s adjoint, trivariate ns={d.ns} nt={d.nt} nb={d.nb} (b={d.b}) FP64; "
                  "31-point diagonal FD Hessian batch",
        "value": 1.0 / t_adj, "unit": "grad_evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_adj * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "nv": d.nv, "ns": d.ns, "nt": d.nt, "nb": d.nb, "b": d.b, "a": d.a,
                   "theta_dim": int(th.size), "pobtaf_schedule": choice["alg"]},
        "hessian_diag_batch_s": t_hess, "hessian_diag_evaluations": len(pts),
        "hessian_diag_f_evals_per_s": len(pts) / t_hess, "graph_vs_eager_rel": abs(fv[0] - f_eager) / abs(f_eager),
        "adjoint_vs_central_diff_cost": 31 * (t_hess / len(pts)) / t_adj,
        "pobtaf_ms": ms, "pobtaf_tflops": tf, "pobtaf_frac_of_peak": tf / FP64_PEAK_TFLOPS,
        "handle_gb": eng.lib.dh_memory_bytes(eng.h) / 1e9,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_ckpt(args):
    """The spatio-temporal workload with the factor checkpointed every K time blocks
    (`dalia_amd.ckpt.STInlaCkpt`): adjoint gradient and central differences, device memory."""
    import torch

    from dalia_amd import _lib, synth
    from dalia_amd.ckpt import STInlaCkpt, full_handle_bytes

    wl = synth.WORKLOADS[args.workload]
    m, nt, nb, seed = wl
    ns = m * m
    K = args.checkpoint
    t0 = time.perf_counter()
    data = synth.make_st_model(m, nt, nb, seed)
    eng = STInlaCkpt(data, K=K)
    torch.cuda.synchronize()
    mem = eng.memory_bytes()
    full = full_handle_bytes(nt, ns, nb)
    log(f"set-up {time.perf_counter() - t0:.1f} s; checkpointed every {K} blocks: {mem / 1e9:.1f} GB of device "
        f"memory for the solver against {full / 1e9:.1f} GB with the whole factor stored")
    theta = data.theta.copy()
    lib = eng.lib
    for _ in range(args.warmup):
        eng.value_and_grad(theta)
    torch.cuda.synchronize()
    _lib.check(lib.dh_ckpt_timing_enable(eng.ck, _lib.TIMED_KERNELS["k_pobtaf_dag"]), "dh_ckpt_timing_enable")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f, g = eng.value_and_grad(theta)
    torch.cuda.synchronize()
    t_adj = (time.perf_counter() - t0) / args.steps
    tot, cnt = ctypes.c_double(), ctypes.c_longlong()
    _lib.check(lib.dh_ckpt_timing_read(eng.ck, ctypes.byref(tot), ctypes.byref(cnt)), "dh_ckpt_timing_read")
    _lib.check(lib.dh_ckpt_timing_enable(eng.ck, -1), "dh_ckpt_timing_enable")
    log(f"adjoint gradient (checkpointed): {t_adj * 1e3:.1f} ms/eval, f={f:.12g}")
    fd_info = {}
    if args.fd > 0:
        t0 = time.perf_counter()
        f_fd, g_fd = eng.fd_gradient(theta, h=1e-4)
        torch.cuda.synchronize()
        t_fd = time.perf_counter() - t0
        relerr = float(np.linalg.norm(g - g_fd) / np.linalg.norm(g_fd))
        fd_info = {"fd_grad_evals_per_s": 1.0 / t_fd, "fd_ms_per_grad": t_fd * 1e3,
                   "adjoint_vs_fd_speedup": t_fd / t_adj, "adjoint_vs_fd_relerr": relerr}
        log(f"central differences (checkpointed f): {t_fd * 1e3:.0f} ms, speed-up {t_fd / t_adj:.2f}x, "
            f"rel. difference {relerr:.2e}")
    # per evaluation: the forward pass factors every segment, the backward pass all but the last;
    # flops as run: the zero tiles of the sub-diagonal blocks (the same profile for every block of
    # this model, read from its structure) take part in no product
    S = (nt + K - 1) // K
    prof = structural_profile(data)
    f_prof = lambda n: f_pobtaf_profile(n, ns, nb, [prof] * (n - 1))
    per_block = f_prof(3) - f_prof(2)
    flops_eval = (f_prof(nt) - nb ** 3 / 3) + (S - 1) * K * per_block
    t_launch = tot.value * 1e-3 / max(cnt.value, 1)
    achieved = flops_eval * args.steps / (tot.value * 1e-3) / 1e12
    roofline = {"kernel": "k_pobtaf_dag", "bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                "launches": int(cnt.value), "avg_launch_us": t_launch * 1e6,
                "flops_per_launch": flops_eval * args.steps / max(cnt.value, 1)}
    cpu = None
    if not args.no_cpu:
        threads = min(len(os.sched_getaffinity(0)), 16)
        log(f"CPU baseline on {threads} threads (bounded sample: nt=2)")
        cpu = cpu_baseline(wl, nt, threads, reps=3, nt_s=2)
    out = {
        "metric": metric_name(ns, nt, nb) + f", factor checkpointed every {K} time blocks",
        "value": 1.0 / t_adj, "unit": "grad_evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_adj * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "ns": ns, "nt": nt, "nb": nb, "seed": seed, "theta_dim": 4,
                   "checkpoint_every": K, "segments": S},
        "solver_memory_gb": mem / 1e9, "full_factor_memory_gb": full / 1e9,
        **fd_info, "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="st_ns2025_nt100_nb6")
    ap.add_argument("--fd", type=int, default=1, help="central-difference gradients to time (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--overlap", type=int, default=0, choices=(0, 1, 2),
                    help="work run beside each factorization on a second stream (STInla overlap option)")
    ap.add_argument("--roofline-kernel", default="k_syrk")
    ap.add_argument("--fused", action="store_true",
                    help="batches: the factorizations of each round as one launch over all instances")
    ap.add_argument("--graph", action="store_true",
                    help="time the adjoint evaluation as one captured HIP graph replayed per step")
    ap.add_argument("--checkpoint", type=int, default=0,
                    help="factor checkpointed every K time blocks and recomputed in the backward pass (0: stored)")
    ap.add_argument("--twisted", type=int, default=1, choices=(0, 1),
                    help="two-ended factorizations (blocks eliminated from both ends, meeting in the middle)")
    ap.add_argument("--profile", type=int, default=1, choices=(0, 1),
                    help="zero-tile profiles of the sub-diagonal blocks in the factorization (0: every tile dense)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 2 GiB device-to-device copy (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(one f(theta) evaluation per GPU; see the sharding design notes)", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        n = 1 << 28  # doubles: 2 GiB read + 2 GiB written by torch's elementwise copy kernel
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {n * 8 / 2 ** 30:.1f} GiB three times")
        return

    from dalia_amd import synth as _synth
    if args.workload in _synth.COREG_BATCHES:
        return bench_coreg_batch(args)
    if args.workload in _synth.COREG_WORKLOADS:
        return bench_coreg(args)
    if args.checkpoint > 0:
        return bench_ckpt(args)

    from dalia_amd import _lib, synth
    from dalia_amd.inla import STInla

    wl = synth.WORKLOADS[args.workload]
    m, nt, nb, seed = wl
    ns = m * m
    log(f"building synthetic model {args.workload} (ns={ns}, nt={nt}, nb={nb}, seed={seed})")
    t0 = time.perf_counter()
    data = synth.make_st_model(m, nt, nb, seed)
    eng = STInla(data, overlap=args.overlap, twisted=bool(args.twisted))
    eng.solver.set_tile_profile(bool(args.profile))
    torch.cuda.synchronize()
    log(f"set-up {time.perf_counter() - t0:.1f} s (closed-form prior: {eng.dm.closed_form})")
    theta = data.theta.copy()
    lib = eng.lib
    from dalia_amd import autotune
    choice = autotune.tune(eng)
    log(f"autotuned kernel variants: {choice}")

    # the dominant kernel: its launches inside the timed steps are bracketed by HIP events on the
    # library's stream
    kname = args.roofline_kernel
    if kname == "k_syrk" and choice["alg"] == 1:
        kname = "k_pobtaf_dag"  # the persistent schedule has no separate Schur-update launch
    kind = _lib.TIMED_KERNELS[kname]

    if args.graph:
        # the whole adjoint evaluation as one HIP graph (events recorded inside a graph cannot be
        # timed, so the roofline kernel is timed on eager evaluations of the same step afterwards)
        f_e, g_e = eng.value_and_grad(theta)
        eng.capture_grad()
        step = eng.value_and_grad_graph
    else:
        step = eng.value_and_grad
    for w in range(args.warmup):
        f, g = step(theta)
        log(f"warm-up {w}: f={f:.10g}")
    torch.cuda.synchronize()
    if not args.graph:
        _lib.check(lib.dh_timing_enable(eng.h, kind), "dh_timing_enable")
    t0 = time.perf_counter()
    for k in range(args.steps):
        f, g = step(theta)
    torch.cuda.synchronize()
    t_adj = (time.perf_counter() - t0) / args.steps
    if args.graph:
        _lib.check(lib.dh_timing_enable(eng.h, kind), "dh_timing_enable")
        for _ in range(2):
            eng.value_and_grad(theta)
    tot = ctypes.c_double()
    cnt = ctypes.c_longlong()
    _lib.check(lib.dh_timing_read(eng.h, ctypes.byref(tot), ctypes.byref(cnt)), "dh_timing_read")
    _lib.check(lib.dh_timing_enable(eng.h, -1), "dh_timing_enable")
    if args.graph:
        diff = float(np.linalg.norm(g - g_e) / np.linalg.norm(g_e))
        log(f"graph replay against the eager evaluation: f {abs(f - f_e) / abs(f_e):.1e}, gradient {diff:.1e}")
        if abs(f - f_e) > 1e-12 * abs(f_e) or diff > 1e-10:
            raise RuntimeError("the replayed adjoint evaluation differs from the eager one")
    log(f"adjoint gradient{' (graph replay)' if args.graph else ''}: {t_adj * 1e3:.1f} ms/eval, f={f:.12g}, grad={g}")

    # plain GPU central differences (reference method, h = 1e-4)
    fd_info = {}
    if args.fd > 0:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.fd):
            f_fd, g_fd = eng.fd_gradient(theta, h=1e-4)
        torch.cuda.synchronize()
        t_fd = (time.perf_counter() - t0) / args.fd
        relerr = float(np.linalg.norm(g - g_fd) / np.linalg.norm(g_fd))
        fd_info = {"fd_grad_evals_per_s": 1.0 / t_fd, "fd_ms_per_grad": t_fd * 1e3,
                   "adjoint_vs_fd_speedup": t_fd / t_adj, "adjoint_vs_fd_relerr": relerr}
        log(f"central differences: {t_fd * 1e3:.1f} ms/gradient, adjoint speed-up {t_fd / t_adj:.2f}x, "
            f"rel. difference {relerr:.2e}")

    # pobtaf alone, HIP events on the library's stream (torch's current stream); nothing runs
    # beside it (the overlap option is switched off for this measurement)
    eng._overlap(0)
    eng._assemble(theta, "bta", True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for r in range(3):
        eng._assemble(theta, "bta", True)
        ev0.record()
        _lib.check(lib.dh_cholesky(eng.h, 1), "dh_cholesky")
        ev1.record()
        torch.cuda.synchronize()
        times.append(ev0.elapsed_time(ev1))
    t_pobtaf = float(np.median(times)) * 1e-3
    # algorithmic flops of the factorization as run: the zero tiles of the sub-diagonal blocks
    # (their profile, read back from the library) take part in no product
    F_dense = f_pobtaf(nt, ns, nb)
    F = f_pobtaf_profile(nt, ns, nb, tile_profile(eng)) if args.profile else F_dense
    pobtaf_tflops = F / t_pobtaf / 1e12
    log(f"pobtaf: {t_pobtaf * 1e3:.2f} ms, {pobtaf_tflops:.2f} TFLOP/s ({100 * pobtaf_tflops / FP64_PEAK_TFLOPS:.1f}% "
        f"of FP64 matrix peak) on {F:.4g} flops ({F / F_dense:.3f} of the dense count; dense-equivalent "
        f"{F_dense / t_pobtaf / 1e12:.2f} TFLOP/s)")

    # roofline of the dominant kernel, per launch, from its launches inside the timed steps
    avg_s = tot.value * 1e-3 / max(cnt.value, 1)
    if kname == "k_pobtaf_dag":
        per = F - nb ** 3 / 3  # the whole factorization (as run, profile included) except the tip's own Cholesky
    elif kname == "k_syrk" and choice["syrk"] == 1:
        per = float(ns) ** 3  # k_syrk128: the SYRK alone (arrow and tip run in k_syrk_tail)
        kname = "k_syrk128"
    elif kname == "k_syrk":
        per = syrk_flops_per_launch(nt, ns, nb)
    elif kname == "k_update":
        per, _ = update_flops_per_launch(nt, ns, nb)
    else:
        per = float("nan")
    achieved = per / avg_s / 1e12
    tw = kname == "k_pobtaf_dag" and eng.twisted
    roofline = {"kernel": kname, "bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                # the committed counter summary is of the one-ended launch without zero-tile profiles;
                # the profiler's counter mode does not complete on the two-ended build, so the other
                # configurations have none
                "traffic": None if (tw or args.profile) else committed_traffic(kname, args.workload),
                "launches": int(cnt.value), "avg_launch_us": avg_s * 1e6, "flops_per_launch": per}
    if tw:
        roofline["timed_span"] = ("two-ended factorization: the two-sided k_pobtaf_dag launch, k_twist_join "
                                  "and block m's k_pobtaf_dag launch, between one event pair")
    log(f"roofline {kname}: {cnt.value} launches, avg {avg_s * 1e6:.1f} us, {achieved:.2f} TFLOP/s")
    dgemm = dgemm_ceiling()
    roofline["dgemm_ceiling"] = dgemm
    roofline["frac_of_dgemm"] = achieved / dgemm
    log(f"library DGEMM ceiling (torch.mm FP64, 8192^3): {dgemm:.2f} TFLOP/s; {kname} at {achieved / dgemm:.2f} of it")

    # HBM-bound kernels: the assembly of Qc (k_assemble_rows: the dense diagonal and sub-diagonal
    # block rows) and the solve (pobtas N then C, every launch of dh_solve), events on the stream
    hbm = {}
    _lib.check(lib.dh_timing_enable(eng.h, _lib.TIMED_KERNELS["k_assemble_rows"]), "dh_timing_enable")
    for _ in range(3):
        eng._assemble(theta, "bta", True)
    _lib.check(lib.dh_timing_read(eng.h, ctypes.byref(tot), ctypes.byref(cnt)), "dh_timing_read")
    _lib.check(lib.dh_timing_enable(eng.h, -1), "dh_timing_enable")
    asm_bytes = 8.0 * (2 * nt - 1) * ns * ns
    asm_s = tot.value * 1e-3 / max(cnt.value, 1)
    hbm["k_assemble_rows"] = {"bound": "hbm", "achieved": asm_bytes / asm_s / 1e9, "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": asm_bytes / asm_s / 1e9 / HBM_PEAK_GBS,
                              "avg_launch_us": asm_s * 1e6, "bytes_per_launch": asm_bytes}
    _lib.check(lib.dh_cholesky(eng.h, 1), "dh_cholesky")
    x = eng.mu.clone()
    _lib.check(lib.dh_solve(eng.h, x.data_ptr(), 1), "dh_solve")  # also forms the packed inverse once
    ts = []
    for _ in range(3):
        ev0.record()
        _lib.check(lib.dh_solve(eng.h, x.data_ptr(), 1), "dh_solve")
        ev1.record()
        torch.cuda.synchronize()
        ts.append(ev0.elapsed_time(ev1) * 1e-3)
    sol_s = float(np.median(ts))
    sol_bytes = 16.0 * (nt * ns * (ns + 1) / 2 + (nt - 1) * ns * ns + nt * nb * ns + nb * (nb + 1) / 2)
    hbm["pobtas"] = {"bound": "hbm", "achieved": sol_bytes / sol_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": sol_bytes / sol_s / 1e9 / HBM_PEAK_GBS, "ms": sol_s * 1e3, "bytes": sol_bytes,
                     "launches_per_solve": solve_launches(nt, nb, bool(eng.twisted))}
    log(f"assembly {asm_s * 1e3:.2f} ms ({hbm['k_assemble_rows']['achieved']:.0f} GB/s); "
        f"pobtas {sol_s * 1e3:.2f} ms ({hbm['pobtas']['achieved']:.0f} GB/s)")

    cpu = None
    if not args.no_cpu:
        threads = min(len(os.sched_getaffinity(0)), 16)
        log(f"CPU baseline on {threads} threads (bounded sample)")
        cpu = cpu_baseline(wl, nt, threads)
        log(f"CPU baseline {cpu['value']:.3g} grad evals/s")

    out = {
        "metric": metric_name(ns, nt, nb),
        "value": 1.0 / t_adj,
        "unit": "grad_evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_adj * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "ns": ns, "nt": nt, "nb": nb, "seed": seed, "theta_dim": 4,
                   "n_obs": data.n_obs, "kernel_variants": {"pobtaf_schedule": "tile-task" if choice["alg"] == 1 else "per-panel",
                                                        "syrk": choice["syrk"]},
                   "overlap": args.overlap, "graph": bool(args.graph), "twisted": bool(eng.twisted),
                   "tile_profile": bool(args.profile)},
        "pobtaf_ms": t_pobtaf * 1e3,
        "pobtaf_tflops": pobtaf_tflops,
        "pobtaf_frac_of_peak": pobtaf_tflops / FP64_PEAK_TFLOPS,
        "pobtaf_flops": F,
        "pobtaf_flops_dense": F_dense,
        "pobtaf_dense_equiv_tflops": F_dense / t_pobtaf / 1e12,
        **fd_info,
        "roofline": roofline,
        "roofline_hbm": hbm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
