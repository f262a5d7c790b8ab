#!/usr/bin/env python
"""Benchmark: fused Θ-apply throughput at 192³ on one MI355X.

Workload (one "step" = one Θ·X over a 32-vector block): the 10,000-point
synthetic rod network (seed 3, rods of radius 0.331836 × mean bond length,
ε = 8.41 in rods / 1 outside, periodic rasterisation) on a 192³ grid, Γ point,
X = Γ-masked unit-row complex64 block (m = 32, counter-based stream seed 0)
resident in HBM before timing.

Prints ONE JSON line.  `value` = vector-applies/s (m × steps / wall time).
`roofline` prices the fused Θ-apply against HBM with the algorithmic bytes
B_Θ = 32·G³ + 4·G³/m per vector-apply (read 2 components, write 2, ε⁻¹
amortised over the block), timed with HIP events on the launch stream.
`cpu_baseline` times the CPU oracle (SciPy FFT restatement, complex64, all
host cores) on a bounded sample of the same workload.

The same line carries the metric's second half and the rest of the path:
`serving` — solves/s and p50/p99 latency of a fixed 24-request sample through
the batched engines and through the plain one-at-a-time LOBPCG loop, with the
eigenvalue error against committed FP64 oracle values; `paths` — the
preconditioner, fused Chebyshev and KPM steps, Gram and combine at 192³, each
priced on its own algorithmic bytes or flops; `step_ms_p10_p50_p90` — the
spread of the timed Θ steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
G_BENCH = 192
M_BENCH = 32
NET = dict(n_points=10000, seed=3, radius_factor=0.331836)


def algorithmic_bytes_per_vector(G, m):
    return 32 * G ** 3 + 4 * G ** 3 / m


MFMA_F32_PEAK_TFS = 157.3   # dense FP32 matrix cores, MI355X
MFMA_F64_PEAK_TFS = 78.6
TRAFFIC_PROFILE = "profiles/theta_apply_192_m32_pmc.md"


def _launch_seconds(fn, reps, stream):
    """Mean duration of one fn() launch set from HIP events on `stream`."""
    import torch
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def path_rooflines(op, X, Y, G, stream, reps=5):
    """The other kernels of the path on the same structure and grid, each priced
    on its own algorithmic bytes (or flops) per launch."""
    import torch
    from eigenfns_amd.linalg import combine_into, gram
    m = X.shape[0]
    G3 = G ** 3
    out = {}

    def hbm(name, fn, alg_bytes, unit_count, unit):
        t = _launch_seconds(fn, reps, stream)
        gbs = alg_bytes / t / 1e9
        out[name] = {"per_s": round(unit_count / t, 2), "unit": unit, "launch_ms": round(t * 1e3, 3),
                     "alg_bytes_per_launch": int(alg_bytes), "achieved_GBs": round(gbs, 1),
                     "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}

    X8, Y8 = X[:8].contiguous(), Y[:8].contiguous()
    hbm("theta_m8", lambda: op.theta_into(X8, Y8), algorithmic_bytes_per_vector(G, 8) * 8, 8,
        "vector-applies/s")
    hbm("precondition_m32", lambda: op.precondition_into(X, Y),
        algorithmic_bytes_per_vector(G, m) * m, m, "vector-applies/s")
    T0 = X.clone()
    Z = torch.zeros_like(X)
    alpha = 2.0 / 1.0e3
    hbm("cheb_step_m32", lambda: op.cheb_step(X, T0, Z, alpha, 1e-3),
        (80 * G3 + 4 * G3 / m) * m, m, "vector-degrees/s")
    np_ = 16
    mom = torch.zeros((np_, 64), dtype=torch.float64, device=X.device)
    T1k, T0k = X[:np_].contiguous(), Y[:np_].contiguous()
    hbm("kpm_step_16probes", lambda: op.kpm_step(T1k, T0k, mom, 5, alpha),
        (48 * G3 + 4 * G3 / np_) * np_, np_, "probe-degrees/s")
    est = torch.zeros(np_, dtype=torch.float64, device=X.device)
    Zk = Z[:np_]
    hbm("kpm_count_step_16probes", lambda: op.kpm_count_step(T1k, T0k, Zk, est, 1e-3, alpha),
        (64 * G3 + 4 * G3 / np_) * np_, np_, "probe-degrees/s")
    del T0, Z, T1k, T0k, Zk
    # Rayleigh–Ritz-sized dense kernels: 3m = 96 rows at G³
    S = torch.cat([X, Y, X.roll(1, dims=0)]).contiguous()
    ms = S.shape[0]
    t = _launch_seconds(lambda: gram(S, S), 2, stream)
    # algorithmic: 8 real flops per complex multiply-add; the FP64 matrix-core
    # engine executes 6 (three real products per complex entry)
    flops = 8.0 * ms * ms * 2 * G3
    out["gram_96x96"] = {"launch_ms": round(t * 1e3, 3), "TFLOPs": round(flops / t / 1e12, 2),
                         "mfma_TFLOPs": round(0.75 * flops / t / 1e12, 2),
                         "frac_mfma_f64": round(0.75 * flops / t / 1e12 / MFMA_F64_PEAK_TFS, 4),
                         "engine": "mfma_f64"}
    # the opt-in FP32 matrix-core engine (four real products per complex entry,
    # FP32 partial sums of at most 64 products added into FP64) on the same shape
    from eigenfns_amd.linalg import gram_engine
    prev = gram_engine("mfma_f32")
    try:
        t32 = _launch_seconds(lambda: gram(S, S), 2, stream)
        g32 = gram(S, S)
    finally:
        gram_engine(prev)
    g64 = gram(S, S)
    out["gram_96x96_mfma_f32"] = {
        "launch_ms": round(t32 * 1e3, 3), "TFLOPs": round(flops / t32 / 1e12, 2),
        "mfma_TFLOPs": round(flops / t32 / 1e12, 2),
        "frac_mfma_f32": round(flops / t32 / 1e12 / MFMA_F32_PEAK_TFS, 4),
        "max_abs_diff_vs_f64_engine_rel": float((g32 - g64).abs().max() / g64.abs().max())}
    del g32, g64
    C = torch.randn(ms, m, dtype=torch.complex64, device=X.device)
    t = _launch_seconds(lambda: combine_into(C, S, Y), reps, stream)
    byts = 16.0 * G3 * (ms + m)
    out["combine_96to32"] = {"launch_ms": round(t * 1e3, 3), "achieved_GBs": round(byts / t / 1e9, 1),
                             "frac_hbm": round(byts / t / 1e9 / HBM_PEAK_GBS, 4)}
    del S, C
    # small dense FP64 eigensolves of Rayleigh–Ritz size (72 × 72 Hermitian,
    # lowest 24): one matrix and a batch of 16, the library's bisection +
    # inverse-iteration and QL kernels against rocSOLVER's divide and conquer
    # (latency-bound; no roofline)
    from eigenfns_amd.linalg import eigh_batched_device
    gen = torch.Generator(device=X.device).manual_seed(7)
    R = torch.randn((16, 72, 72), dtype=torch.complex128, device=X.device, generator=gen)
    R = R + R.conj().transpose(1, 2)
    for method in ("heevx", "ql", "dc"):
        for nb in (1, 16):
            t = _launch_seconds(lambda: eigh_batched_device(R[:nb], lowest=24, method=method), reps, stream)
            out[f"heev_72_lowest24_x{nb}_{method}"] = {"launch_ms": round(t * 1e3, 3)}
    del R
    return out


def committed_traffic(G, m, chunk):
    """HBM bytes per Θ launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary, if it covers this workload."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), TRAFFIC_PROFILE)
    try:
        text = open(path).read()
    except OSError:
        return None
    vals = {}
    for line in text.splitlines():
        if ":" in line and line.split(":", 1)[0].strip() in ("workload", "traffic_bytes_per_launch"):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    if vals.get("workload") != f"theta_apply_{G}^3_m{m}_chunk{chunk}":
        return None
    try:
        return int(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def serving_sample():
    """The serving metric on a fixed sample: 24 seeded requests (networks of
    64/216/1000 points, grids 32³/48³/64³, random Bloch k, 8–64 bands, all
    arriving at t = 0) through the batched engines, then one at a time through
    the plain LOBPCG loop; eigenvalues of the requests covered by the committed
    FP64 oracle fixture are checked against it."""
    from eigenfns_amd import workloads as wl
    from eigenfns_amd.serving import StructureCache, serve_batched, serve_plain
    specs = wl.serving_sample()
    cache = StructureCache()
    for sp in specs:
        cache.get(sp.n_points, sp.grid)
    solver = dict(wl.SERVING_SOLVER)
    warm = [sp for sp in specs if sp.nev <= 20][:2]      # library, allocator and thread pools
    serve_batched(warm, cache, **solver)
    serve_plain(warm, cache, **solver)
    res, summ = serve_batched(specs, cache, **solver)
    pres, psumm = serve_plain(specs, cache, **solver)
    by = {r.rid: r.vals for r in res}
    diff = max(float(np.max(np.abs(by[p["rid"]] - p["vals"]) / np.abs(p["vals"]))) for p in pres)
    gpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                         "solve_golden.npz")
    errs = []
    if os.path.exists(gpath):
        with np.load(gpath, allow_pickle=False) as z:
            for rid, vals in by.items():
                key = f"srv_{rid}"
                if key in z.files:
                    ref = z[key]
                    errs.append(float(np.max(np.abs(vals[:len(ref)] - ref) / np.abs(ref))))
    return {"solves_per_s": round(summ["solves_per_s"], 3), "p50_s": round(summ["p50_s"], 3),
            "p99_s": round(summ["p99_s"], 3), "wall_s": round(summ["wall_s"], 2),
            "plain_solves_per_s": round(psumm["solves_per_s"], 3),
            "plain_p50_s": round(psumm["p50_s"], 3), "plain_p99_s": round(psumm["p99_s"], 3),
            "speedup_vs_plain": round(summ["solves_per_s"] / psumm["solves_per_s"], 3),
            "max_rel_err_vs_fp64_oracle": max(errs) if errs else None,
            "oracle_checked_requests": len(errs),
            "max_rel_diff_batched_vs_plain": diff, "requests": len(specs),
            "slots": summ["slots"],
            "sample": "24 seeded requests (request_stream seed 11), N∈{64,216,1000}, "
                      "G∈{32,48,64}, k~U[-1/2,1/2)^3, nev 8-64, block 24, guard 8, "
                      "residual tol 1e-4, all arriving at t=0"}


def cpu_c1(eps, L):
    """The oracle's FP64 LOBPCG on C1 (64-point network, 32³, 24 bands), all host cores."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from eigenfns_amd import workloads as wl
    from oracle.lobpcg_np import lobpcg_blocks
    from oracle.operator_np import MaxwellOperatorNP, set_workers
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    set_workers(cores)
    c = wl.C1
    t0 = time.perf_counter()
    vals, _, st = lobpcg_blocks(MaxwellOperatorNP(eps, L), c["nev"], m=c["m"], guard=c["guard"],
                                tol=c["tol"], seed=0, normalized_rq=True)
    return vals, time.perf_counter() - t0, cores, st["theta_applications"]


def cpu_baseline(eps, L, G, budget_s=12.0):
    """Time the oracle's Θ-apply (SciPy, complex64, all cores) on a bounded sample."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import operator_np
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    operator_np.set_workers(cores)
    op = operator_np.MaxwellOperatorNP(eps, L, fft_dtype=np.complex64, out_dtype=np.complex64)
    rng = np.random.default_rng(0)
    X = (rng.standard_normal((1, 2, G, G, G)) + 1j * rng.standard_normal((1, 2, G, G, G))).astype(
        np.complex64)
    op.theta(X)  # warm (plans, pages)
    n = 0
    t0 = time.perf_counter()
    while True:
        op.theta(X)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 64:
            break
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": n / el, "unit": "vector-applies/s", "cores": int(cores), "kind": "port",
            "sample": f"{n} single-vector Θ-applies at {G}^3 (complex64 SciPy FFT, "
                      f"{cores} workers, {el:.1f} s); cpu_count={os.cpu_count()}; {cpu_model}"}


def cpu_side_by_side(stream, budget_s=8.0):
    """Two more shapes timed on the host cores (the oracle, complex64 SciPy FFT,
    all cores) and on the GPU (the product path): the Θ-apply at 128³ on a
    32-vector block (1000-point network) and one KPM doubling step at 256³ for
    one probe (10,000-point network).  CPU legs are bounded samples."""
    import torch
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import operator_np
    from eigenfns_amd import workloads as wl
    from eigenfns_amd.operator import MaxwellOperator
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    operator_np.set_workers(cores)
    rng = np.random.default_rng(0)
    out = {"cores": int(cores)}

    def cpu_loop(fn):
        fn()
        n, t0 = 0, time.perf_counter()
        while True:
            fn()
            n += 1
            el = time.perf_counter() - t0
            if el > budget_s or n >= 32:
                return n, el

    # Θ at 128³, m = 32
    G = 128
    eps, L = wl.network(1000, G)
    op = MaxwellOperator(eps, L)
    X = torch.randn((32, 2, G, G, G), dtype=torch.complex64, device="cuda")
    Y = torch.empty_like(X)
    t = _launch_seconds(lambda: op.theta_into(X, Y), 5, stream)
    opn = operator_np.MaxwellOperatorNP(eps, L, fft_dtype=np.complex64, out_dtype=np.complex64)
    Xc = (rng.standard_normal((1, 2, G, G, G)) + 1j * rng.standard_normal((1, 2, G, G, G))).astype(np.complex64)
    n, el = cpu_loop(lambda: opn.theta(Xc))
    out["theta_128_m32"] = {"gpu_per_s": round(32 / t, 1), "cpu_per_s": round(n / el, 2),
                            "unit": "vector-applies/s",
                            "cpu_sample": f"{n} single-vector applies in {el:.1f} s"}
    del X, Y, op
    # one KPM doubling step at 256³, one probe: T2 = 2(αΘT1 − T1) − T0 and the two dots
    G = 256
    eps, L = wl.network(10000, G)
    op = MaxwellOperator(eps, L)
    alpha = 2.0 / 4.0e3
    T1 = torch.randn((1, 2, G, G, G), dtype=torch.complex64, device="cuda")
    T0 = torch.randn_like(T1)
    mom = torch.zeros((1, 64), dtype=torch.float64, device="cuda")
    t = _launch_seconds(lambda: op.kpm_step(T1, T0, mom, 5, alpha), 5, stream)
    opn = operator_np.MaxwellOperatorNP(eps, L, fft_dtype=np.complex64, out_dtype=np.complex64)
    t1 = (rng.standard_normal((1, 2, G, G, G)) + 1j * rng.standard_normal((1, 2, G, G, G))).astype(np.complex64)
    t0 = t1.copy()

    def cpu_step():
        t2 = 2.0 * (alpha * opn.theta(t1) - t1) - t0
        return np.vdot(t1, t1).real, np.vdot(t2, t1).real
    n, el = cpu_loop(cpu_step)
    out["kpm_step_256_1probe"] = {"gpu_ms": round(t * 1e3, 3), "cpu_ms": round(el / n * 1e3, 1),
                                  "cpu_sample": f"{n} steps in {el:.1f} s"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=G_BENCH)
    ap.add_argument("--m", type=int, default=M_BENCH)
    ap.add_argument("--chunk", type=int, default=0, help="vectors per workspace pass (0 = auto)")
    ap.add_argument("--passes", type=int, default=0, choices=[0, 3, 5],
                    help="Θ pass structure (0 = the tuning table's choice for the grid)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-extra", action="store_true",
                    help="also time Θ at 128³ (m = 32) and a KPM step at 256³ on CPU and GPU")
    ap.add_argument("--no-paths", action="store_true",
                    help="skip the other path kernels (preconditioner, Chebyshev/KPM steps, Gram, combine)")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving sample (batched engines vs the plain LOBPCG loop)")
    ap.add_argument("--cpu-c1", action="store_true",
                    help="also time C1 (32³, 24 bands) end to end on the GPU and with the FP64 "
                         "oracle on the host cores (about a minute of CPU)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one plain copy kernel (efn_axpby, 16 B per complex) of the block "
                         "before timing, to calibrate rocprof FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the Θ path "
              "shards by request, k-point, probe or vector column with no collective on Θ, so "
              "one GPU's kernel at its real shape is the measured unit.", file=sys.stderr)
        return 2

    import torch
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from eigenfns_amd.operator import MaxwellOperator
    from eigenfns_amd.structure import network_eps, filling_fraction
    from eigenfns_amd.linalg import normalize_rows

    G, m = args.grid, args.m
    t0 = time.perf_counter()
    eps, L = network_eps(NET["n_points"], NET["seed"], G, radius_factor=NET["radius_factor"])
    t_raster = time.perf_counter() - t0
    print(f"[bench] network N={NET['n_points']} seed {NET['seed']} rasterised at {G}^3 in "
          f"{t_raster:.1f}s, ff={filling_fraction(eps):.4f}, L={L:.4f}", file=sys.stderr, flush=True)
    op = MaxwellOperator(eps, L, max_chunk=args.chunk, passes=args.passes)
    X = normalize_rows(op.random_block(m, seed=0))
    Y = torch.empty_like(X)
    stream = torch.cuda.current_stream()
    if args.calib_copy:
        from eigenfns_amd.linalg import axpby
        axpby(X, 1.0, out=Y)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy: {X.numel() * 8} B read, {X.numel() * 8} B written",
              file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        op.theta_into(X, Y)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t_start = time.perf_counter()
    ev0.record(stream)
    for i in range(args.steps):
        op.theta_into(X, Y)
        evs[i].record(stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ev_ms = ev0.elapsed_time(ev1)
    step_ms = np.array([ev0.elapsed_time(evs[0])] +
                       [evs[i - 1].elapsed_time(evs[i]) for i in range(1, args.steps)])
    ms_per_step = 1e3 * wall / args.steps
    value = m * args.steps / wall
    # roofline of the fused Θ-apply (one launch set = one step over m vectors)
    launch_s = ev_ms / 1e3 / args.steps
    alg_bytes = algorithmic_bytes_per_vector(G, m) * m
    achieved = alg_bytes / launch_s / 1e9
    line = {
        "metric": "Θ-applies/s at 192³ (% HBM peak); solves/s + p99 latency at 1e-6 eig accuracy",
        "value": round(value, 2),
        "unit": "vector-applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / 43.0, 2),
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"theta_apply_{G}^3_m{m}",
                   "passes": op.passes,
                   "structure": f"synthetic periodic 4-NN rod network N={NET['n_points']} "
                                f"seed {NET['seed']} r={NET['radius_factor']}·mean bond, "
                                f"eps 8.41/1, periodic raster", "grid": G, "block": m,
                   "k_point": "Gamma", "chunk": op.chunk},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": committed_traffic(G, m, op.chunk),
                     "kernel": {3: "fused Θ-apply (xin_split, yzlds, xout_split)" if G == 192 else
                                "fused Θ-apply (xin, yzmid, xout)",
                                5: "fused Θ-apply (zin, ypass, xmid, ypass, zout)"}[op.passes],
                     "alg_bytes_per_launch": int(alg_bytes),
                     "launch_ms": round(launch_s * 1e3, 4)},
        "step_ms_p10_p50_p90": [round(float(np.percentile(step_ms, q)), 4) for q in (10, 50, 90)],
        "cpu_baseline": None,
    }
    if line["roofline"]["traffic"] is not None:
        line["roofline"]["traffic_source"] = TRAFFIC_PROFILE
    if not args.no_paths and G == G_BENCH and m == M_BENCH:
        line["paths"] = path_rooflines(op, X, Y, G, stream)
    if not args.no_serving and G == G_BENCH and m == M_BENCH:
        del X, Y
        torch.cuda.empty_cache()
        line["serving"] = serving_sample()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(eps, L, G)
    if args.cpu_extra:
        line["cpu_side_by_side"] = cpu_side_by_side(stream)
    if args.cpu_c1:
        from eigenfns_amd import workloads as wl
        from eigenfns_amd.solve_bench import run_c1
        gpu = run_c1()
        eps1, L1 = wl.network(wl.C1["n_points"], wl.C1["grid"])
        cvals, tc, cores, n_theta = cpu_c1(eps1, L1)
        gv = np.asarray(gpu.pop("vals"))
        line["c1_end_to_end"] = {**gpu, "cpu_oracle_seconds": round(tc, 1), "cpu_cores": cores,
                                 "cpu_theta_applications": n_theta,
                                 "speedup_vs_cpu": round(tc / gpu["gpu_seconds"], 1),
                                 "max_rel_diff_gpu_vs_cpu_oracle":
                                     float(np.max(np.abs(gv - cvals) / cvals))}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
