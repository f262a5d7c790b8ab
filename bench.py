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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=G_BENCH)
    ap.add_argument("--m", type=int, default=M_BENCH)
    ap.add_argument("--chunk", type=int, default=0, help="vectors per workspace pass (0 = auto)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one plain copy kernel (efn_axpby, 16 B per complex) of the block "
                         "before timing, to calibrate rocprof FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); every GPU box "
              "here has a single MI355X — see the sharding paragraph of DESIGN.md.",
              file=sys.stderr)
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
    op = MaxwellOperator(eps, L, max_chunk=args.chunk)
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
    t_start = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        op.theta_into(X, Y)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ev_ms = ev0.elapsed_time(ev1)
    ms_per_step = 1e3 * wall / args.steps
    value = m * args.steps / wall
    # roofline of the fused Θ-apply (one launch set = one step over m vectors)
    launch_s = ev_ms / 1e3 / args.steps
    alg_bytes = algorithmic_bytes_per_vector(G, m) * m
    achieved = alg_bytes / launch_s / 1e9
    line = {
        "metric": "Θ-applies/s at 192³ (% HBM peak)",
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
                   "structure": f"synthetic periodic 4-NN rod network N={NET['n_points']} "
                                f"seed {NET['seed']} r={NET['radius_factor']}·mean bond, "
                                f"eps 8.41/1, periodic raster", "grid": G, "block": m,
                   "k_point": "Gamma", "chunk": op.chunk},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel": "fused Θ-apply (zin, ypass, xmid, ypass, zout)",
                     "alg_bytes_per_launch": int(alg_bytes),
                     "launch_ms": round(launch_s * 1e3, 4)},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(eps, L, G)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
