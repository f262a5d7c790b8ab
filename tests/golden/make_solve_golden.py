#!/usr/bin/env python
"""Write FP64 oracle eigenvalues for the solve workloads (CPU, this container).

Cases (all from eigenfns_amd.workloads):
  c1        C1: 64-point network, 32³, Γ, lowest 24 bands — oracle LOBPCG
            (FP64 FFT restatement of solver.py:136-382), block 16, guard 8,
            residual tol 1e-8.
  c2dense   C2's network (1000 points) rasterised at 16³, Γ: lowest 128
            eigenvalues of the dense FP64 matrix (n = 8192).
  c3_k<j>   C3's network at 64³ for the two k-points of the path with the
            narrowest blocks (case c3) and for the widest and a median-width
            one (case c3wide): oracle LOBPCG with the configuration's width,
            nev = width − 4, guard 4, tol 1e-8.
  srv_<rid> the serving sample's requests at 32³ with nev ≤ 32: oracle
            LOBPCG, block 32, guard 8, tol 1e-8.
  c5_<rid>  one request per grid 48³, 64³, 96³, 128³ of the C5 stream
            (request_stream(2000, seed 0); the smallest nev, first by id):
            oracle LOBPCG, block 16, guard 8, tol 1e-8.

Output: tests/golden/solve_golden.npz (eigenvalues and the case parameters
as JSON).  Run:  python tests/golden/make_solve_golden.py [case ...]
"""
from __future__ import annotations

import json
import os
import sys
import time
from pathlib import Path

import numpy as np
import scipy.linalg as sla

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from eigenfns_amd import workloads as wl  # noqa: E402
from oracle.dense_np import dense_theta  # noqa: E402
from oracle.lobpcg_np import lobpcg_blocks  # noqa: E402
from oracle.operator_np import MaxwellOperatorNP, set_workers  # noqa: E402

OUT = HERE / "solve_golden.npz"
TOL = 1e-8


def _lobpcg(eps, L, k, nev, m, guard, seed=0):
    op = MaxwellOperatorNP(eps, L, k)
    vals, _, st = lobpcg_blocks(op, nev, m=m, guard=guard, tol=TOL, maxit=2000, seed=seed,
                                normalized_rq=True)
    return np.asarray(vals, np.float64), st


def case_c1():
    eps, L = wl.network(wl.C1["n_points"], wl.C1["grid"])
    vals, _ = _lobpcg(eps, L, (0.0, 0.0, 0.0), wl.C1["nev"], wl.C1["m"], wl.C1["guard"])
    return {"c1": (vals, dict(wl.C1, k=(0.0, 0.0, 0.0), solver="oracle LOBPCG", tol=TOL))}


def case_c2dense():
    G = wl.C2["dense_grid"]
    eps, L = wl.network(wl.C2["n_points"], G)
    M = dense_theta(MaxwellOperatorNP(eps, L))
    nev = wl.C2["nev"]
    w = sla.eigh((M + M.conj().T) / 2, eigvals_only=True, subset_by_index=[0, nev + 1],
                 check_finite=False)
    # the two structurally zero Γ slots are the two lowest (exact zero) eigenvalues
    assert abs(w[0]) < 1e-10 and abs(w[1]) < 1e-10, w[:3]
    return {"c2dense": (np.asarray(w[2:nev + 2]), dict(n_points=wl.C2["n_points"], grid=G, nev=nev,
                                                        k=(0.0, 0.0, 0.0), solver="dense FP64"))}


def case_c3():
    probs = wl.c3_problems()
    order = sorted(range(len(probs)), key=lambda j: (probs[j][1], j))[:2]
    eps, L = wl.network(wl.C3["n_points"], wl.C3["grid"])
    out = {}
    for j in order:
        k, width, nev = probs[j]
        vals, _ = _lobpcg(eps, L, k, nev, width, wl.C3["guard"])
        out[f"c3_k{j}"] = (vals, dict(n_points=wl.C3["n_points"], grid=wl.C3["grid"], k=k,
                                      width=width, nev=nev, index=j, solver="oracle LOBPCG",
                                      tol=TOL))
    return out


def case_c3wide():
    """C3's widest k-point and the one whose width is the median of the path
    (first by index): oracle LOBPCG at their own widths, guard 4."""
    probs = wl.c3_problems()
    widths = sorted(w for _, w, _ in probs)
    med = widths[len(widths) // 2]
    wide = max(range(len(probs)), key=lambda j: (probs[j][1], -j))
    mid = min(range(len(probs)), key=lambda j: (abs(probs[j][1] - med), j))
    eps, L = wl.network(wl.C3["n_points"], wl.C3["grid"])
    out = {}
    for j in (mid, wide):
        k, width, nev = probs[j]
        t0 = time.perf_counter()
        vals, _ = _lobpcg(eps, L, k, nev, width, wl.C3["guard"])
        out[f"c3_k{j}"] = (vals, dict(n_points=wl.C3["n_points"], grid=wl.C3["grid"], k=k,
                                      width=width, nev=nev, index=j, solver="oracle LOBPCG",
                                      tol=TOL))
        print(f"c3_k{j}: width {width}, nev {nev} in {time.perf_counter() - t0:.0f} s", flush=True)
    return out


def case_srv():
    out = {}
    for sp in wl.serving_oracle_subset(wl.serving_sample()):
        eps, L = wl.network(sp.n_points, sp.grid)
        vals, _ = _lobpcg(eps, L, sp.k_frac, sp.nev, wl.SERVING_SOLVER["m"],
                          wl.SERVING_SOLVER["guard"])
        out[f"srv_{sp.rid}"] = (vals, dict(rid=sp.rid, n_points=sp.n_points, grid=sp.grid,
                                           k=sp.k_frac, nev=sp.nev, solver="oracle LOBPCG",
                                           tol=TOL))
    return out


def c5_golden_specs():
    """One request per grid 48³-128³ from the C5 stream (2000 requests, seed 0):
    the smallest nev, first by request id."""
    specs = wl.request_stream(2000, 0, rate=None)
    out = []
    for G in (48, 64, 96, 128):
        c = sorted((s for s in specs if s.grid == G), key=lambda s: (s.nev, s.rid))
        out.append(c[0])
    return out


def case_c5():
    out = {}
    for sp in c5_golden_specs():
        eps, L = wl.network(sp.n_points, sp.grid)
        t0 = time.perf_counter()
        vals, st = _lobpcg(eps, L, sp.k_frac, sp.nev, 16, 8, seed=sp.seed)
        out[f"c5_{sp.rid}"] = (vals, dict(rid=sp.rid, n_points=sp.n_points, grid=sp.grid, k=sp.k_frac,
                                          nev=sp.nev, solver="oracle LOBPCG", tol=TOL, block=16,
                                          guard=8, stream="request_stream(2000, seed=0)"))
        print(f"c5_{sp.rid}: {sp.grid}^3 nev {sp.nev} in {time.perf_counter() - t0:.0f} s", flush=True)
    return out


CASES = {"c1": case_c1, "c2dense": case_c2dense, "c3": case_c3, "c3wide": case_c3wide, "srv": case_srv,
         "c5": case_c5}


def main(names):
    set_workers(os.cpu_count() or 1)
    data = {}
    if OUT.exists():
        with np.load(OUT, allow_pickle=False) as z:
            data = {k: z[k] for k in z.files}
    for name in names:
        t0 = time.perf_counter()
        for key, (vals, meta) in CASES[name]().items():
            data[key] = vals
            data[key + "__meta"] = np.array(json.dumps(meta))
            print(f"{key}: {len(vals)} eigenvalues, λ₁ = {vals[0]:.10f}", flush=True)
        print(f"[{name}] {time.perf_counter() - t0:.0f} s", flush=True)
        np.savez_compressed(OUT, **data)


if __name__ == "__main__":
    main(sys.argv[1:] or list(CASES))
