"""Solve-level measurements of the five workloads on one GPU (one JSON line each).

    python -m eigenfns_amd.solve_bench c1             plain LOBPCG, 32³
    python -m eigenfns_amd.solve_bench c2             plain LOBPCG, 128³, 128 bands (+ 16³ vs dense FP64)
    python -m eigenfns_amd.solve_bench c3 [--n-k N]   64 k-points: one ragged engine, width buckets, one at a time
    python -m eigenfns_amd.solve_bench c4kpm          KPM moments at 256³, degree 12000, 16 probes
    python -m eigenfns_amd.solve_bench c4cheb         ChebSI build stage at 192³ (m = 104, 3000 × 2)
    python -m eigenfns_amd.solve_bench c5 [--n N --rate R]   Poisson request stream, batched vs plain

Eigenvalue accuracy is reported against the FP64 values in
tests/golden/solve_golden.npz where that file covers the case.
"""
from __future__ import annotations

import argparse
import contextlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np
import torch

from . import workloads as wl

GOLDEN = Path(__file__).resolve().parents[1] / "tests" / "golden" / "solve_golden.npz"


def _golden(key):
    if not GOLDEN.exists():
        return None
    with np.load(GOLDEN, allow_pickle=False) as z:
        return z[key] if key in z.files else None


def _rel_err(vals, ref):
    if ref is None:
        return None
    n = min(len(vals), len(ref))
    return float(np.max(np.abs(np.asarray(vals[:n]) - ref[:n]) / np.abs(ref[:n])))


def _progress(msg):
    print(f"[solve_bench] {msg}", file=sys.stderr, flush=True)


def _sync_time(fn):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0


def run_c1(ortho: str = "cholqr", graph: bool = False):
    from .operator import MaxwellOperator
    from .solver import lobpcg_blocks
    c = wl.C1
    eps, L = wl.network(c["n_points"], c["grid"])
    op = MaxwellOperator(eps, L)
    lobpcg_blocks(op, 4, m=8, guard=4, tol=1e-3, verbose=False, ortho=ortho, graph=graph)  # warm
    (vals, _, st), t = _sync_time(lambda: lobpcg_blocks(op, c["nev"], m=c["m"], guard=c["guard"],
                                                        tol=c["tol"], seed=0, verbose=False,
                                                        ortho=ortho, graph=graph))
    return {"case": "C1", "grid": c["grid"], "nev": c["nev"], "m": c["m"], "guard": c["guard"],
            "tol": c["tol"], "ortho": ortho, "graph": graph,
            "small_dense": "device" if graph else os.environ.get("EFN_SMALL_DENSE", "host"),
            "gpu_seconds": round(t, 3), "iterations": st.iterations,
            "theta_applications": st.theta_applications,
            "blocks": [{k: r[k] for k in ("iters", "max_res_locked", "converged", "stagnated")}
                       for r in st.rounds],
            "max_rel_err_vs_fp64_oracle": _rel_err(vals, _golden("c1")), "vals": vals.tolist()}


def run_c2(ortho: str = "cholqr", graph: bool = False):
    from .operator import MaxwellOperator
    from .solver import lobpcg_blocks
    c = wl.C2
    res = {"case": "C2", "grid": c["grid"], "nev": c["nev"], "m": c["m"], "guard": c["guard"],
           "tol": c["tol"], "ortho": ortho, "graph": graph,
           "small_dense": "device" if graph else os.environ.get("EFN_SMALL_DENSE", "host")}
    # parity: the same network at 16³ against the dense FP64 solve
    eps16, L = wl.network(c["n_points"], c["dense_grid"])
    op16 = MaxwellOperator(eps16, L)
    vals16, _, _ = lobpcg_blocks(op16, c["nev"], m=c["m"], guard=c["guard"], tol=c["tol"],
                                 verbose=False, ortho=ortho, graph=graph)
    res["dense16_max_rel_err"] = _rel_err(vals16, _golden("c2dense"))
    del op16
    eps, L = wl.network(c["n_points"], c["grid"])
    op = MaxwellOperator(eps, L)
    with contextlib.redirect_stdout(sys.stderr):    # per-block progress lines
        (vals, _, st), t = _sync_time(lambda: lobpcg_blocks(
            op, c["nev"], m=c["m"], guard=c["guard"], tol=c["tol"], seed=0, verbose=True,
            ortho=ortho, graph=graph))
    res.update({"gpu_seconds": round(t, 2), "iterations": st.iterations,
                "theta_applications": st.theta_applications,
                "lambda_1_128": [float(vals[0]), float(vals[-1])]})
    return res


def run_c3(n_k: int | None = None, plain: bool = True, bucket_step: int = 4, mode: str = "both",
           ragged_slots: int = 64, ragged_engines: int = 2):
    """Every k-point is a request.  "ragged": one engine holds every k-point
    at its own width (guard 4, as the plain solver) and each kernel of an
    iteration is one ragged launch over all of them.  "buckets": widths are
    rounded up to a multiple of `bucket_step` and each bucket is one engine
    (own host thread and stream) whose guard is the bucket minus its largest
    nev (≥ 4), so each request locks its nev = width − 4 bands in one block.
    The plain baseline solves each k-point alone with block = width, guard 4."""
    from .batched import BatchedLOBPCG, Request
    from .operator import MaxwellOperator
    from .solver import lobpcg_blocks
    c = wl.C3
    probs = wl.c3_problems(n_k)
    eps, L = wl.network(c["n_points"], c["grid"])
    buckets = {}
    for j, (k, w, nev) in enumerate(probs):
        b = bucket_step * ((w + bucket_step - 1) // bucket_step)
        buckets.setdefault(b, []).append((j, k, w, nev))

    def batched():
        import threading
        done, errors = {}, []

        def worker(b, items):
            try:
                with torch.cuda.stream(torch.cuda.Stream()):
                    # guard chosen so that one block locks every request's bands
                    guard = b - max(nev for _, _, _, nev in items)
                    eng = BatchedLOBPCG(c["grid"], b, max(guard, 4), min(len(items), 32),
                                        max(nev for *_, nev in items), c["tol"], 400)
                    pending = list(items)
                    while pending or eng.active():
                        while pending and eng.free_slots():
                            j, k, w, nev = pending.pop(0)
                            eng.admit(Request(MaxwellOperator(eps, L, k, max_chunk=-1), nev,
                                              seed=j, rid=j))
                        for req in eng.step():
                            done[req.rid] = req.vals
                            req.op = None
                            _progress(f"C3 batched: k-point {req.rid} done "
                                      f"({len(done)}/{len(probs)})")
                    torch.cuda.current_stream().synchronize()
            except BaseException as exc:
                errors.append(exc)

        from .linalg import one_blas_thread
        ths = [threading.Thread(target=worker, args=(b, it)) for b, it in buckets.items()]
        with one_blas_thread():
            for th in ths:
                th.start()
            for th in ths:
                th.join()
        if errors:
            raise errors[0]
        return done

    def ragged():
        # every k-point at its own width and the plain solver's guard, all of
        # an engine's k-points in the same ragged launches; `ragged_engines`
        # engines (own host thread and stream, k-points dealt round-robin in
        # width order) so one engine's host phases overlap another's kernels
        import threading
        done, steps, errors = {}, [], []
        order = sorted(range(len(probs)), key=lambda j: probs[j][1])
        parts = [order[e::ragged_engines] for e in range(ragged_engines)]

        def worker(items):
            try:
                with torch.cuda.stream(torch.cuda.Stream()):
                    wmax = max(probs[j][1] for j in items)
                    eng = BatchedLOBPCG(c["grid"], wmax, c["guard"], min(len(items), ragged_slots),
                                        max(probs[j][2] for j in items), c["tol"], 400)
                    pending = list(items)
                    while pending or eng.active():
                        while pending and eng.free_slots():
                            j = pending.pop(0)
                            k, w, nev = probs[j]
                            eng.admit(Request(MaxwellOperator(eps, L, k, max_chunk=-1), nev, seed=j, rid=j,
                                              m=w, guard=c["guard"]))
                        for req in eng.step():
                            done[req.rid] = req.vals
                            req.op = None
                            _progress(f"C3 ragged: k-point {req.rid} done ({len(done)}/{len(probs)})")
                    torch.cuda.current_stream().synchronize()
                    steps.append(eng.steps)
            except BaseException as exc:
                errors.append(exc)

        ths = [threading.Thread(target=worker, args=(it,)) for it in parts if it]
        with one_blas_thread():
            for th in ths:
                th.start()
            for th in ths:
                th.join()
        if errors:
            raise errors[0]
        return done, steps

    def warm_up():
        # untimed: the process's first engine pays one-time costs (library and
        # kernel loading, allocator growth, the host pool) that would otherwise
        # land on whichever mode is timed first
        k, w, nev = probs[0]
        eng = BatchedLOBPCG(c["grid"], w, c["guard"], 1, nev, c["tol"], 3)
        eng.admit(Request(MaxwellOperator(eps, L, k, max_chunk=-1), nev, seed=0, rid=0, m=w,
                          guard=c["guard"]))
        for _ in range(8):
            if not eng.active():
                break
            eng.step()
        torch.cuda.synchronize()

    from .linalg import one_blas_thread
    with one_blas_thread():
        warm_up()
    res = {"case": "C3", "grid": c["grid"], "k_points": len(probs)}
    errs = []
    if mode in ("ragged", "both"):
        (done_r, steps), tr = _sync_time(ragged)
        res.update({"ragged_seconds": round(tr, 2), "ragged_k_per_s": round(len(probs) / tr, 3),
                    "ragged_engines": ragged_engines, "ragged_engine_steps": steps})
        for j in range(len(probs)):
            g = _golden(f"c3_k{j}")
            if g is not None and j in done_r:
                errs.append(_rel_err(done_r[j], g))
        done = done_r
        torch.cuda.empty_cache()
    if mode in ("buckets", "both"):
        done, tb = _sync_time(batched)
        res.update({"buckets": {b: len(v) for b, v in buckets.items()},
                    "batched_seconds": round(tb, 2), "batched_k_per_s": round(len(probs) / tb, 3)})
        for j in range(len(probs)):
            g = _golden(f"c3_k{j}")
            if g is not None and j in done:
                errs.append(_rel_err(done[j], g))
        if mode == "both":
            res["max_rel_diff_ragged_vs_buckets"] = max(_rel_err(done_r[j], done[j]) for j in done)
        torch.cuda.empty_cache()
    tb = res.get("batched_seconds") or res.get("ragged_seconds")
    res["max_rel_err_vs_fp64_oracle"] = max(errs) if errs else None
    res["oracle_checked"] = len(errs)
    if plain:
        def one_at_a_time():
            out = {}
            for j, k, w, nev in sorted((j, k, w, nev) for it in buckets.values()
                                       for j, k, w, nev in it):
                op = MaxwellOperator(eps, L, k)
                out[j] = lobpcg_blocks(op, nev, m=w, guard=c["guard"], tol=c["tol"], seed=j,
                                       verbose=False)[0]
                del op
                _progress(f"C3 plain: k-point {j} done ({len(out)}/{len(probs)})")
            return out
        pl, tp = _sync_time(one_at_a_time)
        res.update({"plain_seconds": round(tp, 2), "plain_k_per_s": round(len(probs) / tp, 3),
                    "speedup": round(tp / tb, 2),
                    "speedup_ragged": round(tp / res["ragged_seconds"], 2) if "ragged_seconds" in res else None,
                    "max_rel_diff_batched_vs_plain": max(_rel_err(done[j], pl[j]) for j in pl)})
    return res


def run_c4kpm(degree: int | None = None):
    from .kpm import kpm_moments
    from .operator import MaxwellOperator
    from .chebyshev import lanczos_lambda_max
    c = wl.C4
    degree = degree or c["kpm_degree"]
    G = c["kpm_grid"]
    eps, L = wl.network(c["n_points"], G)
    op = MaxwellOperator(eps, L)
    lam_max = 1.05 * lanczos_lambda_max(op)
    kpm_moments(op, 20, 2, seed=c["kpm_seed"], lam_max=lam_max)  # warm
    with contextlib.redirect_stdout(sys.stderr):    # a line every 1000 degrees
        mom, t = _sync_time(lambda: kpm_moments(op, degree, c["kpm_probes"], seed=c["kpm_seed"],
                                                lam_max=lam_max, verbose=True))
    n_theta = c["kpm_probes"] * (degree // 2)
    per_probe_degree = (48 * G ** 3 + 4 * G ** 3 / c["kpm_probes"])
    return {"case": "C4-KPM", "grid": G, "degree": degree, "probes": c["kpm_probes"],
            "lam_max": lam_max, "seconds": round(t, 2), "theta_applications": n_theta,
            "probe_degrees_per_s": round(n_theta / t, 1),
            "frac_hbm_fused_step": round(per_probe_degree * n_theta / t / 8e12, 4),
            "mu0_over_2G3m2": float(np.mean(mom[:, 0]) / (2 * G ** 3 - 2))}


def run_c4cheb(outers: int | None = None, degree: int | None = None):
    from .chebyshev import lanczos_lambda_max
    from .interior import two_stage
    from .operator import MaxwellOperator
    c = wl.C4
    G = c["cheb_grid"]
    degree = degree or c["build"][0]
    outers = outers or c["build"][1]
    eps, L = wl.network(c["n_points"], G)
    op = MaxwellOperator(eps, L)
    lam_max = 1.05 * lanczos_lambda_max(op)
    lo, hi = c["window"]
    with contextlib.redirect_stdout(sys.stderr):
        (vals, _, rn, info), t = _sync_time(lambda: two_stage(
            op, lo, hi, lam_max, m=c["cheb_m"], build_degree=degree, build_outers=outers,
            polish_degree=c["polish"][0], polish_outers=0, verbose=True))
    hist = info["history"]
    vdeg = c["cheb_m"] * degree * outers
    filt_s = sum(h["outer_seconds"] for h in hist)
    full = c["cheb_m"] * (c["build"][0] * c["build"][1] + c["polish"][0] * c["polish"][1])
    return {"case": "C4-ChebSI build", "grid": G, "m": c["cheb_m"], "degree": degree,
            "outers": outers, "lam_max": lam_max, "seconds": round(t, 1),
            "vector_degrees_per_s": round(vdeg / filt_s, 1), "history": hist,
            "in_window_converged": int(len(vals)),
            "projected_full_protocol_s": round(full / (vdeg / filt_s), 0)}


def run_c5(n: int, rate: float | None, seed: int = 0, plain: bool = True, grids=None,
           engine_voxels=None, start: int = 0, total: int | None = None):
    """`n` requests of the seeded stream of `total` (default n) requests,
    starting at request `start`; arrival times keep their spacing, shifted so
    the slice's first request arrives at t = 0."""
    from .serving import StructureCache, serve_batched, serve_plain
    specs = wl.request_stream(max(total or n, start + n), seed, rate=rate,
                              grids=grids or wl.C5["grids"])[start:start + n]
    t_first = specs[0].t_arrive
    for sp in specs:
        sp.t_arrive -= t_first
    cache = StructureCache()
    for sp in specs:
        cache.get(sp.n_points, sp.grid)
    res, summ = serve_batched(specs, cache, max_slots=32, progress=True,
                              engine_voxels=engine_voxels)
    out = {"case": "C5", "requests": n, "first_request": start, "stream_length": max(total or n, start + n),
           "rate": rate, "grids": list(grids or wl.C5["grids"]), "engine_voxels": engine_voxels,
           "batched": summ}
    if plain:
        pres, psumm = serve_plain(specs, cache, progress=True)
        by = {r.rid: r.vals for r in res}
        out["plain"] = psumm
        out["speedup"] = summ["solves_per_s"] / psumm["solves_per_s"]
        out["max_rel_diff_batched_vs_plain"] = max(_rel_err(by[p["rid"]], p["vals"]) for p in pres)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("case", choices=["c1", "c2", "c3", "c4kpm", "c4cheb", "c5"])
    ap.add_argument("--n-k", type=int, default=None)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--bucket-step", type=int, default=4)
    ap.add_argument("--c3-mode", default="both", choices=["ragged", "buckets", "both"],
                    help="C3: one ragged engine, width buckets, or both")
    ap.add_argument("--slots", type=int, default=64, help="C3: slots per ragged engine")
    ap.add_argument("--engines", type=int, default=2, help="C3: ragged engines (threads and streams)")
    ap.add_argument("--degree", type=int, default=None)
    ap.add_argument("--outers", type=int, default=None)
    ap.add_argument("--n", type=int, default=200)
    ap.add_argument("--rate", type=float, default=None)
    ap.add_argument("--grids", default=None)
    ap.add_argument("--start", type=int, default=0, help="C5: first request of the slice")
    ap.add_argument("--ortho", default="cholqr", choices=["svqb", "cholqr"],
                    help="C1/C2: orthonormalisation of the LOBPCG blocks")
    ap.add_argument("--total", type=int, default=None, help="C5: length of the seeded stream")
    ap.add_argument("--graph", action="store_true",
                    help="C1/C2: steady-state iterations as replayed HIP graphs (device eigensolves)")
    ap.add_argument("--engine-voxels", type=float, default=None,
                    help="C5: cap slots x G^3 per engine (latency setting)")
    args = ap.parse_args(argv)
    if args.case == "c1":
        out = run_c1(args.ortho, args.graph)
    elif args.case == "c2":
        out = run_c2(args.ortho, args.graph)
    elif args.case == "c3":
        out = run_c3(args.n_k, plain=not args.no_plain, bucket_step=args.bucket_step, mode=args.c3_mode,
                     ragged_slots=args.slots, ragged_engines=args.engines)
    elif args.case == "c4kpm":
        out = run_c4kpm(args.degree)
    elif args.case == "c4cheb":
        out = run_c4cheb(args.outers, args.degree)
    else:
        grids = tuple(int(g) for g in args.grids.split(",")) if args.grids else None
        out = run_c5(args.n, args.rate, plain=not args.no_plain, grids=grids,
                     engine_voxels=args.engine_voxels, start=args.start, total=args.total)
    print(json.dumps(out, default=float), flush=True)


if __name__ == "__main__":
    main()
