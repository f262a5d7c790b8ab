"""Continuous-batching request loop for mixed (structure, k-point, band-count) eigen-solves.

A request asks for the lowest `nev` bands of one synthetic rod network
(N ∈ {64, 216, 1000}) on one grid (32³…128³) at one Bloch k.  Requests are
admitted, in arrival order, into the batched LOBPCG engine of their grid size
(`eigenfns_amd.batched.BatchedLOBPCG`) as soon as a slot is free; each engine's
slot count is fixed by an HBM budget (per-slot footprint = the engine's fixed
buffers).  ε per structure stays resident (one ε table per request context,
shared rasterisation per structure) and locked vectors stay in HBM.  Every
loop turn advances every non-empty engine by one LOBPCG iteration, so requests
that arrive while others run join the next turn (continuous batching).

The baseline is the plain GPU path: the same requests solved one at a time, in
arrival order, with `solver.lobpcg_blocks` on the same kernels.
"""
from __future__ import annotations

import sys
import time

import numpy as np

from .batched import BatchedLOBPCG, Request
from .linalg import one_blas_thread
from .operator import MaxwellOperator
from .workloads import NET_SEEDS, Spec, network, request_stream  # noqa: F401 (re-exported)


class StructureCache:
    """ε grids rasterised once per (network, grid); safe to share between threads."""

    def __init__(self):
        import threading
        self._eps = {}
        self._lock = threading.Lock()

    def get(self, n_points: int, grid: int):
        key = (n_points, grid)
        with self._lock:
            if key not in self._eps:
                self._eps[key] = network(n_points, grid)
            return self._eps[key]


def _make_op(cache, sp, workspace: bool):
    eps, L = cache.get(sp.n_points, sp.grid)
    return MaxwellOperator(eps, L, sp.k_frac, max_chunk=0 if workspace else -1)


def _stats(lat):
    lat = np.asarray(lat)
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}


def _slot_bytes(grid, m, guard, max_nev):
    """Device bytes one engine slot holds: [X; W; P] and Θ of it, R, the guard
    carry, the locked buffer and, when combines cannot run in place, two
    temporaries (BatchedLOBPCG)."""
    from .linalg import COMBINE_IN_PLACE_MAX
    cap = int(np.ceil(max_nev / (m - guard))) * (m - guard) + m
    temps = 2 * m if 2 * m > COMBINE_IN_PLACE_MAX else 0
    return (2 * 3 * m + 2 * m + temps + cap) * 2 * grid ** 3 * 8


def plan_slots(grids, m, guard, max_nev, max_slots, hbm_budget, engine_voxels=None):
    """Slots per grid under one HBM budget: every grid gets the same count
    first (slot bytes grow as G³, so equal counts spend the budget in
    proportion to each grid's work per request), then what is left goes to the
    cheapest grids, up to `max_slots` each.

    engine_voxels (optional) caps slots·G³ per engine, i.e. the work of one
    lock-step iteration: fewer large-grid requests run at once, each finishing
    sooner, while the rest wait in the FIFO — a latency (p99) rather than a
    throughput setting."""
    grids = sorted(set(grids))
    cap = {g: max_slots if engine_voxels is None else
           int(min(max_slots, max(1, engine_voxels // g ** 3))) for g in grids}
    cost = {g: _slot_bytes(g, m, guard, max_nev) for g in grids}
    base = int(min(max_slots, max(1, hbm_budget // sum(cost.values()))))
    slots = {g: min(base, cap[g]) for g in grids}
    left = hbm_budget - sum(cost[g] * slots[g] for g in grids)
    for g in grids:                                    # smallest grid first
        extra = int(min(cap[g] - slots[g], max(0, left) // cost[g]))
        slots[g] += extra
        left -= extra * cost[g]
    return slots


def _summary(done, wall, turns, engines):
    lat = [r.t_done - r.t_arrive for r in done]
    timing = {}
    for e in engines.values():
        for k, v in e.timing.items():
            timing[k] = round(timing.get(k, 0.0) + v, 3)
    return {"requests": len(done), "wall_s": wall, "solves_per_s": len(done) / wall,
            "turns": turns, "slots": {g: e.nslots for g, e in engines.items()},
            "engine_steps": {g: e.steps for g, e in engines.items()},
            "host_phase_s": timing, **_stats(lat)}


def engines_for(grids, engines_per_grid="auto") -> dict:
    """Engines per grid size: "auto" gives the largest grid two and every other
    grid one; an int applies to every grid; a dict (grid → count, keys may be
    strings) to the grids it names, one elsewhere.  At least one each."""
    grids = sorted(int(g) for g in grids)
    if engines_per_grid == "auto":
        engines_per_grid = {grids[-1]: 2} if grids else 1
    if isinstance(engines_per_grid, dict):
        per = {int(g): max(1, int(v)) for g, v in engines_per_grid.items()}
        return {g: per.get(g, 1) for g in grids}
    return {g: max(1, int(engines_per_grid)) for g in grids}


def serve_batched(specs, cache=None, m=24, guard=8, tol=1e-4, maxit=400, max_slots=32,
                  hbm_budget=200e9, verbose=False, threads=True, progress=False,
                  engines_per_grid="auto", engine_voxels=None, small_dense="host", max_nev=None,
                  ortho="cholqr"):
    """Run the stream through per-grid batched engines.  Returns (results, summary).

    threads=True: each grid's engine runs in its own host thread on its own HIP
    stream (default priority), so one engine's host work (small eigensolves,
    admission) overlaps the others' kernels, and small kernels of different
    engines can share the GPU.  Requests are admitted FIFO per grid.
    engines_per_grid (an int, a dict grid → count, or "auto": two engines
    for the largest grid of the stream, one for each other grid) > 1 splits a
    grid's slots over that many engines
    (threads, streams) pulling from the same queue, so one half's host work
    overlaps the other half's kernels: 2 measured +16 % solves/s over 1 (48
    requests, 32³–64³) while each step's host work was several milliseconds
    of Python per engine; with the slot views cached and the engine step's
    host work down to the eigensolves, two engines for every grid lose to
    the interpreter contention of six engine threads (serving sample 9.5–11.0
    against 11.5–11.8 solves/s with one each), while two for the largest grid
    only — the engine the stream waits for — measured 12.5 ("auto").

    Slots are sized for the service's band limit `max_nev` (default: the
    request stream configuration's upper bound, 64), not for the requests
    seen so far; a request asking for more is rejected (ValueError) before
    anything runs.  `small_dense` and `ortho` choose where the engines' small
    FP64 eigensolves run and how blocks are orthonormalised (BatchedLOBPCG)."""
    cache = cache or StructureCache()
    grids = sorted({sp.grid for sp in specs})
    from .workloads import C5
    max_nev = int(max_nev or C5["nev_range"][1])
    too_many = [sp.rid for sp in specs if sp.nev > max_nev]
    if too_many:
        raise ValueError(f"requests {too_many[:8]} ask for more than max_nev={max_nev} bands")
    plan = plan_slots(grids, m, guard, max_nev, max_slots, hbm_budget, engine_voxels)
    if not threads:
        with one_blas_thread():
            return _serve_sequential(specs, cache, m, guard, tol, maxit, plan, max_nev, verbose,
                                     small_dense, ortho)
    import threading
    import torch
    queues = {g: sorted([sp for sp in specs if sp.grid == g], key=lambda s: s.t_arrive)
              for g in grids}
    engines, done, errors, turns = {}, [], [], [0]
    lock = threading.Lock()
    t0 = time.perf_counter()

    kmap = engines_for(grids, engines_per_grid)

    def worker(g, e):
        try:
            stream = torch.cuda.Stream()
            k = kmap[g]
            with torch.cuda.stream(stream):
                pending = queues[g]
                nslots = max(1, min(-(-plan[g] // k), -(-len(pending) // k)))
                eng = BatchedLOBPCG(g, m, guard, nslots, max_nev, tol, maxit,
                                    small_dense=small_dense, ortho=ortho)
                with lock:
                    engines[str(g) if k == 1 else f"{g}.{e}"] = eng
                while True:
                    now = time.perf_counter() - t0
                    admit = []
                    with lock:  # the grid's queue is shared by its engines
                        while pending and pending[0].t_arrive <= now and len(admit) < eng.free_slots():
                            admit.append(pending.pop(0))
                        left = len(pending)
                        nxt = pending[0].t_arrive if pending else None
                    for sp in admit:
                        req = Request(_make_op(cache, sp, workspace=False), sp.nev, seed=sp.seed,
                                      rid=sp.rid, meta={"grid": sp.grid, "n_points": sp.n_points})
                        req.t_arrive = t0 + sp.t_arrive
                        eng.admit(req)
                    if not eng.active():
                        if not left:
                            break
                        time.sleep(max(0.0, min(0.01, nxt - now)))
                        continue
                    finished = eng.step()
                    for req in finished:
                        req.op = None   # frees the request's context (device-synchronising)
                    with lock:
                        turns[0] += 1
                        done.extend(finished)
                        if finished and progress:
                            print(f"[serve] batched: {len(done)}/{len(specs)} done "
                                  f"({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
                        if verbose and turns[0] % 50 == 0:
                            print(f"  turn {turns[0]}: done {len(done)}/{len(specs)}", flush=True)
                stream.synchronize()
        except BaseException as exc:  # re-raised in the caller
            errors.append(exc)

    workers = [threading.Thread(target=worker, args=(g, e), name=f"engine-{g}.{e}")
               for g in grids for e in range(kmap[g])]
    with one_blas_thread():   # host BLAS at one thread while the engines' pools run
        for w in workers:
            w.start()
        for w in workers:
            w.join()
    if errors:
        raise errors[0]
    return done, _summary(done, time.perf_counter() - t0, turns[0], engines)


def _serve_sequential(specs, cache, m, guard, tol, maxit, plan, max_nev, verbose,
                      small_dense="host", ortho="cholqr"):
    """One host thread, one stream: every loop turn steps each engine in turn."""
    engines = {}
    pending = sorted(specs, key=lambda s: s.t_arrive)
    done = []
    t0 = time.perf_counter()
    turns = 0
    while pending or any(e.active() for e in engines.values()):
        now = time.perf_counter() - t0
        still = []
        for sp in pending:
            if sp.t_arrive > now:
                still.append(sp)
                continue
            eng = engines.get(sp.grid)
            if eng is None:
                eng = engines[sp.grid] = BatchedLOBPCG(sp.grid, m, guard, plan[sp.grid], max_nev,
                                                       tol, maxit, small_dense=small_dense,
                                                       ortho=ortho)
            if eng.free_slots():
                req = Request(_make_op(cache, sp, workspace=False), sp.nev, seed=sp.seed,
                              rid=sp.rid, meta={"grid": sp.grid, "n_points": sp.n_points})
                req.t_arrive = t0 + sp.t_arrive
                eng.admit(req)
            else:
                still.append(sp)
        pending = still
        active = [e for e in engines.values() if e.active()]
        if not active:
            if pending:
                time.sleep(max(0.0, min(0.01, pending[0].t_arrive - now)))
            continue
        for eng in active:
            for req in eng.step():
                done.append(req)
                req.op = None
        turns += 1
        if verbose and turns % 50 == 0:
            print(f"  turn {turns}: done {len(done)}/{len(specs)} "
                  f"active {[e.active() for e in engines.values()]}", flush=True)
    return done, _summary(done, time.perf_counter() - t0, turns, engines)


def serve_plain(specs, cache=None, m=24, guard=8, tol=1e-4, maxit=400, progress=False):
    """The same stream solved one request at a time with lobpcg_blocks (FIFO)."""
    from .solver import lobpcg_blocks
    cache = cache or StructureCache()
    out = []
    t0 = time.perf_counter()
    for sp in sorted(specs, key=lambda s: s.t_arrive):
        now = time.perf_counter() - t0
        if sp.t_arrive > now:
            time.sleep(sp.t_arrive - now)
        op = _make_op(cache, sp, workspace=True)
        vals, _, st = lobpcg_blocks(op, sp.nev, m=m, guard=guard, tol=tol, maxit=maxit,
                                    seed=sp.seed, verbose=False)
        t_done = time.perf_counter() - t0
        out.append({"rid": sp.rid, "vals": vals, "latency": t_done - sp.t_arrive,
                    "iterations": st.iterations})
        if progress:
            print(f"[serve] plain: {len(out)}/{len(specs)} done ({t_done:.1f}s)", file=sys.stderr,
                  flush=True)
        del op
    wall = time.perf_counter() - t0
    summary = {"requests": len(out), "wall_s": wall, "solves_per_s": len(out) / wall,
               **_stats([o["latency"] for o in out])}
    return out, summary


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=16)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--grids", default="32,48,64,96,128")
    ap.add_argument("--nev-max", type=int, default=64)
    ap.add_argument("--plain", action="store_true")
    ap.add_argument("--slots", type=int, default=32)
    ap.add_argument("--sequential", action="store_true", help="one host thread, one stream")
    ap.add_argument("--engines-per-grid", default="auto",
                    help='"auto" (two engines for the largest grid) or an engine count per grid')
    ap.add_argument("--small-dense", choices=["device", "host", "hybrid"], default="host")
    ap.add_argument("--ortho", choices=["svqb", "cholqr"], default="cholqr")
    ap.add_argument("--no-warm", action="store_true", help="no untimed warm-up solve")
    args = ap.parse_args()
    grids = tuple(int(g) for g in args.grids.split(","))
    specs = request_stream(args.n, args.seed, grids=grids, nev_range=(8, args.nev_max))
    cache = StructureCache()
    for sp in specs:   # structures rasterised before the clock starts, as the service's cache would hold
        cache.get(sp.n_points, sp.grid)
    kw = dict(max_slots=args.slots, threads=not args.sequential,
              engines_per_grid=(args.engines_per_grid if args.engines_per_grid == "auto"
                                else int(args.engines_per_grid)),
              small_dense=args.small_dense, ortho=args.ortho)
    warm = [sp for sp in specs if sp.nev <= 20][:2]
    if warm and not args.no_warm:  # library, allocator and thread pools, untimed (two small requests)
        serve_batched(warm, cache, **kw)
    res, summ = serve_batched(specs, cache, verbose=True, **kw)
    print(json.dumps({"batched": summ}), flush=True)
    if args.plain:
        pres, psumm = serve_plain(specs, cache)
        by = {r.rid: r.vals for r in res}
        diff = max(float(np.max(np.abs(by[p["rid"]] - p["vals"]) / np.abs(p["vals"]))) for p in pres)
        print(json.dumps({"plain": psumm, "max_rel_diff": diff,
                          "speedup": summ["solves_per_s"] / psumm["solves_per_s"]}), flush=True)
