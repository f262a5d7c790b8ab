"""Launch-shape autotuner for Θ: pass structure and workspace chunk per grid.

The library compiles two forms of Θ (and of the preconditioner and the fused
Chebyshev / KPM steps, which share its passes): three passes on the grids
that have the plane pass, five everywhere.  Each launch set runs `chunk`
vectors through one workspace.  Which (passes, chunk) pair is fastest depends
on the grid size; this module times the candidates on a synthetic operator
(a 32-vector block by default) and keeps the winner per grid in a table:

  * the committed table ``theta_tuning.json`` beside this module, measured on
    MI355X with ``python -m eigenfns_amd.autotune --write-package``;
  * a user table (``$EFN_TUNING_CACHE``, default
    ``~/.cache/eigenfns_amd/theta_tuning.json``), which first-use tuning fills
    when ``EFN_AUTOTUNE=1``: a grid missing from both tables is then timed the
    first time an operator of that size is created.

``MaxwellOperator(max_chunk=0)`` takes its chunk and pass structure from
``choice(G)``; an explicit ``max_chunk`` or ``passes`` argument, or a forced
process-wide engine (``operator.theta_engine(3 or 5)``), overrides it.
No reference counterpart: XLA schedules the reference's FFTs itself.
"""
from __future__ import annotations

import json
import os
import threading
import time
from pathlib import Path

import numpy as np

PACKAGE_TABLE = Path(__file__).resolve().parent / "theta_tuning.json"
ARCH = "gfx950"
_lock = threading.Lock()
_cache: dict | None = None


def user_table_path() -> Path:
    env = os.environ.get("EFN_TUNING_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "eigenfns_amd" / "theta_tuning.json"


def _read(path: Path) -> dict:
    try:
        data = json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return {}
    return data.get(ARCH, {}) if isinstance(data, dict) else {}


def _tables() -> dict:
    global _cache
    with _lock:
        if _cache is None:
            merged = dict(_read(PACKAGE_TABLE))
            merged.update(_read(user_table_path()))   # the user's measurements win
            _cache = merged
        return _cache


def reload() -> None:
    """Forget the cached tables (they are read again on the next lookup)."""
    global _cache
    with _lock:
        _cache = None


def default_chunk(grid_size: int) -> int:
    """The library's own default: the three-component workspace within 192 MB."""
    per = 3 * grid_size ** 3 * 8
    return int(min(1024, max(1, (192 << 20) // per)))


def candidates(grid_size: int, max_workspace: float = 1.1e9):
    """(passes, chunk) pairs worth timing at this grid size."""
    from . import _lib
    G = int(grid_size)
    per = 3 * G ** 3 * 8
    c0 = default_chunk(G)
    chunks = sorted({c for c in (1, max(1, c0 // 2), c0, 2 * c0, 4 * c0) if c * per <= max_workspace}
                    | {c0})
    passes = [5]
    if _has_three_pass(G, _lib.load()):
        passes = [3, 5]
    return [(p, c) for p in passes for c in chunks]


def _has_three_pass(G: int, lib) -> bool:
    return lib.efn_theta_three_pass(int(G)) == 1


def tune(grid_size: int, m: int = 32, reps: int = 5, warmup: int = 2, cands=None,
         verbose: bool = False) -> dict:
    """Time Θ on an m-vector block for every candidate; returns the winner's
    entry {passes, chunk, vector_applies_per_s, m, timed: [...]}."""
    import torch
    from .operator import MaxwellOperator, theta_engine
    G = int(grid_size)
    rng = np.random.default_rng(G)
    eps = np.where(rng.random((G, G, G)) < 0.3, 8.41, 1.0).astype(np.float32)
    prev = theta_engine(0)    # per-context choices apply only under the automatic engine
    timed = []
    try:
        X = None
        for passes, chunk in (cands or candidates(G)):
            op = MaxwellOperator(eps, 0.13 * G, max_chunk=chunk, passes=passes, tuned=False)
            if X is None:
                X = op.random_block(m, seed=1)
                Y = torch.empty_like(X)
            for _ in range(warmup):
                op.theta_into(X, Y)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                op.theta_into(X, Y)
            torch.cuda.synchronize()
            rate = m * reps / (time.perf_counter() - t0)
            timed.append({"passes": op.passes, "chunk": op.chunk, "vector_applies_per_s": round(rate, 1)})
            if verbose:
                print(f"[autotune] G={G} passes={op.passes} chunk={op.chunk}: {rate:,.0f} vector-applies/s",
                      flush=True)
            del op
        del X, Y
        torch.cuda.empty_cache()
    finally:
        theta_engine(prev)
    return dict(pick(timed), m=m, timed=timed)


def pick(timed, within: float = 0.02) -> dict:
    """The smallest workspace among the candidates within `within` of the
    fastest (a bigger chunk is kept only for a gain above timing noise)."""
    top = max(e["vector_applies_per_s"] for e in timed)
    close = [e for e in timed if e["vector_applies_per_s"] >= (1 - within) * top]
    best = min(close, key=lambda e: (e["chunk"], -e["vector_applies_per_s"]))
    return {k: best[k] for k in ("passes", "chunk", "vector_applies_per_s")}


def _write(path: Path, G: int, entry: dict) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    try:
        data = json.loads(path.read_text())
    except (OSError, ValueError):
        data = {}
    data.setdefault(ARCH, {})[str(int(G))] = entry
    tmp = path.with_name(path.name + ".tmp")
    tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
    tmp.replace(path)


def choice(grid_size: int) -> dict | None:
    """The table's (passes, chunk) for this grid, tuning it first when
    EFN_AUTOTUNE=1 and no table has it; None when there is no entry."""
    G = str(int(grid_size))
    entry = _tables().get(G)
    if entry is None and os.environ.get("EFN_AUTOTUNE") == "1":
        entry = tune(int(grid_size))
        _write(user_table_path(), int(grid_size), entry)
        reload()
    return entry


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="time Θ launch shapes and record the winners")
    ap.add_argument("grids", nargs="*", type=int, default=[32, 48, 64, 96, 128, 192, 256])
    ap.add_argument("--m", type=int, default=32)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--write-package", action="store_true",
                    help="record into the package table (theta_tuning.json) instead of the user table")
    args = ap.parse_args(argv)
    path = PACKAGE_TABLE if args.write_package else user_table_path()
    for G in args.grids:
        entry = tune(G, m=args.m, reps=args.reps, verbose=True)
        _write(path, G, entry)
        print(json.dumps({"grid": G, "passes": entry["passes"], "chunk": entry["chunk"],
                          "vector_applies_per_s": entry["vector_applies_per_s"]}), flush=True)
    reload()


if __name__ == "__main__":
    main()
