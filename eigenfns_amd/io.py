"""Checkpoint / resume for long solves (reference eigenfns/io.py:19-67 and the
interior driver's state file, scripts/run_interior.py:94-126).

`BlockCheckpointer` keeps the reference's on-disk format so either code can
resume the other's runs: per locked block ``<tag>_block<iii>_vals.npy`` and
``<tag>_block<iii>_vecs.npy`` (append-only, never rewritten), and
``<tag>_meta.json`` with ``n_blocks``; every file is written to a temporary
name and renamed, so a crash leaves either the old or the new checkpoint.
Vectors are copied device → host once per block.

`save_interior_state` / `load_interior_state` hold the bandpass ChebSI basis
and its protocol position (stage, outer step, Θ count, λ_max, history) in one
``.npz`` written the same way; loading executes nothing from the file.
"""
from __future__ import annotations

import json
import time
from pathlib import Path

import numpy as np


def _host(a):
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _replace_atomically(final: Path, write) -> None:
    """Run write(tmp) on a sibling temporary file, then rename it over `final`
    (a crash leaves the old file or the new one, never a partial one).  The
    temporary name keeps the final suffix, so numpy writes exactly that path."""
    tmp = final.with_name(f"{final.stem}.tmp{final.suffix}")
    write(tmp)
    tmp.replace(final)


class BlockCheckpointer:
    """Locked eigenpairs persisted block by block; a later run resumes from them.

    Protocol of the reference's checkpointer (eigenfns/io.py:19-67), used by
    ``lobpcg_blocks_resumable``: ``load() -> (vals, vecs, n_blocks, meta)``
    and ``save_block(i, vals, vecs, extra=None)``.  On-disk format, shared
    with the reference so either code resumes the other's runs:
    ``<tag>_block<iii>_vals.npy`` / ``<tag>_block<iii>_vecs.npy`` per block
    (never rewritten) and ``<tag>_meta.json`` whose ``n_blocks`` counts the
    blocks that are complete on disk, plus the caller's metadata, the block's
    ``extra`` entries and ``saved_at``.
    """

    def __init__(self, directory, tag: str, meta: dict | None = None):
        self.dir = Path(directory)
        self.dir.mkdir(parents=True, exist_ok=True)
        self.tag = tag
        self.meta = dict(meta or {})

    # the format's file names
    def _file(self, suffix: str) -> Path:
        return self.dir / f"{self.tag}_{suffix}"

    def _block_file(self, i: int, what: str) -> Path:
        return self._file(f"block{i:03d}_{what}.npy")

    def _read_meta(self) -> dict | None:
        path = self._file("meta.json")
        return json.loads(path.read_text()) if path.exists() else None

    def _read_block(self, i: int):
        files = [self._block_file(i, what) for what in ("vals", "vecs")]
        absent = [f.name for f in files if not f.exists()]
        if absent:
            raise RuntimeError(f"checkpoint corrupt: the metadata counts block {i} but "
                               f"{', '.join(absent)} is missing")
        vals, vecs = (np.load(f, allow_pickle=False) for f in files)
        if vals.shape[0] != vecs.shape[0]:
            raise RuntimeError(f"checkpoint corrupt: block {i} holds {vals.shape[0]} eigenvalues "
                               f"and {vecs.shape[0]} vectors")
        return vals, vecs

    def load(self):
        """(vals, vecs, n_blocks, meta) of every saved block in order; vals and vecs
        are None (and n_blocks 0) when nothing has been saved.  RuntimeError when
        a block the metadata counts is missing or inconsistent."""
        meta = self._read_meta()
        if meta is None:
            return None, None, 0, self.meta
        count = int(meta.get("n_blocks", 0))
        if count == 0:
            return None, None, 0, meta
        parts = [self._read_block(i) for i in range(count)]
        return (np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]),
                count, meta)

    def save_block(self, i: int, vals, vecs, extra: dict | None = None) -> None:
        """Write block i's arrays, then the metadata that makes it count (so a crash
        between the two leaves the previous checkpoint intact)."""
        for what, arr in (("vals", vals), ("vecs", vecs)):
            data = _host(arr)
            _replace_atomically(self._block_file(i, what), lambda tmp, d=data: np.save(tmp, d))
        record = {**self.meta, **(extra or {}), "n_blocks": i + 1,
                  "saved_at": time.strftime("%Y-%m-%d %H:%M:%S")}
        text = json.dumps(record, indent=1)
        _replace_atomically(self._file("meta.json"), lambda tmp: tmp.write_text(text))


def save_interior_state(path, X, stage: str, outer: int, n_theta: int, lam_max: float,
                        history: list) -> None:
    """The ChebSI basis and protocol position (run_interior.py:116-121)."""
    fields = dict(X=_host(X), stage=np.array(stage), outer=np.array(int(outer)),
                  n_theta=np.array(int(n_theta)), lam_max=np.array(float(lam_max)),
                  history=np.array(json.dumps(history)))
    _replace_atomically(Path(path), lambda tmp: np.savez(tmp, **fields))


def load_interior_state(path):
    """dict(X (host), stage, outer, n_theta, lam_max, history) or None if absent."""
    path = Path(path)
    if not path.exists():
        return None
    with np.load(path, allow_pickle=False) as st:
        return {"X": st["X"], "stage": str(st["stage"]), "outer": int(st["outer"]),
                "n_theta": int(st["n_theta"]), "lam_max": float(st["lam_max"]),
                "history": json.loads(str(st["history"]))}
