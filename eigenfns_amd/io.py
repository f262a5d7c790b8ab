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


def _atomic_save_npy(path: Path, arr) -> None:
    tmp = path.with_suffix(".tmp.npy")
    np.save(tmp, _host(arr))
    tmp.replace(path)


class BlockCheckpointer:
    """Persist locked eigenpairs block by block; resume a partial solve (io.py:19-67)."""

    def __init__(self, directory, tag: str, meta: dict | None = None):
        self.dir = Path(directory)
        self.dir.mkdir(parents=True, exist_ok=True)
        self.tag = tag
        self.meta = dict(meta or {})

    def _meta_path(self) -> Path:
        return self.dir / f"{self.tag}_meta.json"

    def _block_paths(self, i: int) -> tuple[Path, Path]:
        return (self.dir / f"{self.tag}_block{i:03d}_vals.npy",
                self.dir / f"{self.tag}_block{i:03d}_vecs.npy")

    def load(self):
        """(vals, vecs, n_blocks, meta); vals/vecs None when there is no checkpoint.
        Raises RuntimeError if the metadata names a block whose files are missing."""
        pm = self._meta_path()
        if not pm.exists():
            return None, None, 0, self.meta
        meta = json.loads(pm.read_text())
        n_blocks = int(meta.get("n_blocks", 0))
        vals, vecs = [], []
        for i in range(n_blocks):
            pv, px = self._block_paths(i)
            if not (pv.exists() and px.exists()):
                raise RuntimeError(f"checkpoint corrupt: missing block {i} files")
            vals.append(np.load(pv, allow_pickle=False))
            vecs.append(np.load(px, allow_pickle=False))
        if not vals:
            return None, None, 0, meta
        return np.concatenate(vals), np.concatenate(vecs), n_blocks, meta

    def save_block(self, i: int, vals, vecs, extra: dict | None = None) -> None:
        """Append block i: the two arrays first, then the metadata that counts it."""
        pv, px = self._block_paths(i)
        _atomic_save_npy(pv, vals)
        _atomic_save_npy(px, vecs)
        meta = dict(self.meta)
        meta.update(extra or {})
        meta["n_blocks"] = i + 1
        meta["saved_at"] = time.strftime("%Y-%m-%d %H:%M:%S")
        tmp = self._meta_path().with_suffix(".tmp.json")
        tmp.write_text(json.dumps(meta, indent=1))
        tmp.replace(self._meta_path())


def save_interior_state(path, X, stage: str, outer: int, n_theta: int, lam_max: float,
                        history: list) -> None:
    """The ChebSI basis and protocol position (run_interior.py:116-121)."""
    path = Path(path)
    tmp = path.with_name(path.stem + ".tmp.npz")
    np.savez(tmp, X=_host(X), stage=np.array(stage), outer=np.array(int(outer)),
             n_theta=np.array(int(n_theta)), lam_max=np.array(float(lam_max)),
             history=np.array(json.dumps(history)))
    tmp.replace(path)


def load_interior_state(path):
    """dict(X (host), stage, outer, n_theta, lam_max, history) or None if absent."""
    path = Path(path)
    if not path.exists():
        return None
    with np.load(path, allow_pickle=False) as st:
        return {"X": st["X"], "stage": str(st["stage"]), "outer": int(st["outer"]),
                "n_theta": int(st["n_theta"]), "lam_max": float(st["lam_max"]),
                "history": json.loads(str(st["history"]))}
