"""Checkpoint formats (reference eigenfns/io.py:19-67, run_interior.py:116-121)."""
import json

import numpy as np
import pytest

from eigenfns_amd.io import BlockCheckpointer, load_interior_state, save_interior_state


def test_block_checkpointer_round_trip_and_layout(tmp_path):
    ck = BlockCheckpointer(tmp_path, "run", meta={"grid": 8})
    assert ck.load()[0] is None
    rng = np.random.default_rng(0)
    blocks = [(np.sort(rng.random(3)), (rng.standard_normal((3, 2, 4, 4, 4)) + 0j).astype(np.complex64))
              for _ in range(2)]
    for i, (v, x) in enumerate(blocks):
        ck.save_block(i, v, x, extra={"n_locked": 3 * (i + 1)})
    # the reference's file names and metadata
    assert (tmp_path / "run_block001_vecs.npy").exists() and not list(tmp_path.glob("*.tmp*"))
    meta = json.loads((tmp_path / "run_meta.json").read_text())
    assert meta["n_blocks"] == 2 and meta["grid"] == 8 and meta["n_locked"] == 6
    vals, vecs, n, meta2 = BlockCheckpointer(tmp_path, "run").load()
    assert n == 2 and np.array_equal(vals, np.concatenate([b[0] for b in blocks]))
    assert np.array_equal(vecs, np.concatenate([b[1] for b in blocks]))
    (tmp_path / "run_block000_vals.npy").unlink()
    with pytest.raises(RuntimeError, match="corrupt"):
        ck.load()


def test_interior_state_round_trip(tmp_path):
    X = (np.arange(2 * 2 * 8) + 1j).reshape(2, 2, 2, 2, 2).astype(np.complex64)
    hist = [{"stage": "build", "outer": 0, "in_window": 5}]
    p = tmp_path / "state.npz"
    save_interior_state(p, X, "build", 0, 1234, 56.5, hist)
    st = load_interior_state(p)
    assert np.array_equal(st["X"], X) and st["stage"] == "build" and st["outer"] == 0
    assert st["n_theta"] == 1234 and st["lam_max"] == 56.5 and st["history"] == hist
    assert load_interior_state(tmp_path / "absent.npz") is None
