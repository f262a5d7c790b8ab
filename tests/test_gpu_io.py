"""Resume from checkpoints on the GPU: LOBPCG blocks (solver.py:385-411) and the
two-stage ChebSI state (run_interior.py:94-126)."""
import json
import shutil

import numpy as np
import pytest

from oracle.dense_np import dense_eigh
from oracle.operator_np import MaxwellOperatorNP

pytestmark = pytest.mark.gpu


def _eps(G=8, seed=6):
    rng = np.random.default_rng(seed)
    return np.where(rng.random((G, G, G)) < 0.3, 8.41, 1.0).astype(np.float32)


def test_lobpcg_resume_after_a_crash_matches_full_run(tmp_path):
    from eigenfns_amd.io import BlockCheckpointer
    from eigenfns_amd.operator import MaxwellOperator
    from eigenfns_amd.solver import lobpcg_blocks_resumable
    op = MaxwellOperator(_eps(), 1.0)
    w = dense_eigh(MaxwellOperatorNP(_eps(), 1.0))[0]
    full_dir = tmp_path / "full"
    vals, vecs, _ = lobpcg_blocks_resumable(op, 24, BlockCheckpointer(full_dir, "t"), m=16,
                                            guard=8, tol=1e-6, verbose=False)
    assert json.loads((full_dir / "t_meta.json").read_text())["n_blocks"] == 3
    # a run that died after its first locked block
    crash = tmp_path / "crash"
    shutil.copytree(full_dir, crash)
    meta = json.loads((crash / "t_meta.json").read_text())
    meta["n_blocks"] = 1
    (crash / "t_meta.json").write_text(json.dumps(meta))
    rvals, rvecs, _ = lobpcg_blocks_resumable(op, 24, BlockCheckpointer(crash, "t"), resume=True,
                                              m=16, guard=8, tol=1e-6, verbose=False)
    assert np.abs(rvals - w[:24]).max() / w[23] < 1e-6
    assert np.abs(rvals - vals).max() / w[23] < 1e-6
    assert json.loads((crash / "t_meta.json").read_text())["n_blocks"] == 3
    # a finished checkpoint returns its stored values at once
    again, _, _ = lobpcg_blocks_resumable(op, 24, BlockCheckpointer(crash, "t"), resume=True,
                                          m=16, guard=8)
    stored = np.concatenate([np.load(crash / f"t_block{i:03d}_vals.npy") for i in range(3)])
    assert np.array_equal(again, stored[:24])


def test_two_stage_resumes_from_state_file(tmp_path):
    """An interrupted ChebSI run resumed from its state file gives bitwise the
    eigenvalues and history of an uninterrupted run (deterministic kernels,
    exact basis round trip); converged pairs match the dense FP64 spectrum."""
    from eigenfns_amd.chebyshev import lanczos_lambda_max
    from eigenfns_amd.interior import two_stage
    from eigenfns_amd.io import load_interior_state
    from eigenfns_amd.operator import MaxwellOperator
    eps = _eps()
    op = MaxwellOperator(eps, 1.0)
    w = dense_eigh(MaxwellOperatorNP(eps, 1.0))[0]
    lo, hi = 0.5 * (w[9] + w[10]), 0.5 * (w[15] + w[16])         # 6 low states
    lam_max = 1.05 * lanczos_lambda_max(op)
    kw = dict(m=16, build_degree=300, polish_degree=3000, polish_outers=2, tol=1e-4, chunk=8,
              verbose=False)
    va, xa, ra, ia = two_stage(op, lo, hi, lam_max, build_outers=2, **kw)
    state = tmp_path / "state.npz"
    two_stage(op, lo, hi, lam_max, build_outers=1, state_path=state, **{**kw, "polish_outers": 0})
    st = load_interior_state(state)
    assert st["stage"] == "build" and st["outer"] == 0 and st["lam_max"] == lam_max
    vb, xb, rb, ib = two_stage(op, lo, hi, lam_max, build_outers=2, state_path=state, resume=True,
                               **kw)
    strip = lambda h: [(r["stage"], r["outer"], r["in_window"], r["converged"]) for r in h]
    assert strip(ia["history"]) == strip(ib["history"])
    assert np.array_equal(va, vb) and np.array_equal(ra, rb)
    print("converged in-window pairs:", len(va), va)
    if len(va):
        ref = w[(w >= lo) & (w <= hi)]
        nearest = np.array([ref[np.argmin(np.abs(ref - v))] for v in va])
        assert np.abs(va - nearest).max() / hi < 1e-6
