"""Pins the oracle's filtered subspace iteration (eigenfns/interior.py:357-427,
``bandpass_subspace``) and two-stage driver (scripts/run_interior.py:154-216)
against a dense FP64 eigendecomposition at G=8: the window is placed between
dense eigenvalues so it holds a known set of states, which the iteration must
return — converged, inside the window, with eigenvalues and eigenvectors of
the dense spectrum."""
import numpy as np
import pytest

from oracle import interior_np
from oracle.dense_np import dense_eigh
from oracle.operator_np import MaxwellOperatorNP


@pytest.fixture(scope="module")
def g8():
    rng = np.random.default_rng(3)
    G = 8
    eps = np.asarray(1.0 + 7.57 * (rng.random((G, G, G)) > 0.7), np.float32)
    op = MaxwellOperatorNP(eps, 1.0)
    w, V = dense_eigh(op)
    return op, w, V


def test_bandpass_subspace_returns_exactly_the_window_states(g8):
    op, w, V = g8
    lo, hi = (w[4] + w[5]) / 2, (w[11] + w[12]) / 2
    lam_max = 1.05 * w.max()
    Xr, info = interior_np.bandpass_subspace(op, lo, hi, lam_max, m=24, degree=300, max_outer=8,
                                             res_tol=1e-6, target_count=7, seed=0)
    lam, rn = info["lam"], info["res"]
    conv = (lam >= lo) & (lam <= hi) & (rn < 1e-6)
    assert conv.sum() == 7
    assert np.abs(np.sort(lam[conv]) - w[5:12]).max() < 1e-9 * w[11]
    # the converged vectors span the dense window eigenspace
    Q = Xr.reshape(Xr.shape[0], -1)[conv]
    Q = Q / np.linalg.norm(Q, axis=1, keepdims=True)
    P = V[:, 5:12].conj().T @ Q.T
    assert np.abs(np.linalg.svd(P, compute_uv=False) - 1).max() < 1e-9
    # stop rule: the target count ends the outer loop
    assert info["outer_stats"][-1]["converged"] >= 7


def test_bandpass_subspace_stagnation_stop(g8):
    # without a target the loop stops once the converged count has not grown
    # for an outer after the second (interior.py:420-421)
    op, w, V = g8
    lo, hi = (w[4] + w[5]) / 2, (w[11] + w[12]) / 2
    _, info = interior_np.bandpass_subspace(op, lo, hi, 1.05 * w.max(), m=16, degree=200,
                                            max_outer=8, res_tol=1e-6, seed=1)
    counts = [h["converged"] for h in info["outer_stats"]]
    n = len(counts)
    assert n <= 8
    if n < 8:
        assert n >= 3 and counts[-1] <= max(counts[:-1])
