"""Pins the oracle's Lanczos λ_max estimate (eigenfns/chebyshev.py:28-64).

The estimate is θ_max(T_k) + β_k of a k-step Lanczos run.  Both terms are
fixed independently of the three-term recurrence: the Rayleigh–Ritz values of
the Krylov space K_k(Θ, v₀), built here by Arnoldi with full
re-orthogonalisation on the dense FP64 matrix, are the eigenvalues of T_k, and
β_k is the norm of the part of Θq_k outside that space.  Dropping β_k, using
the wrong Ritz value or the wrong start vector fails these checks.
"""
import numpy as np
import pytest

from oracle import chebyshev_np
from oracle.dense_np import dense_theta
from oracle.operator_np import MaxwellOperatorNP
from oracle.rng_np import random_block


@pytest.fixture(scope="module")
def small():
    rng = np.random.default_rng(3)
    G = 8
    eps = np.asarray(1.0 + 7.57 * (rng.random((G, G, G)) > 0.7), np.float32)
    op = MaxwellOperatorNP(eps, 1.0, (0.1, 0.0, 0.3))
    M = dense_theta(op)
    return op, (M + M.conj().T) / 2


def _krylov_estimate(M, v0, k):
    n = M.shape[0]
    Q = np.zeros((n, k), np.complex128)
    Q[:, 0] = v0 / np.linalg.norm(v0)
    for j in range(1, k):
        w = M @ Q[:, j - 1]
        for _ in range(2):  # full re-orthogonalisation, twice
            w = w - Q[:, :j] @ (Q[:, :j].conj().T @ w)
        Q[:, j] = w / np.linalg.norm(w)
    H = Q.conj().T @ M @ Q
    r = M @ Q[:, -1]
    for _ in range(2):
        r = r - Q @ (Q.conj().T @ r)
    return float(np.linalg.eigvalsh((H + H.conj().T) / 2)[-1] + np.linalg.norm(r))


@pytest.mark.parametrize("iters", [3, 6, 10])
def test_lanczos_equals_krylov_rayleigh_ritz_plus_residual(small, iters):
    op, M = small
    for seed in (0, 7):
        v0 = random_block("normal", seed, 0, 1, op.grid_size, op.mask).astype(np.complex128).ravel()
        want = _krylov_estimate(M, v0, iters)
        got = chebyshev_np.lanczos_lambda_max(op, iters=iters, seed=seed, n_seeds=1)
        assert abs(got - want) < 1e-9 * want, (iters, seed, got, want)


def test_lanczos_takes_the_max_over_seeds(small):
    op, M = small
    ests = []
    for s in range(3):
        v0 = random_block("normal", 5 + 1000 * s, 0, 1, op.grid_size, op.mask)
        ests.append(_krylov_estimate(M, v0.astype(np.complex128).ravel(), 6))
    got = chebyshev_np.lanczos_lambda_max(op, iters=6, seed=5, n_seeds=3)
    assert abs(got - max(ests)) < 1e-9 * max(ests)


def test_lanczos_converges_to_the_dense_maximum(small):
    # the top Ritz value converges first; the estimate brackets λ_max from above
    # once β is small against the gap, as the reference's callers assume
    op, M = small
    lmax = float(np.linalg.eigvalsh(M)[-1])
    est = chebyshev_np.lanczos_lambda_max(op, iters=24)
    assert lmax * (1 - 1e-9) <= est < 1.5 * lmax
