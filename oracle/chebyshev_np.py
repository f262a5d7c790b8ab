"""Oracle: Lanczos λ_max and the Jackson-damped KPM step-trace count (NumPy).

Restates eigenfns/chebyshev.py:
  lanczos_lambda_max  chebyshev.py:28-64   (θ_max + β_last over seeds; callers ×1.05)
  kpm_count_below     chebyshev.py:213-267 (Jackson-damped step trace, optional deflation)
Random start vectors and probes come from ``rng_np`` (the reference's JAX
threefry streams are not reproducible).

Pinned by:
  tests/test_oracle_chebyshev.py  Lanczos: θ_max + β_k equals the Rayleigh–Ritz
      maximum of the Krylov space K_k(Θ, v₀) (Arnoldi with full
      re-orthogonalisation on the dense FP64 matrix) plus the norm of Θq_k
      outside it, for k = 3, 6, 10 and two seeds; the max over seeds; the
      bracket of the dense λ_max at 24 steps.
  tests/test_oracle_spectral.py   KPM count against the exact count of dense
      eigenvalues below λ_b (within the stochastic error and the Jackson
      smearing).
"""
from __future__ import annotations

import numpy as np

from .linalg_np import deflate, flat
from .interior_np import jackson, step_coeffs
from .rng_np import random_block


def lanczos_lambda_max(op, iters=24, seed=0, n_seeds=2):
    if n_seeds > 1:
        return max(lanczos_lambda_max(op, iters, seed + 1000 * s, n_seeds=1)
                   for s in range(n_seeds))
    v = random_block("normal", seed, 0, 1, op.grid_size, op.mask).astype(np.complex128)
    v = v / np.linalg.norm(flat(v))
    alphas, betas = [], []
    v_prev = None
    beta = 0.0
    for _ in range(iters):
        w = op.theta(v)
        alpha = float(np.real(np.sum(flat(v).conj() * flat(w))))
        w = w - alpha * v - (beta * v_prev if v_prev is not None else 0.0)
        beta_new = float(np.linalg.norm(flat(w)))
        alphas.append(alpha)
        betas.append(beta_new)
        if beta_new < 1e-20:
            break
        v_prev = v
        v = w / beta_new
        beta = beta_new
    T = np.diag(alphas) + np.diag(betas[:-1], 1) + np.diag(betas[:-1], -1)
    return float(np.linalg.eigvalsh(T)[-1] + betas[-1])


def kpm_count_coeffs(lam_b, lam_max, degree):
    """Jackson-damped step coefficients of chebyshev.py:229-242 (FP64)."""
    xb = (2 * lam_b - lam_max) / lam_max
    return step_coeffs(xb, degree) * jackson(degree)


def kpm_count_below(op, lam_b, lam_max, degree=500, n_probe=30, seed=1, locked=None):
    coef = kpm_count_coeffs(lam_b, lam_max, degree)
    Z = random_block("rademacher", seed, 0, n_probe, op.grid_size, op.mask).astype(np.complex128)
    if locked is not None:
        Z = deflate(deflate(Z, locked), locked)

    def Bx(V):
        return (2.0 / lam_max) * op.theta(V) - V

    zf = flat(Z)
    T0, T1 = Z, Bx(Z)
    est = coef[0] * np.real(np.sum(zf.conj() * flat(T0), axis=1))
    est = est + coef[1] * np.real(np.sum(zf.conj() * flat(T1), axis=1))
    for k in range(2, degree + 1):
        T2 = 2.0 * Bx(T1) - T0
        est = est + coef[k] * np.real(np.sum(zf.conj() * flat(T2), axis=1))
        T0, T1 = T1, T2
    return float(est.mean()), float(est.std(ddof=1) / np.sqrt(n_probe)), est
