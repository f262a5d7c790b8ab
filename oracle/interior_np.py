"""Oracle: bandpass Chebyshev filter, Rayleigh–Ritz extraction, filtered SI (NumPy).

Restates eigenfns/interior.py:
  jackson            interior.py:312-316  (also exp_kpm_analyze.py:22-26)
  step_coeffs        interior.py:319-325  (also exp_kpm_analyze.py:29-36)
  bandpass_coeffs    interior.py:328-333
  apply_cheb_poly    interior.py:336-354
  rr_extract         interior.py:37-88   (SVQB → A = XᴴΘX → dead penalty 1e8 →
                                           eigh → rotate → normalised RQ, relres)
  bandpass_subspace  interior.py:357-427
(the two-stage protocol of scripts/run_interior.py:154-216 is checked on the
device against a dense FP64 eigendecomposition directly).

Pinned by:
  tests/test_oracle_spectral.py  Jackson kernel closed form (g₀ = 1, monotone,
      the kernel is a non-negative density); step coefficients against
      quadrature of the indicator; apply_cheb_poly against V·diag(p(λ))·Vᴴ·X
      from a dense eigendecomposition at G=8; rr_extract on an invariant
      subspace (tests/test_interior.py:80-92 of the reference).
  tests/test_oracle_interior.py  bandpass_subspace at G=8 returns exactly the
      dense window states (eigenvalues to 1e-9, subspace angles to 1e-9) and
      honours the target and stagnation stop rules.
"""
from __future__ import annotations

import numpy as np

from .linalg_np import combine, flat, gram, ortho_block, rownorm
from .rng_np import random_block


def jackson(p: int) -> np.ndarray:
    k = np.arange(p + 1)
    g = ((p - k + 1) * np.cos(np.pi * k / (p + 1))
         + np.sin(np.pi * k / (p + 1)) / np.tan(np.pi / (p + 1)))
    return g / (p + 1)


def step_coeffs(xb: float, p: int) -> np.ndarray:
    """Chebyshev coefficients of 1_{x < xb} on [-1, 1], orders 0..p."""
    k = np.arange(1, p + 1)
    tb = np.arccos(np.clip(xb, -1.0, 1.0))
    c = np.empty(p + 1)
    c[0] = 1 - tb / np.pi
    c[1:] = -2 * np.sin(k * tb) / (k * np.pi)
    return c


def bandpass_coeffs(lam_lo, lam_hi, lam_max, degree):
    x_lo = 2 * lam_lo / lam_max - 1
    x_hi = 2 * lam_hi / lam_max - 1
    return jackson(degree) * (step_coeffs(x_hi, degree) - step_coeffs(x_lo, degree))


def apply_cheb_poly(op, X, coef, lam_max, coef_dtype=np.float64):
    """Y = Σ_k c_k T_k(B) X, B = (2Θ − λ_max)/λ_max, rows normalised."""
    inv_l = 2.0 / lam_max
    c = np.asarray(coef, coef_dtype)

    def Bx(V):
        return inv_l * op.theta(V) - V

    T0 = np.asarray(X)
    T1 = Bx(T0)
    Y = c[0] * T0 + c[1] * T1
    for k in range(2, len(coef)):
        T2 = 2.0 * Bx(T1) - T0
        Y = Y + c[k] * T2
        T0, T1 = T1, T2
    return Y / np.maximum(rownorm(Y), 1e-30)[:, None, None, None, None]


def rr_extract(op, X):
    """Returns (lam (m,), Xr, relres (m,)), sorted by λ, dead rows at +inf."""
    X = ortho_block(np.asarray(X))
    HX = op.theta(X)
    A = gram(X, HX)
    A = (A + A.conj().T) / 2
    dead = rownorm(X) < 0.5
    A[np.diag_indices_from(A)] += np.where(dead, 1e8, 0.0)
    wA, VA = np.linalg.eigh(A)
    Xr = combine(VA, X)
    Hr = op.theta(Xr)
    xn2 = rownorm(Xr) ** 2
    lam = np.real(np.sum(flat(Xr).conj() * flat(Hr), axis=1)) / np.maximum(xn2, 1e-30)
    R = Hr - lam[:, None, None, None, None] * Xr
    rn = rownorm(R) / np.maximum(rownorm(Hr), 1e-30)
    bad = rownorm(Xr) < 0.5
    return np.where(bad, np.inf, lam), Xr, np.where(bad, np.inf, rn)


def bandpass_subspace(op, lam_lo, lam_hi, lam_max, m=64, degree=2000, max_outer=8,
                      res_tol=1e-4, target_count=None, seed=0, X0=None):
    if X0 is not None:
        X = np.asarray(X0)
        m = X.shape[0]
    else:
        X = random_block("normal", seed, 0, m, op.grid_size, op.mask).astype(np.complex128)
    coef = bandpass_coeffs(lam_lo, lam_hi, lam_max, degree)
    hist = []
    best = -1
    lam = Xr = rn = None
    for outer in range(max_outer):
        X = apply_cheb_poly(op, X, coef, lam_max)
        lam, Xr, rn = rr_extract(op, X)
        inwin = (lam >= lam_lo) & (lam <= lam_hi) & np.isfinite(lam)
        conv = inwin & (rn < res_tol)
        hist.append({"outer": outer, "in_window": int(inwin.sum()),
                     "converged": int(conv.sum())})
        if target_count is not None and conv.sum() >= target_count:
            break
        if int(conv.sum()) <= best and outer >= 2:
            break
        best = max(best, int(conv.sum()))
        X = Xr
    return Xr, {"lam": lam, "res": rn, "outer_stats": hist}
