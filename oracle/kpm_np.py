"""Oracle: KPM Chebyshev moments with the doubling identities, and their analysis.

Restates scripts/exp/exp_kpm_dos.py:125-163 (moment collection) and
scripts/exp/exp_kpm_analyze.py:39-70 (counting function, Jackson DOS on the
Chebyshev–Gauss grid, smearing width).

  μ₀ = ⟨z,z⟩, μ₁ = ⟨z,Bz⟩, and for k = 1..⌊p/2⌋ with T_{k+1} = 2BT_k − T_{k−1}:
  μ_{2k} = 2⟨T_k,T_k⟩ − μ₀,   μ_{2k+1} = 2⟨T_{k+1},T_k⟩ − μ₁,
  B = (2Θ − λ_max)/λ_max.

Pinned by:
  tests/test_oracle_spectral.py  the doubling identities against the direct
      recurrence μ_k = ⟨z, T_k(B) z⟩; μ_k = Σ_i |⟨v_i, z⟩|²·T_k(x_i) from a
      dense eigendecomposition at G=8; μ₀ = 2G³ − 2 for Rademacher probes at
      Γ; the counting function at the top of the spectrum equals 2G³ − 2.
  tests/test_oracle_kpm.py       dos_curve: its quadrature over λ is exactly
      μ₀ (Chebyshev orthogonality on the midpoint grid); one state's curve
      has the Jackson kernel's exact mean g₁·a and second moment
      (1 + g₂T₂(a))/2; its integral below λ_b is the counting function.
      smearing_width: the kernel's exact standard deviation to O(1/p).
"""
from __future__ import annotations

import numpy as np

from .interior_np import jackson, step_coeffs
from .linalg_np import flat
from .rng_np import random_block


def _dots(A, B):
    return np.real(np.sum(flat(A).conj() * flat(B), axis=1)).astype(np.float64)


def kpm_moments(op, degree, n_probe, seed, lam_max, probes=None):
    """(n_probe, 2·⌊p/2⌋+2) FP64 moments; ``probes`` overrides the generated block."""
    half = degree // 2
    if probes is None:
        Z = random_block("rademacher", seed, 0, n_probe, op.grid_size, op.mask)
    else:
        Z = np.asarray(probes)
    Z = Z.astype(np.complex128)
    inv_l = 2.0 / lam_max

    def Bx(V):
        return inv_l * op.theta(V) - V

    mom = np.zeros((Z.shape[0], 2 * half + 2))
    t0, t1 = Z, Bx(Z)
    mu0 = _dots(t0, t0)
    mu1 = _dots(t0, t1)
    mom[:, 0], mom[:, 1] = mu0, mu1
    for k in range(1, half + 1):
        t2 = 2.0 * Bx(t1) - t0
        mom[:, 2 * k] = 2.0 * _dots(t1, t1) - mu0
        mom[:, 2 * k + 1] = 2.0 * _dots(t2, t1) - mu1
        t0, t1 = t1, t2
    return mom


def kpm_moments_direct(op, degree, Z, lam_max):
    """μ_k = ⟨z, T_k(B) z⟩ for k = 0..degree by the plain three-term recurrence."""
    inv_l = 2.0 / lam_max
    Z = np.asarray(Z).astype(np.complex128)

    def Bx(V):
        return inv_l * op.theta(V) - V

    mom = np.zeros((Z.shape[0], degree + 1))
    t0, t1 = Z, Bx(Z)
    mom[:, 0] = _dots(Z, t0)
    if degree >= 1:
        mom[:, 1] = _dots(Z, t1)
    for k in range(2, degree + 1):
        t2 = 2.0 * Bx(t1) - t0
        mom[:, k] = _dots(Z, t2)
        t0, t1 = t1, t2
    return mom


def counting(mom, lam_max, lams, degree=None):
    p = (mom.shape[1] - 1) if degree is None else degree
    g = jackson(p)
    xs = (2.0 * np.asarray(lams) - lam_max) / lam_max
    est = np.empty((mom.shape[0], len(xs)))
    for j, xb in enumerate(xs):
        est[:, j] = mom[:, :p + 1] @ (step_coeffs(xb, p) * g)
    return est.mean(0), est.std(0, ddof=1) / np.sqrt(mom.shape[0]), est


def dos_curve(mom, lam_max, n_pts=4000, degree=None):
    p = (mom.shape[1] - 1) if degree is None else degree
    g = jackson(p)
    mu = mom.mean(0)[:p + 1] * g
    theta = (np.arange(n_pts) + 0.5) * np.pi / n_pts
    x = np.cos(theta)
    T = np.cos(np.outer(np.arange(p + 1), theta))
    rho_x = (mu[0] + 2 * (mu[1:] @ T[1:])) / (np.pi * np.sin(theta))
    lam = (x + 1) * lam_max / 2
    order = np.argsort(lam)
    return lam[order], (rho_x * 2 / lam_max)[order]


def smearing_width(lam, lam_max, degree):
    x = np.clip(2 * lam / lam_max - 1, -1 + 1e-12, 1 - 1e-12)
    return np.pi / degree * np.sqrt(1 - x ** 2) * lam_max / 2
