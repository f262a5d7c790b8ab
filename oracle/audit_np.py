"""Oracle: deflated-probe completeness audit "I2 v2" (NumPy, FP64).

Test infrastructure only (imported by tests/); restates
scripts/exp/exp_i2_v2.py:
  bandpass_rho       :53-60    ρ(λ) = Σ_k c_k T_k(x), c = Jackson-damped bandpass
  bandpass_trace     :127-150  per-probe Σ_k c_k Re⟨z, T_k(B) z⟩, z Rademacher,
                               Γ-masked, deflated twice against the locked set
  edge_leakage       :159-187  L = ∫_outside ρ(λ)·DOS(λ) dλ from KPM moments,
                               each side of the window integrated on its own
                               over `margin` (the reference's 0.6 by default)
  audit_verdict      :197-213  missed = mean − L; pass if |missed| <
                               max(0.5, 2·√(se² + se_L²)) (and the sub-gap
                               interval, if given, |missed| < 0.5)

Pinned by (tests/test_oracle_audit.py): for a fixed probe, the estimator equals
Σ_i |⟨v_i, z⟩|²·ρ(λ_i) from a dense FP64 eigendecomposition (and deflation
removes exactly the locked states' terms); edge_leakage of the exact moments
of a flat density of states equals the quadrature of ρ over the two outside
pieces (closed form for the flat DOS).
"""
from __future__ import annotations

import numpy as np

from .interior_np import bandpass_coeffs, jackson
from .linalg_np import deflate


def bandpass_rho(lam, lam_lo, lam_hi, lam_max, degree):
    coef = bandpass_coeffs(lam_lo, lam_hi, lam_max, degree)
    x = np.clip(2 * np.asarray(lam, np.float64) / lam_max - 1, -1, 1)
    th = np.arccos(x)
    return np.cos(np.outer(th, np.arange(len(coef)))) @ coef


def bandpass_trace(op, Z, lam_lo, lam_hi, lam_max, degree, locked=None):
    """Per-probe estimates for the given probe block Z (m, 2, G, G, G)."""
    coef = bandpass_coeffs(lam_lo, lam_hi, lam_max, degree)
    Z = np.asarray(Z, np.complex128)
    if locked is not None and len(locked):
        Z = deflate(deflate(Z, locked), locked)
    inv_l = 2.0 / lam_max

    def Bx(V):
        return inv_l * op.theta(V) - V

    zf = Z.reshape(Z.shape[0], -1)

    def dot(T):
        return np.real(np.sum(zf.conj() * T.reshape(T.shape[0], -1), axis=1))

    T0, T1 = Z, Bx(Z)
    est = coef[0] * dot(T0) + coef[1] * dot(T1)
    for k in range(2, degree + 1):
        T2 = 2.0 * Bx(T1) - T0
        est = est + coef[k] * dot(T2)
        T0, T1 = T1, T2
    return est


def edge_leakage(moments, kpm_lam_max, lam_lo, lam_hi, lam_max, degree, n_grid=6000,
                 margin=0.6):
    """(L, se_L): the window filter's weight on the states outside [lo, hi]
    predicted from KPM moments (per-probe rows) of the density of states."""
    mom = np.asarray(moments, np.float64)
    p = mom.shape[1] - 1
    gj = jackson(p)
    mu = mom.mean(0)[:p + 1] * gj
    se = mom.std(0, ddof=1)[:p + 1] / np.sqrt(mom.shape[0]) * gj
    grid = np.linspace(max(lam_lo - margin, 1e-3), lam_hi + margin, n_grid)
    xg = np.clip(2 * grid / kpm_lam_max - 1, -1 + 1e-12, 1 - 1e-12)
    th = np.arccos(xg)
    T = np.cos(np.outer(np.arange(p + 1), th))
    dos = (mu[0] + 2 * (mu[1:] @ T[1:])) / (np.pi * np.sin(th)) * 2 / kpm_lam_max
    dos_se = np.sqrt(se[0] ** 2 + 4 * ((se[1:] ** 2) @ (T[1:] ** 2))) \
        / (np.pi * np.sin(th)) * 2 / kpm_lam_max
    rho = bandpass_rho(grid, lam_lo, lam_hi, lam_max, degree)
    lo, hi = grid < lam_lo, grid > lam_hi

    def piece(y):
        return float(np.trapezoid(y[lo], grid[lo])) + float(np.trapezoid(y[hi], grid[hi]))

    return piece(rho * dos), piece(rho * dos_se)


def audit_verdict(window, sub=None):
    """window / sub: dicts with deflated_estimate, se, predicted_edge_leakage, leakage_se."""
    def missed(r):
        return r["deflated_estimate"] - r["predicted_edge_leakage"]

    tot_se = float(np.hypot(window["se"], window["leakage_se"]))
    ok = abs(missed(window)) < max(0.5, 2 * tot_se)
    if sub is not None:
        ok = ok and abs(missed(sub)) < 0.5
    return {"missed_window": missed(window), "missed_window_se": tot_se,
            "missed_subgap": missed(sub) if sub else None, "pass": bool(ok)}
