"""Deflated-probe completeness audit ("I2 v2", scripts/exp/exp_i2_v2.py).

Given the converged eigenvectors found in a window [λ_lo, λ_hi] (the locked
set), Rademacher probes z deflated twice against them give

    E[Σ_k c_k Re⟨z, T_k(B) z⟩] = Σ_{missed in window} ρ(λ_i) + L,

with ρ the Jackson-damped bandpass of the window (c its Chebyshev
coefficients) and L = Σ_{states outside} ρ(λ_i) the edge leakage, predicted
from the KPM density of states.  missed = mean − L is the number of window
states the solve did not find.

The per-probe sum runs in the fused KPM-count step (one operator pass per
degree, Re⟨z, T_k z⟩ accumulated in FP64 on the device, no per-degree host
sync); deflation against the HBM-resident locked set uses the FP64-accumulated
Gram; the leakage integral is host NumPy.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from .interior import bandpass_coeffs, jackson
from .linalg import deflate, rowdot, rownorm2


def bandpass_rho(lam, lam_lo, lam_hi, lam_max, degree):
    """ρ(λ) of the window filter (exp_i2_v2.py:53-60)."""
    coef = bandpass_coeffs(lam_lo, lam_hi, lam_max, degree)
    x = np.clip(2 * np.asarray(lam, np.float64) / lam_max - 1, -1, 1)
    return np.cos(np.outer(np.arccos(x), np.arange(len(coef)))) @ coef


def bandpass_trace(op, lam_lo, lam_hi, lam_max, degree, n_probe, seed=23, locked=None,
                   chunk=0, verbose=False):
    """Per-probe estimates Σ_k c_k Re⟨z, T_k(B) z⟩ (exp_i2_v2.py:127-150).

    Probe rows are rows 0..n_probe−1 of the counter-based Rademacher stream
    `seed`, Γ-masked, deflated twice against `locked` ((n, 2, G, G, G) device
    block or None).  Returns a NumPy array (n_probe,)."""
    coef = bandpass_coeffs(lam_lo, lam_hi, lam_max, degree)
    inv_l = 2.0 / lam_max
    chunk = chunk or n_probe
    out = np.empty(n_probe)
    t0 = time.perf_counter()
    for s in range(0, n_probe, chunk):
        c = min(chunk, n_probe - s)
        Z = op.random_block(c, seed, kind="rademacher", row0=s)
        if locked is not None and len(locked):
            Z = deflate(deflate(Z, locked), locked)
        T0 = Z.clone()
        T1 = op.theta_affine(Z, inv_l, -1.0)
        est = (coef[0] * rownorm2(Z) + coef[1] * rowdot(Z, T1).real).contiguous()
        for k in range(2, degree + 1):
            op.kpm_count_step(T1, T0, Z, est, float(coef[k]), inv_l)
            T0, T1 = T1, T0
            if verbose and k % 2000 == 0:
                torch.cuda.synchronize()
                print(f"    probes {s}..{s + c - 1} k={k}/{degree} "
                      f"{time.perf_counter() - t0:.0f}s", flush=True)
        out[s:s + c] = est.cpu().numpy()
        del Z, T0, T1, est
    return out


def edge_leakage(moments, kpm_lam_max, lam_lo, lam_hi, lam_max, degree, n_grid=6000,
                 margin=0.6):
    """(L, se_L): the filter's weight on the states outside the window from KPM
    moments (rows per probe) of the density of states, each outside side
    integrated on its own over `margin` λ-units (exp_i2_v2.py:159-187; the
    reference's 0.6 suits its µm⁻² windows — the margin must cover the
    filter's transition tails, ≈ (π/degree)·λ_max/2)."""
    mom = np.asarray(moments, np.float64)
    p = mom.shape[1] - 1
    gj = jackson(p)
    mu = mom.mean(0)[:p + 1] * gj
    se = mom.std(0, ddof=1)[:p + 1] / np.sqrt(mom.shape[0]) * gj
    grid = np.linspace(max(lam_lo - margin, 1e-3), lam_hi + margin, n_grid)
    x = np.clip(2 * grid / kpm_lam_max - 1, -1 + 1e-12, 1 - 1e-12)
    th = np.arccos(x)
    T = np.cos(np.outer(np.arange(p + 1), th))
    scale = 2 / kpm_lam_max / (np.pi * np.sin(th))
    dos = (mu[0] + 2 * (mu[1:] @ T[1:])) * scale
    dos_se = np.sqrt(se[0] ** 2 + 4 * ((se[1:] ** 2) @ (T[1:] ** 2))) * scale
    rho = bandpass_rho(grid, lam_lo, lam_hi, lam_max, degree)
    outside = (grid < lam_lo, grid > lam_hi)
    leak = sum(float(np.trapezoid((rho * dos)[m], grid[m])) for m in outside)
    leak_se = sum(float(np.trapezoid((rho * dos_se)[m], grid[m])) for m in outside)
    return leak, leak_se


def completeness_audit(op, locked, lam_found, lam_lo, lam_hi, lam_max, degree=24000,
                       n_probe=12, seed=23, kpm=None, sub=None, chunk=4, verbose=False,
                       leak_margin=0.6):
    """The I2 v2 gate (exp_i2_v2.py:114-213).  `kpm` = (moments, kpm_lam_max)
    for the leakage prediction (None: no prediction, no verdict); `sub` =
    (sub_lo, sub_hi), a sub-interval whose edges sit in a spectral gap.
    Returns the results per interval and the pass verdict."""
    lam_found = np.asarray(lam_found, np.float64)
    intervals = [("window", lam_lo, lam_hi)]
    if sub is not None:
        intervals.append(("sub_gap", sub[0], sub[1]))
    results = {}
    t_all = time.perf_counter()
    for name, a, b in intervals:
        t0 = time.perf_counter()
        est = bandpass_trace(op, a, b, lam_max, degree, n_probe, seed, locked, chunk, verbose)
        leak = leak_se = None
        if kpm is not None:
            leak, leak_se = edge_leakage(kpm[0], kpm[1], a, b, lam_max, degree,
                                         margin=leak_margin)
        results[name] = {
            "interval": [a, b],
            "n_found_inside": int(((lam_found >= a) & (lam_found <= b)).sum()),
            "found_rho_weight": float(bandpass_rho(lam_found, a, b, lam_max, degree).sum()),
            "deflated_estimate": float(est.mean()),
            "se": float(est.std(ddof=1) / np.sqrt(n_probe)),
            "per_probe": est.tolist(),
            "predicted_edge_leakage": leak, "leakage_se": leak_se,
            "missed_estimate": (float(est.mean()) - leak) if leak is not None else None,
            "wall_seconds": time.perf_counter() - t0,
        }
    win, sb = results["window"], results.get("sub_gap")
    out = {"degree": degree, "probes": n_probe,
           "n_deflated": 0 if locked is None else int(len(locked)), "results": results,
           "wall_seconds": time.perf_counter() - t_all}
    if kpm is not None:
        tot_se = float(np.hypot(win["se"], win["leakage_se"]))
        out.update({
            "missed_window": win["missed_estimate"], "missed_window_se": tot_se,
            "missed_subgap": sb["missed_estimate"] if sb else None,
            "pass": bool(abs(win["missed_estimate"]) < max(0.5, 2 * tot_se)
                         and (sb is None or abs(sb["missed_estimate"]) < 0.5))})
    return out
