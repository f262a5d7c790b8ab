"""Lanczos λ_max, lowpass Chebyshev subspace iteration and KPM eigenvalue counting.

Drop-in for eigenfns.chebyshev (reference eigenfns/chebyshev.py).  Operator
applications and recurrences run in the fused HIP passes; FP64 small algebra
on the device (see eigenfns_amd.linalg).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .linalg import (combine, deflate, gram, gram_herm, eigh, normalize_rows, ortho_block, rowdot,
                     rownorm, rownorm2, row_axpby, axpby)


def lanczos_lambda_max(op, iters: int = 24, seed: int = 0, n_seeds: int = 2) -> float:
    """Probable upper bound on λ_max: max over seeds of θ_max + ‖r‖ (chebyshev.py:28-64).

    Not certified; callers multiply by 1.05 as the reference does."""
    if n_seeds > 1:
        return max(lanczos_lambda_max(op, iters, seed + 1000 * s, n_seeds=1)
                   for s in range(n_seeds))
    v = normalize_rows(op.random_block(1, seed, kind="normal"))
    alphas, betas = [], []
    v_prev = None
    beta = 0.0
    for _ in range(iters):
        w = op.theta(v)
        alpha = float(rowdot(v, w).real.item())
        # w ← w − α v − β v_prev
        axpby(w, 1.0, v, -alpha, out=w)
        if v_prev is not None:
            axpby(w, 1.0, v_prev, -beta, out=w)
        beta_new = float(torch.sqrt(rownorm2(w)).item())
        alphas.append(alpha)
        betas.append(beta_new)
        if beta_new < 1e-20:
            break
        v_prev = v
        v = axpby(w, 1.0 / beta_new)
        beta = beta_new
    T = np.diag(alphas) + np.diag(betas[:-1], 1) + np.diag(betas[:-1], -1)
    return float(np.linalg.eigvalsh(T)[-1] + betas[-1])


def _cheb_filter_chunk(op, X, degree: int, lam_c: float, lam_max: float):
    """[λ_c, λ_max]-damping Chebyshev filter of `degree` (chebyshev.py:67-92).

    T_{k+1} = (2/e)(Θ − c)T_k − T_{k−1} is the fused step with alpha = 2/e,
    evaluated as 2·((1/e)Θ T_k − (c/e) T_k) − T_{k−1}; rows are rescaled every
    32 degrees when they exceed 1e30 and normalised at the end."""
    e = (lam_max - lam_c) / 2.0
    c = (lam_max + lam_c) / 2.0
    Xk_prev = X.clone()
    Xk = op.theta_affine(X, 1.0 / e, -c / e)            # Y = (ΘX − cX)/e
    for j in range(degree - 1):
        # Xn = (2/e)(Θ Xk − c Xk) − Xk_prev  → written into Xk_prev's buffer
        Tmp = op.theta_affine(Xk, 2.0 / e, -2.0 * c / e)
        axpby(Tmp, 1.0, Xk_prev, -1.0, out=Xk_prev)
        Xk_prev, Xk = Xk, Xk_prev
        if (j + 1) % 32 == 0:
            s = torch.amax(torch.abs(Xk.reshape(Xk.shape[0], -1)), dim=1).double()
            sc = torch.where(s > 1e30, 1.0 / torch.clamp(s, min=1e-30), torch.ones_like(s))
            if bool((sc != 1.0).any()):
                row_axpby(Xk, sc, out=Xk)
                row_axpby(Xk_prev, sc, out=Xk_prev)
    return normalize_rows(Xk)


@dataclass
class ChebStats:
    theta_applications: int = 0
    outer: list = field(default_factory=list)
    lambda_max: float = 0.0
    wall_seconds: float = 0.0


def chebsi_bottom_up(op, nev: int, guard: int = 48, degree: int = 200, tol: float = 1e-4,
                     max_outer: int = 30, chunk: int = 64, seed: int = 0,
                     lam_max: float | None = None, verbose: bool = True):
    """Lowest `nev` eigenpairs by bottom-up ChebSI with locking (chebyshev.py:103-210)."""
    t0 = time.perf_counter()
    m = nev + guard
    stats = ChebStats()
    if lam_max is None:
        lam_max = 1.05 * lanczos_lambda_max(op)
        stats.theta_applications += 24
    stats.lambda_max = lam_max
    X = ortho_block(op.random_block(m, seed, kind="normal"))
    n_locked = 0
    locked_vals = np.empty((0,))
    lam = None
    for outer in range(max_outer):
        if lam is None:
            lam_c = lam_max / 50.0
        else:
            hi = lam[min(m - 1, nev + guard // 2 - 1)]
            lam_c = float(min(hi * 1.1, lam_max / 4))
        act = X[n_locked:]
        outs = []
        for s in range(0, act.shape[0], chunk):
            piece = act[s:s + chunk]
            outs.append(_cheb_filter_chunk(op, piece, degree, lam_c, lam_max))
            stats.theta_applications += degree * piece.shape[0]
        act = torch.cat(outs, dim=0)
        if n_locked:
            act = deflate(act, X[:n_locked])
            act = deflate(act, X[:n_locked])
        act = ortho_block(act)
        Hact = op.theta(act)
        stats.theta_applications += act.shape[0]
        A = gram_herm(act, Hact)
        dead = rownorm(act) < 0.5
        A = (A + A.conj().T) / 2 + torch.diag(torch.where(dead, 1e4, 0.0).to(A.dtype))
        wA, VA = eigh(A)
        act = combine(VA, act)
        Hact = combine(VA, Hact)
        lam_act = wA.cpu().numpy()
        R = row_axpby(Hact, torch.ones_like(wA), act, -wA)
        rn = (rownorm(R) / torch.clamp(rownorm(Hact), min=1e-30)).cpu().numpy()
        want = nev - n_locked
        conv_prefix = 0
        for i in range(min(want, len(rn))):
            if rn[i] < tol:
                conv_prefix += 1
            else:
                break
        X = torch.cat([X[:n_locked], act], dim=0)
        lam_full = np.concatenate([locked_vals, lam_act])
        if conv_prefix:
            locked_vals = lam_full[:n_locked + conv_prefix]
            n_locked += conv_prefix
        lam = lam_full
        stats.outer.append({"outer": outer, "locked": n_locked, "lam_c": lam_c,
                            "median_res": float(np.median(rn))})
        if verbose:
            print(f"  outer {outer:2d}: locked {n_locked:4d}/{nev}  lam_c {lam_c:8.4f}  "
                  f"res q50 {np.quantile(rn, 0.5):.1e}  "
                  f"elapsed {time.perf_counter() - t0:6.1f}s", flush=True)
        if n_locked >= nev:
            break
    stats.wall_seconds = time.perf_counter() - t0
    return lam[:nev], X[:nev], stats


def kpm_coeffs(lam_b: float, lam_max: float, degree: int) -> np.ndarray:
    """Jackson-damped Chebyshev coefficients of 1_{λ<λ_b} on [0, λmax] (chebyshev.py:229-242)."""
    from .interior import jackson, step_coeffs
    xb = (2 * lam_b - lam_max) / lam_max
    return step_coeffs(xb, degree) * jackson(degree)


def kpm_count_below(op, lam_b: float, lam_max: float, degree: int = 500, n_probe: int = 30,
                    seed: int = 1, locked=None) -> tuple[float, float]:
    """Stochastic count of eigenvalues below λ_b (chebyshev.py:213-267).

    Returns (estimate, standard_error).  With `locked` the probes are deflated
    twice against it (completeness-gate mode).  The coefficient sum
    Σ c_k·Re⟨z, T_k z⟩ accumulates in FP64 on the device inside the fused step."""
    coef = kpm_coeffs(lam_b, lam_max, degree)
    Z = op.random_block(n_probe, seed, kind="rademacher")
    if locked is not None:
        Z = deflate(deflate(Z, locked), locked)
    inv_l = 2.0 / lam_max
    T0 = Z.clone()
    T1 = op.theta_affine(Z, inv_l, -1.0)
    est = coef[0] * rownorm2(Z) + coef[1] * rowdot(Z, T1).real
    est = est.contiguous()
    for k in range(2, degree + 1):
        op.kpm_count_step(T1, T0, Z, est, float(coef[k]), inv_l)
        T0, T1 = T1, T0
    per_probe = est.cpu().numpy()
    return float(per_probe.mean()), float(per_probe.std(ddof=1) / np.sqrt(n_probe))
