"""KPM density of states: Chebyshev moments on the device, analysis on the host.

Lifts the reference's script-only KPM DOS (scripts/exp/exp_kpm_dos.py:125-163)
into the library, plus its analysis (scripts/exp/exp_kpm_analyze.py:22-70).

Each degree is one fused HIP pass set (efn_kpm_step): the recurrence
T_{k+1} = 2·B·T_k − T_{k−1} with B = (2/λmax)Θ − I is computed in the
operator's last pass, which also reduces ⟨T_k,T_k⟩ and Re⟨T_{k+1},T_k⟩ in
FP64 and writes μ_{2k} = 2⟨T_k,T_k⟩ − μ₀ and μ_{2k+1} = 2⟨T_{k+1},T_k⟩ − μ₁
straight into a device moment array — no host synchronisation per degree
(the reference makes two).
"""
from __future__ import annotations

import json
import time

import numpy as np
import torch

from .linalg import rowdot, rownorm2
from .interior import jackson, step_coeffs


def kpm_moments(op, degree: int, n_probe: int, seed: int = 7, lam_max: float | None = None,
                chunk: int = 0, verbose: bool = False, return_meta: bool = False):
    """Per-probe moments (n_probe, 2·⌊degree/2⌋+2) in FP64 (numpy).

    Probes are Rademacher ±1 rows of the counter-based stream `seed` (row =
    probe index, so results do not depend on `chunk`), Γ-masked as
    exp_kpm_dos.py:139-140.  `lam_max` defaults to 1.05·Lanczos
    (exp_kpm_dos.py:84)."""
    from .chebyshev import lanczos_lambda_max
    if lam_max is None:
        lam_max = 1.05 * lanczos_lambda_max(op)
    half = degree // 2
    n_mom = 2 * half + 2
    chunk = int(chunk) if chunk and chunk > 0 else int(n_probe)
    inv_l = 2.0 / lam_max
    out = np.zeros((n_probe, n_mom), np.float64)
    t_start = time.perf_counter()
    for s in range(0, n_probe, chunk):
        c = min(chunk, n_probe - s)
        T0 = op.random_block(c, seed, kind="rademacher", row0=s)
        T1 = op.theta_affine(T0, inv_l, -1.0)
        mom = torch.zeros((c, n_mom), dtype=torch.float64, device=T0.device)
        mom[:, 0] = rownorm2(T0)
        mom[:, 1] = rowdot(T0, T1).real
        for k in range(1, half + 1):
            op.kpm_step(T1, T0, mom, k, inv_l)
            T0, T1 = T1, T0
            if verbose and k % 1000 == 0:
                torch.cuda.synchronize()
                print(f"  probes {s}..{s + c - 1}: k={k}/{half} "
                      f"elapsed {time.perf_counter() - t_start:.1f}s", flush=True)
        out[s:s + c] = mom.cpu().numpy()
        del T0, T1, mom
    if return_meta:
        meta = dict(grid=op.grid_size, degree=degree, probes=n_probe, chunk=chunk,
                    lam_max=lam_max, seed=seed, wall_seconds=time.perf_counter() - t_start)
        return out, meta
    return out


def kpm_dos(op, degree: int, n_probe: int, seed: int = 7, lam_max: float | None = None,
            chunk: int = 0, verbose: bool = False) -> dict:
    """The fields of the reference's <tag>_kpm.npz: moments, lam_max, meta (JSON)."""
    mom, meta = kpm_moments(op, degree, n_probe, seed, lam_max, chunk, verbose, return_meta=True)
    return {"moments": mom, "lam_max": meta["lam_max"], "meta": json.dumps(meta)}


# ---- analysis (exp_kpm_analyze.py:39-70), host FP64 ----------------------------
def counting(mom, lam_max, lams, degree=None):
    """N(λ) per probe → (mean, standard error, per-probe) over thresholds `lams`."""
    mom = np.asarray(mom)
    p = (mom.shape[1] - 1) if degree is None else degree
    g = jackson(p)
    xs = (2.0 * np.asarray(lams, np.float64) - lam_max) / lam_max
    coef = np.stack([step_coeffs(xb, p) * g for xb in xs], axis=1)
    est = mom[:, :p + 1] @ coef
    return est.mean(0), est.std(0, ddof=1) / np.sqrt(mom.shape[0]), est


def dos_curve(mom, lam_max, n_pts=4000, degree=None):
    """Jackson-damped DOS on the Chebyshev–Gauss grid (per-probe mean)."""
    mom = np.asarray(mom)
    p = (mom.shape[1] - 1) if degree is None else degree
    mu = mom.mean(0)[:p + 1] * jackson(p)
    theta = (np.arange(n_pts) + 0.5) * np.pi / n_pts
    T = np.cos(np.outer(np.arange(p + 1), theta))
    rho_x = (mu[0] + 2 * (mu[1:] @ T[1:])) / (np.pi * np.sin(theta))
    lam = (np.cos(theta) + 1) * lam_max / 2
    order = np.argsort(lam)
    return lam[order], (rho_x * 2 / lam_max)[order]


def smearing_width(lam, lam_max, degree):
    """Jackson kernel width in λ units at λ (π/p in θ)."""
    x = np.clip(2 * np.asarray(lam) / lam_max - 1, -1 + 1e-12, 1 - 1e-12)
    return np.pi / degree * np.sqrt(1 - x ** 2) * lam_max / 2
