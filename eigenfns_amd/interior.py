"""Interior eigenpairs: bandpass Chebyshev filtered subspace iteration + Rayleigh–Ritz.

Drop-in for the bandpass/RR part of eigenfns.interior (reference
eigenfns/interior.py:37-238, :312-466).  The filter's per-degree update
T_{k+1} = 2·B·T_k − T_{k−1}, Y += c_k·T_{k+1} is one fused HIP pass set
(efn_cheb_step): one HBM round trip per degree.  On MI355X the basis is
device-resident at every configuration of interest (192³ × 104 vectors =
11.8 GB of 288 GB), so the `_hosted` variants keep their signatures but run
the device path.  The rejected bake-off methods (folded spectrum, polish,
MINRES shift-invert) are out of scope.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from .linalg import (SVQB_DROP, combine, combine_into, dead_spread, eigh, gram, gram_herm, normalize_rows,
                     ortho_block, row_axpby, rowdot, rownorm, rownorm2, axpby)


# ---------------------------------------------------------------- coefficients
def jackson(p: int) -> np.ndarray:
    """Jackson damping factors g_k, k = 0..p (interior.py:312-316)."""
    k = np.arange(p + 1)
    g = ((p - k + 1) * np.cos(np.pi * k / (p + 1))
         + np.sin(np.pi * k / (p + 1)) / np.tan(np.pi / (p + 1)))
    return g / (p + 1)


_jackson = jackson


def step_coeffs(xb: float, p: int) -> np.ndarray:
    """Chebyshev coefficients of 1_{x < xb} on [−1, 1], orders 0..p (interior.py:319-325)."""
    k = np.arange(1, p + 1)
    tb = np.arccos(np.clip(xb, -1.0, 1.0))
    c = np.empty(p + 1)
    c[0] = 1 - tb / np.pi
    c[1:] = -2 * np.sin(k * tb) / (k * np.pi)
    return c


_step_coeffs = step_coeffs


def bandpass_coeffs(lam_lo: float, lam_hi: float, lam_max: float, degree: int) -> np.ndarray:
    """Jackson-damped Chebyshev coefficients of 1_{lam_lo < λ < lam_hi} (interior.py:328-333)."""
    x_lo = 2 * lam_lo / lam_max - 1
    x_hi = 2 * lam_hi / lam_max - 1
    return jackson(degree) * (step_coeffs(x_hi, degree) - step_coeffs(x_lo, degree))


# ---------------------------------------------------------------- filter
def _apply_cheb_poly(op, X, coef: np.ndarray, lam_max: float):
    """Y = Σ_k coef_k T_k(B) X, B = (2Θ − λ_max)/λ_max; rows normalised (interior.py:336-354).

    len(coef) − 1 operator applications; coefficients cast to FP32 as the
    reference does."""
    inv_l = 2.0 / lam_max
    c = np.asarray(coef, np.float32)
    if isinstance(X, list):
        X = X.pop()
    T0 = X.clone()
    T1 = op.theta_affine(X, inv_l, -1.0)
    Y = axpby(T0, float(c[0]), T1, float(c[1]))
    for k in range(2, len(coef)):
        op.cheb_step(T1, T0, Y, inv_l, float(c[k]))
        T0, T1 = T1, T0
    del T0, T1
    return normalize_rows(Y, out=Y)


# ------------------------------------------------- host-resident basis helpers
def _to_device_rows(Xh, start: int, rows: int):
    part = Xh[start:start + rows]
    if isinstance(part, torch.Tensor):
        return part.to(device="cuda", dtype=torch.complex64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(part, np.complex64)).cuda()


def _stream_rotate(C, Xh, rows: int = 8):
    """out[i] = Σ_j C[j, i]·Xh[j] for a host block Xh (interior.py:93-109): row
    chunks of `rows` vectors move to the device, each output chunk accumulates
    there through the combine kernel (out = Cᵀ·S + out) and comes back once.
    C: (m_in, m_out); returns a host complex64 array."""
    C = np.asarray(C)
    m_in, m_out = C.shape
    out = np.empty((m_out,) + tuple(Xh.shape[1:]), np.complex64)
    for so in range(0, m_out, rows):
        ro = min(rows, m_out - so)
        acc = torch.zeros((ro,) + tuple(Xh.shape[1:]), dtype=torch.complex64, device="cuda")
        for si in range(0, m_in, rows):
            Xc = _to_device_rows(Xh, si, rows)
            Cb = torch.from_numpy(np.ascontiguousarray(C[si:si + Xc.shape[0], so:so + ro],
                                                       np.complex64)).cuda()
            combine_into(Cb, Xc, acc, X=acc, alpha=1.0)
            del Xc
        out[so:so + ro] = acc.cpu().numpy()
        del acc
    return out


def gram_hosted(Ah, Bh=None, rows: int = 8):
    """FP64 ⟨A_i, B_j⟩ of host-resident blocks (interior.py:112-126), chunk by
    chunk through the device Gram; B defaults to A."""
    if Bh is None:
        Bh = Ah
    ma, mb = int(Ah.shape[0]), int(Bh.shape[0])
    Gm = np.zeros((ma, mb), np.complex128)
    for si in range(0, ma, rows):
        Ac = _to_device_rows(Ah, si, rows)
        for sj in range(0, mb, rows):
            Bc = _to_device_rows(Bh, sj, rows)
            Gm[si:si + Ac.shape[0], sj:sj + Bc.shape[0]] = gram(Ac, Bc).cpu().numpy()
            del Bc
        del Ac
    return Gm


def svqb_hosted(Xh, passes: int = 2, rows: int = 8):
    """SVQB of a host block (interior.py:129-138): FP64 Gram, host eigh, keep
    w > 1e-5·max w, rotate with V·w^(−½); dropped directions become zero rows."""
    for _ in range(passes):
        Gm = gram_hosted(Xh, rows=rows)
        Gm = (Gm + Gm.conj().T) / 2
        w, V = np.linalg.eigh(Gm)
        keep = w > SVQB_DROP * max(w.max(), 1e-300)
        scale = np.where(keep, 1.0 / np.sqrt(np.where(keep, w, 1.0)), 0.0)
        Xh = _stream_rotate(V * scale[None, :], Xh, rows)
    return Xh


# ---------------------------------------------------------------- RR extract
def gram_chunked(A, B, rows: int = 8):
    """⟨A_i, B_j⟩ as host FP64 (interior.py:432-441).  One kernel call on the device."""
    return gram(A, B).cpu().numpy()


def ortho_block_lowmem(X, passes: int = 2, rows: int = 8):
    """SVQB (interior.py:444-466); identical to ortho_block on a device-resident basis."""
    return ortho_block(X, passes=passes)


def rr_extract(op, X, theta_chunk: int = 8):
    """Orthonormalise X, Rayleigh–Ritz on Θ, return (lam, Xr, relres) (interior.py:37-88).

    SVQB → A = XᴴΘX (FP64) → dead rows +1e8 → FP64 eigh → rotate → normalised
    Rayleigh quotient ⟨x,Θx⟩/⟨x,x⟩ and rn = ‖Θx − λx‖/‖Θx‖ from a fresh Θ
    pass.  Dead rows come back with λ = rn = +inf.  Sorted by λ (eigh order)."""
    if isinstance(X, list):
        X = X.pop()
    X = ortho_block(X)
    HX = op.theta(X)
    A = gram_herm(X, HX)
    dead = rownorm(X) < 0.5
    A = (A + A.conj().T) / 2 + torch.diag(torch.where(dead, 1e8 * dead_spread(A.shape[0], A.device),
                                                      0.0).to(A.dtype))
    wA, VA = eigh(A)
    Xr = torch.empty_like(X)
    combine_into(VA, X, Xr)
    del X, HX
    Hr = op.theta(Xr)
    xn2 = rownorm2(Xr)
    lam_d = rowdot(Xr, Hr).real / torch.clamp(xn2, min=1e-30)
    R = row_axpby(Hr, torch.ones_like(lam_d), Xr, -lam_d)
    rn_d = rownorm(R) / torch.clamp(rownorm(Hr), min=1e-30)
    del Hr, R
    bad = (torch.sqrt(xn2) < 0.5).cpu().numpy()
    lam = np.where(bad, np.inf, lam_d.cpu().numpy())
    rn = np.where(bad, np.inf, rn_d.cpu().numpy())
    return lam, Xr, rn


def rr_extract_hosted(op, Xh, theta_chunk: int = 8):
    """rr_extract for a host array input (interior.py:141-178); runs on the device."""
    lam, Xr, rn = rr_extract(op, Xh if isinstance(Xh, torch.Tensor) else torch.as_tensor(Xh).cuda(),
                             theta_chunk)
    return lam, Xr, rn


# ---------------------------------------------------------------- filtered SI
def _random_start(op, m: int, seed: int):
    return op.random_block(m, seed, kind="normal")


def bandpass_subspace(op, lam_lo: float, lam_hi: float, lam_max: float, m: int = 64,
                      degree: int = 2000, max_outer: int = 8, res_tol: float = 1e-4,
                      target_count: int | None = None, chunk: int = 16, theta_chunk: int = 8,
                      seed: int = 0, X0=None, verbose: bool = True):
    """Filtered subspace iteration with the bandpass polynomial (interior.py:357-427).

    Each outer: filter (degree fused steps per vector, `chunk` vectors per
    launch set), RR on Θ, count converged pairs in [lam_lo, lam_hi]; stop on
    `target_count` or when the converged count stops growing (after outer 2)."""
    if X0 is not None:
        X = X0.pop() if isinstance(X0, list) else X0
        m = int(X.shape[0])
    else:
        X = _random_start(op, m, seed)
    coef = bandpass_coeffs(lam_lo, lam_hi, lam_max, degree)
    n_theta = 0
    t0 = time.perf_counter()
    hist = []
    best = -1
    lam = Xr = rn = None
    for outer in range(max_outer):
        outs = []
        for s in range(0, X.shape[0], chunk):
            outs.append(_apply_cheb_poly(op, X[s:s + chunk], coef, lam_max))
            n_theta += degree * int(X[s:s + chunk].shape[0])
        del X
        X = torch.cat(outs, dim=0)
        del outs
        lam, Xr, rn = rr_extract(op, X, theta_chunk)
        del X
        n_theta += 2 * m
        inwin = (lam >= lam_lo) & (lam <= lam_hi) & np.isfinite(lam)
        conv = inwin & (rn < res_tol)
        hist.append({"outer": outer, "in_window": int(inwin.sum()), "converged": int(conv.sum()),
                     "median_res_inwin": float(np.median(rn[inwin])) if inwin.any() else np.nan})
        if verbose:
            print(f"  bandpass outer {outer}: in-window {inwin.sum():3d} converged "
                  f"{conv.sum():3d} med-res {hist[-1]['median_res_inwin']:.1e} "
                  f"elapsed {time.perf_counter() - t0:.0f}s", flush=True)
        if target_count is not None and conv.sum() >= target_count:
            break
        if int(conv.sum()) <= best and outer >= 2:
            break
        best = max(best, int(conv.sum()))
        X = Xr
    return Xr, {"method": "bandpass", "degree": degree, "m": m,
                "lam_window": (lam_lo, lam_hi), "theta_applications": n_theta,
                "wall_seconds": time.perf_counter() - t0, "outer_stats": hist,
                "lam": lam, "res": rn}


def bandpass_subspace_hosted(op, lam_lo: float, lam_hi: float, lam_max: float, m: int = 104,
                             degree: int = 3000, max_outer: int = 2, res_tol: float = 1e-4,
                             chunk: int = 8, theta_chunk: int = 8, seed: int = 0, X0h=None,
                             verbose: bool = True):
    """interior.py:181-238 signature; the basis is device-resident on MI355X.

    Unlike bandpass_subspace it has no stagnation stop: it stops when every
    in-window pair is converged (interior.py:233-234)."""
    X = _random_start(op, m, seed) if X0h is None else (
        X0h if isinstance(X0h, torch.Tensor) else torch.as_tensor(np.asarray(X0h)).cuda())
    m = int(X.shape[0])
    coef = bandpass_coeffs(lam_lo, lam_hi, lam_max, degree)
    n_theta = 0
    t0 = time.perf_counter()
    hist = []
    lam = rn = None
    for outer in range(max_outer):
        X = torch.cat([_apply_cheb_poly(op, X[s:s + chunk], coef, lam_max)
                       for s in range(0, m, chunk)], dim=0)
        n_theta += degree * m
        lam, X, rn = rr_extract(op, X, theta_chunk)
        n_theta += 2 * m
        inwin = (lam >= lam_lo) & (lam <= lam_hi) & np.isfinite(lam)
        conv = inwin & (rn < res_tol)
        hist.append({"outer": outer, "in_window": int(inwin.sum()), "converged": int(conv.sum()),
                     "median_res_inwin": float(np.median(rn[inwin])) if inwin.any() else np.nan})
        if verbose:
            print(f"  bandpass-hosted outer {outer}: in-window {inwin.sum():3d} converged "
                  f"{conv.sum():3d} elapsed {time.perf_counter() - t0:.0f}s", flush=True)
        if inwin.any() and conv.sum() == inwin.sum():
            break
    return X, {"method": "bandpass_hosted", "degree": degree, "m": m,
               "lam_window": (lam_lo, lam_hi), "theta_applications": n_theta,
               "wall_seconds": time.perf_counter() - t0, "outer_stats": hist,
               "lam": lam, "res": rn}


def two_stage(op, lam_lo: float, lam_hi: float, lam_max: float, m: int = 104,
              build_degree: int = 3000, build_outers: int = 2, polish_degree: int = 12000,
              polish_outers: int = 4, tol: float = 1e-4, chunk: int = 16, seed: int = 0,
              verbose: bool = True, state_path=None, resume: bool = False):
    """Two-stage bandpass ChebSI protocol of scripts/run_interior.py:154-227.

    Build stage: `build_outers` filter+RR outers at `build_degree` from a random
    start; polish stage: up to `polish_outers` at `polish_degree` on the same
    basis, stopping once every in-window pair has rn < tol.  With `state_path`
    the basis and protocol position are checkpointed after every outer step
    (run_interior.py:116-121); `resume=True` continues from that file, with
    its λ_max.  Returns (vals, vecs, rn, info) for the converged in-window
    pairs, sorted by λ."""
    from .io import load_interior_state, save_interior_state
    history = []
    X = None
    lam = rn = None
    n_theta = 0
    start_stage, outer0 = "build", 0
    if resume and state_path is not None:
        st = load_interior_state(state_path)
        if st is not None:
            X = torch.as_tensor(st["X"]).to(op._dev)
            start_stage, outer0 = st["stage"], st["outer"] + 1
            n_theta, lam_max, history = st["n_theta"], st["lam_max"], st["history"]
            if verbose:
                print(f"  resumed: stage={start_stage} next outer={outer0} n_theta={n_theta}",
                      flush=True)
    stages = (("build", build_degree, build_outers), ("polish", polish_degree, polish_outers))
    for stage, degree, outers in stages:
        if stage == "build" and start_stage == "polish":
            continue
        first = outer0 if stage == start_stage else 0
        coef = bandpass_coeffs(lam_lo, lam_hi, lam_max, degree)
        for outer in range(first, outers):
            t0 = time.perf_counter()
            if X is None:
                X = _random_start(op, m, seed)
            pieces = []
            for s in range(0, m, chunk):
                pieces.append(_apply_cheb_poly(op, X[s:s + chunk], coef, lam_max))
                if verbose:
                    print(f"  {stage} outer {outer}: filtered rows {s}..{min(s + chunk, m) - 1} "
                          f"[{time.perf_counter() - t0:.1f}s]", flush=True)
            X = torch.cat(pieces, dim=0)
            del pieces
            lam, X, rn = rr_extract(op, X)
            n_theta += degree * m + 2 * m
            inwin = np.isfinite(lam) & (lam >= lam_lo) & (lam <= lam_hi)
            conv = inwin & (rn < tol)
            rec = {"stage": stage, "outer": outer, "in_window": int(inwin.sum()),
                   "converged": int(conv.sum()),
                   "med_res_inwin": float(np.median(rn[inwin])) if inwin.any() else None,
                   "max_res_inwin": float(np.max(rn[inwin])) if inwin.any() else None,
                   "outer_seconds": time.perf_counter() - t0}
            history.append(rec)
            if verbose:
                med = rec["med_res_inwin"]
                print(f"  {stage} outer {outer}: in-window {rec['in_window']:3d} "
                      f"converged {rec['converged']:3d} med-res "
                      f"{med if med is not None else float('nan'):.1e} "
                      f"[{rec['outer_seconds']:.1f}s]", flush=True)
            if state_path is not None:
                save_interior_state(state_path, X, stage, outer, n_theta, lam_max, history)
            if stage == "polish" and inwin.any() and conv.sum() == inwin.sum():
                break
    if X is None:
        X = _random_start(op, m, seed)
    if lam is None:  # resumed past the last outer step: extract from the saved basis
        lam, X, rn = rr_extract(op, X)
        n_theta += 2 * m
    inwin = np.isfinite(lam) & (lam >= lam_lo) & (lam <= lam_hi)
    conv = inwin & (rn < tol)
    idx = np.where(conv)[0][np.argsort(lam[conv], kind="stable")]
    win = np.where(inwin)[0][np.argsort(lam[inwin], kind="stable")]
    return lam[idx], X[torch.as_tensor(idx, device=X.device)], rn[idx], {
        "history": history, "theta_applications": n_theta, "lam_max": lam_max,
        "window_vals": lam[win], "window_res": rn[win]}
