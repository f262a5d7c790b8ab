"""Deflated block LOBPCG for the transverse Maxwell operator, device-resident.

Drop-in for eigenfns.solver (reference eigenfns/solver.py).  The algorithm is
the reference's (solver.py:136-382): guard warm start, re-deflation every
iteration, SVQB with the 1e-5 drop threshold, fixed-shape P whitening,
fresh Θ(W), Θ(P) for the Rayleigh–Ritz matrix, dead-row penalty, lock the
lowest m−guard pairs, monotone final sort.  What changes is where things
live: [X; W; P] and [ΘX; ΘW; ΘP] are two contiguous device blocks, so the
Rayleigh–Ritz matrix is one Gram kernel and each update one combine kernel;
the locked set stays in HBM (no host streaming).  The small FP64 algebra
works on Gram matrices the device sums in FP64: the orthonormalisation of W
and P is a pivoted Cholesky-QR on the device by default (ortho="cholqr"; the
reference's SVQB eigendecompositions with ortho="svqb"), and the
Rayleigh–Ritz eigenproblem (at most 3m × 3m) runs in host LAPACK as in the
reference (solver.py:119-121, 263-265, 301), or on the device (the
library's QL kernel up to 96 × 96, rocSOLVER above) with
EFN_SMALL_DENSE=device and always when the iteration is graph-captured
(graph=True).  Rayleigh
quotients are normalised by ‖x‖² (the reference's interior.py:69-74 fix)
unless ``normalized_rq=False``.

One deliberate difference: a block whose lockable residuals have reached
the FP32 floor just above `tol` (below STAGNATION_FACTOR·tol, no better by
10 % over the last `stagnation` iterations) is locked there instead of
iterating to `maxit`; the block's entry in ``SolveStats.rounds`` records
``converged`` (worst locked residual < tol) and ``stagnated``, and a
RuntimeWarning names the blocks that locked above tol.  ``stagnation=0``
restores the reference's run-to-maxit behaviour.
"""
from __future__ import annotations

import gc
import time
import warnings
from dataclasses import dataclass, field

import numpy as np
import torch

from .linalg import (COMBINE_IN_PLACE_MAX, DEFLATE_CHUNK, RR_DEAD_PENALTY, SVQB_DROP, apply_chunked,
                     combine, combine_into, combine_multi_split, dead_spread, deflate, deflate_chunk_rows,
                     eigh, eigh_batched_device, QL_MAX_N,
                     gram, gram_herm, normalize_rows, ortho_block, ortho_coeffs, row_axpby,
                     row_axpby_norm_strided, rowdot, rownorm, rownorm2, rowstats_strided,
                     svqb_coeffs, ORTHO_METHODS)

__all__ = ["gram", "combine", "deflate", "ortho_block", "apply_chunked", "deflate_chunk_rows",
           "SolveStats", "StagnationLock", "lobpcg_blocks", "lobpcg_blocks_resumable", "_flat",
           "_SVQB_DROP", "_RR_DEAD_PENALTY", "_DEFLATE_CHUNK"]

_SVQB_DROP = SVQB_DROP
_RR_DEAD_PENALTY = RR_DEAD_PENALTY
_DEFLATE_CHUNK = DEFLATE_CHUNK


def _flat(X):
    return X.reshape(X.shape[0], -1)


@dataclass
class SolveStats:
    theta_applications: int = 0
    rounds: list = field(default_factory=list)
    wall_seconds: float = 0.0
    precond_applications: int = 0
    iterations: int = 0

    @property
    def converged(self) -> bool:
        """Every locked block met the residual tolerance."""
        return all(r.get("converged", True) for r in self.rounds)


# A block stops at the FP32 residual floor when its worst lockable residual is
# below STAGNATION_FACTOR·tol and the best of the last `stagnation` iterations
# is not 10 % better than the best before them.
STAGNATION_WINDOW = 20
STAGNATION_FACTOR = 10.0


class StagnationLock:
    """Per-block record of the worst lockable residual; ``stalled`` says when
    the block has stopped improving at the FP32 floor near `tol`."""

    def __init__(self, tol: float, window: int = STAGNATION_WINDOW,
                 factor: float = STAGNATION_FACTOR):
        self.tol = float(tol)
        self.window = int(window)
        self.factor = float(factor)
        self.hist: list = []

    def reset(self) -> None:
        self.hist = []

    def stalled(self, worst: float) -> bool:
        self.hist.append(float(worst))
        w = self.window
        if w <= 0 or len(self.hist) <= w or not worst < self.factor * self.tol:
            return False
        return min(self.hist[-w:]) > 0.9 * min(self.hist[:-w])


def _deflate_inplace(Y, L, nl):
    """Y ← Y − L(LᴴY) against the first nl locked rows, in place."""
    if L is None or nl == 0:
        return Y
    Lc = L[:nl]
    Cg = gram(Lc, Y)
    combine_into(-Cg, Lc, Y, X=Y, alpha=1.0)
    return Y


def _svqb_into(Y, tmp, drop=SVQB_DROP, passes=2, method="svqb"):
    """SVQB of Y (two sweeps), result written back into Y; tmp has Y's shape
    (unused when the combine can run in place).  method="cholqr": Cholesky-QR
    coefficients in place of the eigendecomposition."""
    for _ in range(passes):
        Cw = ortho_coeffs(gram_herm(Y), drop, method)
        _combine_back(Cw, Y, tmp)
    return Y


def _combine_back(C, Y, tmp):
    """Y ← Cᵀ·Y: in place up to COMBINE_IN_PLACE_MAX rows, else through tmp."""
    if int(Y.shape[0]) <= COMBINE_IN_PLACE_MAX:
        combine_into(C, Y, Y)
    else:
        combine_into(C, Y, tmp)
        Y.copy_(tmp)


def lobpcg_blocks(op, nev: int, m: int = 96, guard: int = 48, tol: float = 1e-4,
                  maxit: int = 400, seed: int = 0, precond_shift: str = "median",
                  verbose: bool = True, log_every: int = 0, initial_locked=None,
                  initial_vals=None, on_block=None, locked_storage: str = "auto",
                  theta_chunk: int = 0, normalized_rq: bool = True,
                  stagnation: int = STAGNATION_WINDOW, ortho: str = "cholqr", graph: bool = False):
    """Bottom-up deflated block LOBPCG.  Returns (eigenvalues, locked_vectors, stats).

    Locks the lowest (m − guard) Ritz pairs of each block once their relative
    residuals pass `tol` (or, with `stagnation` > 0, once they stall at the
    FP32 floor below STAGNATION_FACTOR·tol); the guard Ritz vectors seed the
    next block.
    `initial_locked`/`initial_vals` resume a partial solve;
    `on_block(n_locked_total, lam_block, X_block)` fires after each lock.
    `locked_storage` and `theta_chunk` are accepted for API compatibility: the
    locked set is always HBM-resident and Θ is chunked by the operator.
    `ortho`: "cholqr" (default: pivoted FP64 Cholesky-QR on the device for the
    SVQB steps and the P whitening, no host round trip; the Rayleigh–Ritz
    eigensolve is unchanged) or "svqb" (the reference's eigendecomposition-based
    orthonormalisation, solver.py:113-126).
    `graph`: the steady-state iteration (everything between two residual
    read-backs: preconditioner, projections, Cholesky-QR, Θ, the Rayleigh–Ritz
    matrix and its eigensolve on the device, the update, Θ(X) and the next
    residuals) is captured once per block as one HIP graph and replayed — one
    submission per iteration instead of some forty launches, no host sync but
    the residual read-back — when 3m ≤ 96 (the library's QL eigensolver);
    above that as two graphs around rocSOLVER's eigensolve (uncaptured);
    needs ortho="cholqr"; results are bitwise those of the uncaptured loop
    with device eigensolves."""
    if m <= guard:
        raise ValueError("m must exceed guard")
    if graph and ortho != "cholqr":
        raise ValueError("graph=True needs ortho='cholqr' (no host eigensolve inside the iteration)")
    if ortho not in ORTHO_METHODS:
        raise ValueError(f"ortho must be one of {ORTHO_METHODS}, got {ortho!r}")
    G = op.grid_size
    shape = (2, G, G, G)
    dev = op.random_block(0, seed).device
    n_lock_per = m - guard
    cap = int(np.ceil(nev / n_lock_per)) * n_lock_per + m
    n_init = 0
    if initial_locked is not None:
        init = torch.as_tensor(initial_locked) if not isinstance(initial_locked, torch.Tensor) \
            else initial_locked
        n_init = int(init.shape[0])
        cap += n_init
    locked_buf = torch.empty((cap,) + shape, dtype=torch.complex64, device=dev)
    n_locked_now = 0
    if n_init:
        locked_buf[:n_init] = init.to(device=dev, dtype=torch.complex64)
        n_locked_now = n_init
    locked_vals = (np.empty((0,)) if initial_vals is None
                   else np.asarray(initial_vals, np.float64))
    if verbose:
        print(f"  locked storage: device ({cap * 16 * G ** 3 / 2 ** 30:.1f} GiB capacity, "
              f"{cap} rows)", flush=True)
    # fixed device buffers
    S = torch.empty((3 * m,) + shape, dtype=torch.complex64, device=dev)    # [X; W; P]
    HS = torch.empty_like(S)                                                 # [ΘX; ΘW; ΘP]
    # [T1; T2]: the new X and P rows of one Rayleigh–Ritz update come from one
    # combine over S (2m outputs, S read once)
    T12 = torch.empty((2 * m,) + shape, dtype=torch.complex64, device=dev)
    T1, T2 = T12[:m], T12[m:]
    X, W, P = S[:m], S[m:2 * m], S[2 * m:]
    HX = HS[:m]
    n = 2 * G ** 3  # complex entries per vector
    stats = SolveStats()
    t0 = time.perf_counter()
    carry = None
    block_idx = 0

    def rq(Xb, HXb):
        lam = rowdot(Xb, HXb).real
        if normalized_rq:
            lam = lam / torch.clamp(rownorm2(Xb), min=1e-30)
        return lam

    while locked_vals.size < nev:
        L = locked_buf if n_locked_now > 0 else None
        nl = n_locked_now
        # X = [carry, masked random unit rows]
        nc = 0 if carry is None else int(carry.shape[0])
        if nc:
            X[:nc] = carry
        if nc < m:
            Z = op.random_block(m, seed, kind="normal", row0=block_idx * m)[:m - nc]
            normalize_rows(Z, out=X[nc:])
        block_idx += 1
        _deflate_inplace(X, L, nl)
        _svqb_into(X, T1, method=ortho)
        op.theta_into(X, HX)
        stats.theta_applications += m
        have_p = False
        n_lock = min(n_lock_per, nev - locked_vals.size)
        stag = StagnationLock(tol, stagnation)
        stalled = False
        it = 0

        def residual_stats():
            # one read of X and ΘX for ⟨x,Θx⟩, |x|², |Θx|²; R = ΘX − λX (into T2)
            # with |R|² in the same pass (solver.py:228-238)
            st = rowstats_strided(X, HX, m, m, m, m, n)
            lam = st[:, 0] / torch.clamp(st[:, 2], min=1e-30) if normalized_rq else st[:, 0]
            rr2 = row_axpby_norm_strided(HX, m, torch.ones_like(lam), X, m, -lam, m, m, n, T2, m)
            return torch.stack([lam, st[:, 2], rr2, st[:, 3]])

        def small_eigh(A):
            # the update reads the lowest m eigenvectors only; asking the
            # device solver for no more also skips the dead-row penalty cluster
            if graph:
                w, V = eigh_batched_device(((A + A.conj().T) / 2)[None], lowest=m)
                return w[0], V[0]
            return eigh(A)

        def iteration_tail(have_p):
            """The device work of one iteration after the host's convergence
            check, ending with the next iteration's residual statistics."""
            A, nb = tail_rr_matrix(have_p)
            return tail_update(small_eigh(A)[1], nb)

        def tail_rr_matrix(have_p):
            """Preconditioner, projections, orthonormalisation, Θ of W and P and
            the penalised Rayleigh–Ritz matrix."""
            R = T2
            # W = K⁻¹R, deflated, projected out of X twice, SVQB (solver.py:251-257)
            op.precondition_into(R, W)
            _deflate_inplace(W, L, nl)
            for _ in range(2):
                combine_into(-gram(X, W), X, W, X=W, alpha=1.0)
            _svqb_into(W, T1, method=ortho)
            # P projected out of [X, W] twice, fixed-shape whitening (solver.py:258-274)
            if have_p:
                XW = S[:2 * m]
                for _ in range(2):
                    combine_into(-gram(XW, P), XW, P, X=P, alpha=1.0)
                # a P with no direction above the drop threshold becomes zero
                # rows, which the dead-row penalty below removes from the
                # Rayleigh–Ritz: the reference's [X, W] basis without a host
                # read of the kept count
                if ortho == "cholqr":
                    _combine_back(ortho_coeffs(gram_herm(P), SVQB_DROP, ortho), P, T1)
                else:
                    wp, Vp = eigh(gram_herm(P))
                    keepp = wp > SVQB_DROP * torch.clamp(wp.max(), min=1e-300)
                    scale = torch.where(keepp, 1.0 / torch.sqrt(torch.where(keepp, wp,
                                                                            torch.ones_like(wp))),
                                        torch.zeros_like(wp))
                    _combine_back(Vp * scale[None, :], P, T1)
            nb = (3 if have_p else 2) * m
            Sb, HSb = S[:nb], HS[:nb]
            op.theta_into(Sb[m:], HSb[m:])
            A = gram_herm(Sb, HSb)
            A = (A + A.conj().T) / 2
            dead = rownorm(Sb) < 0.5
            # max(1e4, 100·max|diag A|) (solver.py:299), kept on the device
            penalty = torch.clamp(100.0 * torch.diagonal(A).abs().max(), min=RR_DEAD_PENALTY)
            A = A + torch.diag(torch.where(dead, penalty * dead_spread(nb, A.device), 0.0).to(A.dtype))
            return A, nb

        def tail_update(VA, nb):
            """The Rayleigh–Ritz update of X and P, deflation, Θ(X) and the next
            residual statistics."""
            Sb = S[:nb]
            # Xn = S·c/‖S·c‖, Pn = S·c'/‖S·c'‖ (c' = c with its X rows zeroed;
            # solver.py:298-307).  S is orthonormal by construction and dead rows
            # get zero weight, so ‖S·c‖ = ‖c‖ to FP32 rounding: the norms are
            # folded into the coefficients and one split combine writes both
            # updates into S's X and P rows (in place when 2m ≤ 64).
            Cx = VA[:, :m]
            Cp = Cx.clone()
            Cp[:m] = 0
            C2 = torch.cat([Cx / torch.clamp(torch.linalg.vector_norm(Cx, dim=0), min=1e-20),
                            Cp / torch.clamp(torch.linalg.vector_norm(Cp, dim=0), min=1e-20)], dim=1)
            if 2 * m <= COMBINE_IN_PLACE_MAX:
                combine_multi_split(C2[None], [Sb], [X], [P], m)
            else:
                combine_multi_split(C2[None], [Sb], [T1], [T2], m)
                X.copy_(T1)
                P.copy_(T2)
            _deflate_inplace(X, L, nl)
            op.theta_into(X, HX)
            return residual_stats()

        # steady-state iterations (P present) of this block: one eager run
        # (sizes every scratch buffer), then one capture, then replays.  Up to
        # 3m = 96 the whole iteration is one graph (the library's QL
        # eigensolver allocates nothing); above, rocSOLVER (which copies host
        # data inside a call and so cannot be captured) runs uncaptured
        # between two graphs on the same stream, without a host sync
        g, g_stats, warm = None, None, False
        one_graph = 3 * m <= QL_MAX_N
        sdev = residual_stats()
        for it in range(maxit):
            host = sdev.cpu().numpy()
            lam_h = host[0]
            rn_h = np.sqrt(host[2]) / np.maximum(np.sqrt(host[3]), 1e-30)
            dead_rows = np.sqrt(host[1]) < 0.5
            lam_h = np.where(dead_rows, np.inf, lam_h)
            rn_h = np.where(dead_rows, np.inf, rn_h)
            order = np.argsort(lam_h)
            worst = float(rn_h[order[:n_lock]].max())
            if worst < tol:
                break
            if stag.stalled(worst):
                stalled = True
                break
            if log_every and it % log_every == 0:
                leak = float(gram(L[:nl], X).abs().max().item()) if nl else 0.0
                lo = np.sort(lam_h)
                print(f"    it {it:3d}  lam[0,{n_lock - 1},{m - 1}] = {lo[0]:.4f} "
                      f"{lo[n_lock - 1]:.4f} {lo[-1]:.4f}  res q50/q90 "
                      f"{np.quantile(rn_h, 0.5):.1e}/{np.quantile(rn_h, 0.9):.1e}  "
                      f"leak {leak:.1e}", flush=True)
            if graph and have_p and warm:
                if g is None and one_graph:
                    g = (torch.cuda.CUDAGraph(),)
                    with torch.cuda.graph(g[0]):
                        g_stats = iteration_tail(True)
                elif g is None:
                    g = (torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph())
                    with torch.cuda.graph(g[0]):
                        g_A, g_nb = tail_rr_matrix(True)
                    g_V = torch.empty((g_A.shape[0], m), dtype=g_A.dtype, device=g_A.device)
                    with torch.cuda.graph(g[1]):
                        g_stats = tail_update(g_V, g_nb)
                g[0].replay()
                if not one_graph:
                    g_V.copy_(small_eigh(g_A)[1])
                    g[1].replay()
                sdev = g_stats
            else:
                sdev = iteration_tail(have_p)
                warm = warm or have_p
            stats.precond_applications += m
            stats.theta_applications += (2 if have_p else 1) * m + m
            have_p = True
            stats.iterations += 1
        g = g_stats = g_A = g_V = None
        # honest final values with a fresh ΘX (solver.py:316-330)
        op.theta_into(X, HX)
        stats.theta_applications += m
        lam = rq(X, HX)
        Rf = row_axpby(HX, torch.ones_like(lam), X, -lam, out=T2)
        rn = (rownorm(Rf) / torch.clamp(rownorm(HX), min=1e-30)).cpu().numpy()
        lam = lam.cpu().numpy()
        xnorm = rownorm(X).cpu().numpy()
        lam = np.where(xnorm < 0.5, np.inf, lam)
        rn = np.where(xnorm < 0.5, np.inf, rn)
        order = np.argsort(lam)
        if not np.isfinite(lam[order[:n_lock]]).all():
            raise RuntimeError(
                f"block produced only {int(np.isfinite(lam).sum())} live Ritz pairs but needs "
                f"{n_lock} to lock — rank collapse")
        lock_idx = torch.as_tensor(order[:n_lock], device=dev)
        carry_idx = torch.as_tensor(order[n_lock:], device=dev)
        locked_buf[n_locked_now:n_locked_now + n_lock] = X[lock_idx]
        carry = X[carry_idx].clone()
        n_locked_now += n_lock
        locked_vals = np.concatenate([locked_vals, lam[order[:n_lock]]])
        worst = float(rn[order[:n_lock]].max())
        stats.rounds.append({"iters": it + 1, "locked": int(locked_vals.size),
                             "max_res_locked": worst, "converged": bool(worst < tol),
                             "stagnated": stalled})
        if on_block is not None:
            on_block(int(locked_vals.size), lam[order[:n_lock]],
                     locked_buf[n_locked_now - n_lock:n_locked_now].cpu().numpy())
        if verbose:
            print(f"  locked {locked_vals.size:4d}/{nev} (+{n_lock})  iters {it + 1:3d}  "
                  f"lam [{lam[order[0]]:.4f}..{lam[order[n_lock - 1]]:.4f}]  "
                  f"worst res {rn[order[:n_lock]].max():.1e}  "
                  f"elapsed {time.perf_counter() - t0:6.1f}s", flush=True)
    stats.wall_seconds = time.perf_counter() - t0
    above = [i for i, r in enumerate(stats.rounds) if not r["converged"]]
    if above:
        warnings.warn(f"lobpcg_blocks: {len(above)} of {len(stats.rounds)} blocks locked with "
                      f"residuals above tol={tol:g} (blocks {above}; worst "
                      f"{max(stats.rounds[i]['max_res_locked'] for i in above):.2e})",
                      RuntimeWarning, stacklevel=2)
    if (np.diff(locked_vals) < -1e-6 * np.abs(locked_vals[1:])).any():
        worst = float(np.min(np.diff(locked_vals)))
        print(f"WARNING: locked eigenvalues non-monotone (worst step {worst:.3e}) — a band was "
              f"recovered out of order; re-sorting.", flush=True)
    order = np.argsort(locked_vals, kind="stable")
    locked = locked_buf[:n_locked_now][torch.as_tensor(order, device=dev)]
    del S, HS, T12, T1, T2, locked_buf
    gc.collect()
    return locked_vals[order][:nev], locked, stats


def lobpcg_blocks_resumable(op, nev, checkpointer=None, resume=False, **kw):
    """`lobpcg_blocks` with per-block checkpointing and resume (solver.py:385-411).

    `checkpointer` follows the reference's BlockCheckpointer protocol:
    ``load() -> (vals, vecs, n_blocks, meta)`` and
    ``save_block(i, lam_block, X_block, extra=...)``."""
    if checkpointer is None:
        return lobpcg_blocks(op, nev, **kw)
    init_vals = init_vecs = None
    n_blocks = 0
    if resume:
        init_vals, init_vecs, n_blocks, _ = checkpointer.load()
        if init_vals is not None:
            print(f"  resuming: {len(init_vals)} bands from {n_blocks} checkpointed blocks",
                  flush=True)
            if len(init_vals) >= nev:
                return (np.asarray(init_vals)[:nev],
                        torch.as_tensor(np.asarray(init_vecs)[:nev]).cuda(), SolveStats())
    state = {"i": n_blocks}

    def on_block(n_locked_total, lam_block, X_block):
        checkpointer.save_block(state["i"], lam_block, X_block, extra={"n_locked": n_locked_total})
        state["i"] += 1

    return lobpcg_blocks(op, nev, initial_locked=init_vecs, initial_vals=init_vals,
                         on_block=on_block, **kw)
