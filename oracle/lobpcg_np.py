"""Oracle: deflated block LOBPCG (NumPy, FP64 vectors and FP64 small algebra).

Restates eigenfns/solver.py:136-382 (``lobpcg_blocks``) step for step:
guard warm start from the previous block's carry (solver.py:217-220), deflation
against the locked set each iteration (:253, :305), SVQB (:113-126), fixed-shape
P whitening (:258-274), Rayleigh–Ritz with the dead-row penalty
max(1e4, 100·max|diag A|) (:295-301), lock the lowest m−guard pairs (:331-340),
final monotone sort (:371-381).  Random rows come from ``rng_np`` instead of
JAX threefry (row offset = block index × m), so iterates are not the
reference's, only the converged eigenpairs are comparable.

Pinned by (tests/test_oracle_spectral.py: test_lobpcg_matches_dense,
test_lobpcg_homogeneous_analytic): the dense FP64 spectrum at G=8
(tests/test_solver.py:36-48 of the reference, tightened to 1e-10), the
homogeneous analytic spectrum, monotonicity and orthonormality of the locked
set.
"""
from __future__ import annotations

import numpy as np

from .linalg_np import (RR_DEAD_PENALTY, SVQB_DROP, combine, deflate, flat, gram,
                        ortho_block, rownorm)
from .rng_np import random_block


def lobpcg_blocks(op, nev, m=96, guard=48, tol=1e-4, maxit=400, seed=0,
                  normalized_rq=False, verbose=False):
    """Returns (vals (nev,), vecs (nev,2,G,G,G) complex128, stats dict)."""
    G = op.grid_size
    mask = op.mask
    dtype = np.complex128
    locked = np.zeros((0, 2, G, G, G), dtype)
    locked_vals = np.empty((0,))
    carry = None
    n_theta = 0
    block_idx = 0
    rounds = []

    def rand_rows(n):
        Z = random_block("normal", seed, block_idx * m, m, G, mask)[:n].astype(dtype)
        return Z / np.maximum(rownorm(Z), 1e-30)[:, None, None, None, None]

    def rq(X, HX):
        lam = np.real(np.sum(flat(X).conj() * flat(HX), axis=1))
        if normalized_rq:
            lam = lam / np.maximum(rownorm(X) ** 2, 1e-30)
        return lam

    while locked_vals.size < nev:
        L = locked if locked.shape[0] else None
        if carry is not None and carry.shape[0] > 0:
            X = np.concatenate([carry, rand_rows(m - carry.shape[0])], axis=0)
        else:
            X = rand_rows(m)
        block_idx += 1
        X = ortho_block(deflate(X, L))
        HX = op.theta(X).astype(dtype); n_theta += m
        P = None
        n_lock = min(m - guard, nev - locked_vals.size)
        it = 0
        for it in range(maxit):
            lam = rq(X, HX)
            R = HX - lam[:, None, None, None, None] * X
            rn = rownorm(R) / np.maximum(rownorm(HX), 1e-30)
            dead_rows = rownorm(X) < 0.5
            lam_h = np.where(dead_rows, np.inf, lam)
            rn = np.where(dead_rows, np.inf, rn)
            order = np.argsort(lam_h)
            if (rn[order[:n_lock]] < tol).all():
                break
            W = op.precondition(R).astype(dtype) * mask[None, None]
            W = deflate(W, L)
            for _ in range(2):
                W = W - combine(gram(X, W), X)
            W = ortho_block(W)
            if P is not None:
                for _ in range(2):
                    P = P - combine(gram(X, P), X) - combine(gram(W, P), W)
                Gp = gram(P, P)
                Gp = (Gp + Gp.conj().T) / 2
                wp, Vp = np.linalg.eigh(Gp)
                keepp = wp > SVQB_DROP * max(wp.max(), 1e-300)
                if not keepp.any():
                    P = None
                else:
                    scale = np.where(keepp, 1.0 / np.sqrt(np.where(keepp, wp, 1.0)), 0.0)
                    P = combine(Vp * scale[None, :], P)
            blocks = [X, W] if P is None else [X, W, P]
            HB = [HX] + [op.theta(B).astype(dtype) for B in blocks[1:]]
            n_theta += m * (len(blocks) - 1)
            S = np.concatenate(blocks, axis=0)
            A = gram(S, np.concatenate(HB, axis=0))
            A = (A + A.conj().T) / 2
            dead = rownorm(S) < 0.5
            penalty = max(RR_DEAD_PENALTY, 100.0 * float(np.abs(np.diag(A)).max()))
            A[np.diag_indices_from(A)] += np.where(dead, penalty, 0.0)
            wA, VA = np.linalg.eigh(A)
            C = VA[:, :m]
            Xn = combine(C, *blocks)
            Xn = deflate(Xn, L)
            Xn = Xn / np.maximum(rownorm(Xn), 1e-20)[:, None, None, None, None]
            Cp = C.copy()
            Cp[:m] = 0
            P = combine(Cp, *blocks)
            P = P / np.maximum(rownorm(P), 1e-20)[:, None, None, None, None]
            X = Xn
            HX = op.theta(X).astype(dtype); n_theta += m
        HX = op.theta(X).astype(dtype); n_theta += m
        lam = rq(X, HX)
        Rf = HX - lam[:, None, None, None, None] * X
        rn = rownorm(Rf) / np.maximum(rownorm(HX), 1e-30)
        xnorm = rownorm(X)
        lam = np.where(xnorm < 0.5, np.inf, lam)
        rn = np.where(xnorm < 0.5, np.inf, rn)
        order = np.argsort(lam)
        if not np.isfinite(lam[order[:n_lock]]).all():
            raise RuntimeError("rank collapse")
        lock_idx = order[:n_lock]
        carry = X[order[n_lock:]]
        locked = np.concatenate([locked, X[lock_idx]], axis=0)
        locked_vals = np.concatenate([locked_vals, lam[lock_idx]])
        rounds.append({"iters": it + 1, "locked": int(locked_vals.size),
                       "max_res_locked": float(rn[lock_idx].max())})
        if verbose:
            print(f"  oracle locked {locked_vals.size}/{nev} iters {it + 1}", flush=True)
    order = np.argsort(locked_vals, kind="stable")
    return locked_vals[order][:nev], locked[order][:nev], {
        "theta_applications": n_theta, "rounds": rounds}
