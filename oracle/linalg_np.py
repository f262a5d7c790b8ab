"""Oracle: block linear algebra of the LOBPCG/ChebSI path (NumPy, FP64).

Restates eigenfns/solver.py:
  SVQB_DROP, RR_DEAD_PENALTY  solver.py:27-28
  gram                        solver.py:35-43
  combine                     solver.py:49-57
  deflate_chunk_rows          solver.py:85-92
  deflate                     solver.py:95-110
  ortho_block                 solver.py:113-126
and rownorms / Rayleigh quotients used at solver.py:228-238, interior.py:75-82.

Pinned by (tests/test_oracle_linalg.py): orthonormality ‖XᴴX − I‖ after
SVQB; rank-dropped rows come back as exact zero rows; the deflation projector
leaves Lᴴ·deflate(X) ≈ 0 and is idempotent; combine equals an explicit sum
of scaled rows; gram equals np.vdot entry by entry.
"""
from __future__ import annotations

import numpy as np

SVQB_DROP = 1e-5
RR_DEAD_PENALTY = 1e4
DEFLATE_CHUNK = 128


def flat(X):
    X = np.asarray(X)
    return X.reshape(X.shape[0], -1)


def gram(A, B):
    """⟨A_i, B_j⟩ = Σ conj(A_i)·B_j, accumulated in FP64."""
    return flat(A).astype(np.complex128).conj() @ flat(B).astype(np.complex128).T


def combine(C, *blocks):
    """out_i = Σ_j C[j, i]·S_j over the concatenated blocks."""
    S = np.concatenate([flat(b) for b in blocks], axis=0)
    out = np.asarray(C).T @ S
    return out.reshape((C.shape[1],) + np.asarray(blocks[0]).shape[1:])


def deflate_chunk_rows(grid_size: int, target_bytes: float = 0.75e9) -> int:
    bytes_per_vec = 2 * grid_size ** 3 * 8
    return max(16, int(target_bytes // bytes_per_vec))


def deflate(X, locked, chunk_rows=DEFLATE_CHUNK):
    """X ← X − L(LᴴX), chunk by chunk over the locked rows (one pass)."""
    if locked is None or np.asarray(locked).shape[0] == 0:
        return X
    X = np.array(X, copy=True)
    n = locked.shape[0]
    for s in range(0, n, chunk_rows):
        chunk = np.asarray(locked[s:s + chunk_rows])
        C = gram(chunk, X)
        X = X - combine(C, chunk).astype(X.dtype)
    return X


def svqb_coeffs(Gm, drop=SVQB_DROP):
    """SVQB whitening coefficients: V·w^{-1/2} with dropped directions zeroed."""
    Gm = (Gm + Gm.conj().T) / 2
    w, V = np.linalg.eigh(Gm)
    keep = w > drop * max(w.max(), 1e-300)
    scale = np.where(keep, 1.0 / np.sqrt(np.where(keep, w, 1.0)), 0.0)
    return V * scale[None, :]


def ortho_block(X, passes=2, drop=SVQB_DROP):
    """SVQB orthonormalisation, `passes` sweeps (solver.py:113-126)."""
    for _ in range(passes):
        C = svqb_coeffs(gram(X, X), drop)
        X = combine(C, X).astype(np.asarray(X).dtype)
    return X


def rownorm(X):
    return np.sqrt(np.real(np.sum(np.abs(flat(X).astype(np.complex128)) ** 2, axis=1)))


def rowdot_re(A, B):
    return np.real(np.sum(flat(A).astype(np.complex128).conj()
                          * flat(B).astype(np.complex128), axis=1))
