"""Oracle: dense Θ by probing with unit vectors (FP64).

Follows the reference's dense-matrix oracle, tests/test_operator.py:16-24
(column i of M is Θ applied to the i-th unit vector), batched over columns.
The small-grid eigenvalue reference of tests/test_solver.py:20-33 and the
FP64 dense solve of a 16³ grid are built from it.
"""
from __future__ import annotations

import numpy as np


def dense_theta(op, batch=256):
    G = op.grid_size
    n = 2 * G ** 3
    M = np.zeros((n, n), np.complex128)
    for s in range(0, n, batch):
        c = min(batch, n - s)
        E = np.zeros((c, n), np.complex128)
        E[np.arange(c), s + np.arange(c)] = 1.0
        M[:, s:s + c] = op.theta(E.reshape(c, 2, G, G, G)).reshape(c, n).T
    return M


def dense_eigh(op, drop_gamma=True):
    """Sorted eigenvalues/eigenvectors of the Hermitian part; at Γ the two zeroed
    slots are dropped (tests/test_solver.py:32-33)."""
    M = dense_theta(op)
    w, V = np.linalg.eigh((M + M.conj().T) / 2)
    if drop_gamma and np.allclose(op.k_frac, 0.0):
        # the two structurally-zero Γ slots carry exact zero eigenvalues
        zero_slots = np.argsort(np.abs(w))[:2]
        keep = np.setdiff1d(np.arange(len(w)), zero_slots)
        w, V = w[keep], V[:, keep]
    order = np.argsort(w)
    return w[order], V[:, order]
