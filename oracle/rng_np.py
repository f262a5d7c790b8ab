"""Oracle: the counter-based random blocks the HIP kernels generate (NumPy).

No reference counterpart: the reference draws its random blocks from JAX
threefry (solver.py:196-205, chebyshev.py:39-44/244-248,
exp_kpm_dos.py:138-140), which cannot be reproduced without JAX.  The build
replaces it with a counter-based generator whose value depends only on
(seed, global row, component, voxel) — so a block is the same whatever the
chunking or batching — and restates it here so CPU and GPU runs start from
the same vectors.

Element e = ((row·2 + comp)·G³ + voxel) gets h = splitmix64(splitmix64(seed) + e·γ),
γ = 0x9E3779B97F4A7C15 (the splitmix64 stream of Steele, Lea & Flood 2014).
  normal:      u1 = (h>>40 + ½)/2²⁴, u2 = ((h>>16)&0xFFFFFF + ½)/2²⁴,
               z = √(−2 ln u1)·(cos 2πu2 + i sin 2πu2)  (Box–Muller, FP64 → complex64)
  rademacher:  z = −1 if bit 63 of h else +1 (real)
Both are multiplied by the Γ mask (|k+G| > 0).

Pinned by (tests/test_oracle_rng.py): splitmix64 against the published first
outputs for seed 1234567; sample moments (mean 0, E|z|² = 2, Rademacher mean
and exact ±1 values); independence from the chunking of rows.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x):
    with np.errstate(over="ignore"):
        z = (np.asarray(x, np.uint64) + _GAMMA).astype(np.uint64)
        z = ((z ^ (z >> np.uint64(30))) * _M1).astype(np.uint64)
        z = ((z ^ (z >> np.uint64(27))) * _M2).astype(np.uint64)
        return z ^ (z >> np.uint64(31))


def _hashes(seed, row0, nrows, n_vox):
    key = splitmix64(np.uint64(seed & 0xFFFFFFFFFFFFFFFF))
    e = (np.uint64(row0) * np.uint64(2 * n_vox)
         + np.arange(nrows * 2 * n_vox, dtype=np.uint64))
    with np.errstate(over="ignore"):
        return splitmix64(key + e * _GAMMA)


def random_block(kind, seed, row0, nrows, grid_size, mask=None):
    """(nrows, 2, G, G, G) complex64 block; ``mask`` is the (G,G,G) Γ mask or None."""
    G = int(grid_size)
    n_vox = G ** 3
    h = _hashes(seed, row0, nrows, n_vox)
    if kind == "normal":
        u1 = ((h >> np.uint64(40)).astype(np.float64) + 0.5) * 2.0 ** -24
        u2 = (((h >> np.uint64(16)) & np.uint64(0xFFFFFF)).astype(np.float64) + 0.5) * 2.0 ** -24
        r = np.sqrt(-2.0 * np.log(u1))
        z = (r * np.cos(2 * np.pi * u2)).astype(np.float32) \
            + 1j * (r * np.sin(2 * np.pi * u2)).astype(np.float32)
    elif kind == "rademacher":
        z = np.where((h >> np.uint64(63)) != 0, -1.0, 1.0).astype(np.float32) + 0j
    else:
        raise ValueError(f"unknown kind {kind!r}")
    z = z.astype(np.complex64).reshape(nrows, 2, G, G, G)
    if mask is not None:
        z = z * np.asarray(mask, np.float32)[None, None]
    return z
