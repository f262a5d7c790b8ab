"""Oracle: the binary "pen-like" rasterizer (NumPy, FP64 geometry).

Restates eigenfns/structure.py:29-151: ``box_size_for_n`` (:29-31) and
``rasterize_penlike`` (:51-151) — z-warp by the aspect ratio, flat-capped
cylinders, voxel-centre membership in FP64, montage (clipped index ranges with
the empty-range fallback of :95-102) or periodic (wrapped assignment, :114-129,
:143-146) index handling.

Pinned by (tests/test_oracle_structure.py): the golden voxel masks written by
tests/golden/make_raster_golden.py from the reference's own module on the
synthetic networks (bit-exact), and the face-straddling rod test of
tests/test_structure.py:61-78 of the reference.
"""
from __future__ import annotations

import numpy as np

D0_UM = 0.8
L_REF_UM = 11.44
N_REF = 1000


def box_size_for_n(n_vertices: int) -> float:
    return (n_vertices / N_REF) ** (1.0 / 3.0) * L_REF_UM


def rasterize_penlike(rods, grid_size, box_size, minor_radius=0.2252, aspect_ratio=2.5,
                      eps_rod=2.9275 ** 2, eps_bg=1.0, periodic=False):
    rods = np.asarray(rods, dtype=np.float64)
    G = int(grid_size)
    s = float(aspect_ratio)
    b = float(minor_radius)
    grid = np.full((G, G, G), eps_bg, dtype=np.float32)
    dx = box_size / G
    coords = (np.arange(G, dtype=np.float64) + 0.5) * dx - box_size / 2.0

    def rng_clip(lo, hi):
        i0 = int(np.searchsorted(coords, lo, side="left"))
        i1 = int(np.searchsorted(coords, hi, side="right") - 1)
        i0, i1 = max(i0, 0), min(i1, G - 1)
        if i1 < i0:
            mid = min(max(int(np.searchsorted(coords, 0.5 * (lo + hi))), 0), G - 1)
            i0 = i1 = mid
        return i0, i1

    def rng_wrap(lo, hi):
        return (int(np.floor((lo + box_size / 2.0) / dx - 0.5)),
                int(np.ceil((hi + box_size / 2.0) / dx - 0.5)))

    pad = b * max(1.0, s) + dx
    for rod in rods:
        p1w, p2w = rod[:3], rod[3:]
        p1u = p1w.copy(); p1u[2] /= s
        p2u = p2w.copy(); p2u[2] /= s
        vu = p2u - p1u
        Lu = float(np.linalg.norm(vu))
        if Lu <= 0.0:
            continue
        nu = vu / Lu
        rng = rng_wrap if periodic else rng_clip
        ix0, ix1 = rng(min(p1w[0], p2w[0]) - pad, max(p1w[0], p2w[0]) + pad)
        iy0, iy1 = rng(min(p1w[1], p2w[1]) - pad, max(p1w[1], p2w[1]) + pad)
        iz0, iz1 = rng(min(p1w[2], p2w[2]) - pad, max(p1w[2], p2w[2]) + pad)
        if periodic:
            gx, gy, gz = (np.arange(ix0, ix1 + 1), np.arange(iy0, iy1 + 1),
                          np.arange(iz0, iz1 + 1))
            X, Y, Z = np.meshgrid((gx + 0.5) * dx - box_size / 2.0,
                                  (gy + 0.5) * dx - box_size / 2.0,
                                  (gz + 0.5) * dx - box_size / 2.0, indexing="ij")
        else:
            X, Y, Z = np.meshgrid(coords[ix0:ix1 + 1], coords[iy0:iy1 + 1],
                                  coords[iz0:iz1 + 1], indexing="ij")
        RX, RY, RZ = X - p1u[0], Y - p1u[1], Z / s - p1u[2]
        t = RX * nu[0] + RY * nu[1] + RZ * nu[2]
        rX, rY, rZ = RX - t * nu[0], RY - t * nu[1], RZ - t * nu[2]
        mask = (t >= 0.0) & (t <= Lu) & (rX * rX + rY * rY + rZ * rZ <= b * b)
        if mask.any():
            if periodic:
                ii, jj, kk = np.nonzero(mask)
                grid[gx[ii] % G, gy[jj] % G, gz[kk] % G] = eps_rod
            else:
                grid[ix0:ix1 + 1, iy0:iy1 + 1, iz0:iz1 + 1][mask] = eps_rod
    return grid
