"""Rod networks and their binary ε grids (the input stage of the hot path).

Drop-in for eigenfns.structure (reference eigenfns/structure.py:29-155):
``box_size_for_n``, ``load_rods``, ``rasterize_penlike``, ``filling_fraction``.
The rasterizer is the reference's per-rod voxel-centre membership test in
FP64, re-implemented with the same operation order so masks are bit-identical
(pinned against golden masks produced by the reference module itself).

``synthetic_network`` builds the benchmark structures (no reference
counterpart: the reference reads rod files): N points uniform in the periodic
unit cube from ``numpy.random.default_rng(seed)``, each joined to its 4
nearest periodic neighbours; rods are the minimum-image segments (both
periodic images emitted for face-crossing rods, the rod-file convention of
structure.py:37-38), radius = factor × mean unique bond length, box
L = box_size_for_n(N), centred at 0.  Rasterised with ``periodic=True``.
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np

D0_UM = 0.8
L_REF_UM = 11.44
N_REF = 1000


def box_size_for_n(n_vertices: int) -> float:
    """Density-matched cubic box side (µm) for an N-vertex network (structure.py:29-31)."""
    return (n_vertices / N_REF) ** (1.0 / 3.0) * L_REF_UM


def load_rods(path):
    """(rods (R,6) float64, N, box_size) from a 6-column rod file (structure.py:34-48)."""
    path = Path(path)
    rods = np.loadtxt(path)
    if rods.ndim != 2 or rods.shape[1] != 6:
        raise ValueError(f"{path}: expected (R,6) rod endpoints, got {rods.shape}")
    m = re.search(r"N(\d+)", path.name) or re.search(r"ak(\d+)", path.name)
    if not m:
        raise ValueError(f"{path.name}: cannot infer N from filename (need 'N<digits>')")
    n = int(m.group(1))
    return rods, n, box_size_for_n(n)


def _index_range_clipped(coords, lo, hi, G):
    i0 = int(np.searchsorted(coords, lo, side="left"))
    i1 = int(np.searchsorted(coords, hi, side="right") - 1)
    i0, i1 = max(i0, 0), min(i1, G - 1)
    if i1 < i0:  # empty range: fall back to the voxel nearest the midpoint
        mid = min(max(int(np.searchsorted(coords, 0.5 * (lo + hi))), 0), G - 1)
        i0 = i1 = mid
    return i0, i1


def rasterize_penlike(rods, grid_size: int, box_size: float, minor_radius: float = 0.2252,
                      aspect_ratio: float = 2.5, eps_rod: float = 2.9275 ** 2,
                      eps_bg: float = 1.0, periodic: bool = False) -> np.ndarray:
    """Binary 'pen-like' rasterization (structure.py:51-151).

    Cylinders of radius `minor_radius` in the unwarped space z' = z/aspect;
    voxel centres inside any rod (axial 0 ≤ t ≤ L, radial ≤ r, FP64) get
    `eps_rod`.  periodic=False clips index ranges at the box (montage
    convention); periodic=True wraps the assignment."""
    rods = np.asarray(rods, dtype=np.float64)
    G = int(grid_size)
    s = float(aspect_ratio)
    b = float(minor_radius)
    grid = np.full((G, G, G), eps_bg, dtype=np.float32)
    dx = box_size / G
    half = box_size / 2.0
    coords = (np.arange(G, dtype=np.float64) + 0.5) * dx - half
    pad = b * max(1.0, s) + dx
    bb = b * b
    for rod in rods:
        p1w, p2w = rod[:3], rod[3:]
        p1u = p1w.copy()
        p1u[2] /= s
        p2u = p2w.copy()
        p2u[2] /= s
        vu = p2u - p1u
        Lu = float(np.linalg.norm(vu))
        if Lu <= 0.0:
            continue
        nu = vu / Lu
        lo = np.minimum(p1w, p2w) - pad
        hi = np.maximum(p1w, p2w) + pad
        if periodic:
            r0 = [int(np.floor((lo[a] + half) / dx - 0.5)) for a in range(3)]
            r1 = [int(np.ceil((hi[a] + half) / dx - 0.5)) for a in range(3)]
            gi = [np.arange(r0[a], r1[a] + 1) for a in range(3)]
            cx, cy, cz = [(g + 0.5) * dx - half for g in gi]
        else:
            rr = [_index_range_clipped(coords, lo[a], hi[a], G) for a in range(3)]
            cx, cy, cz = [coords[r[0]:r[1] + 1] for r in rr]
        # broadcast instead of meshgrid; same FP64 operations in the same order
        RX = (cx - p1u[0])[:, None, None]
        RY = (cy - p1u[1])[None, :, None]
        RZ = (cz / s - p1u[2])[None, None, :]
        t = RX * nu[0] + RY * nu[1] + RZ * nu[2]
        rX, rY, rZ = RX - t * nu[0], RY - t * nu[1], RZ - t * nu[2]
        mask = (t >= 0.0) & (t <= Lu) & (rX * rX + rY * rY + rZ * rZ <= bb)
        if mask.any():
            if periodic:
                ii, jj, kk = np.nonzero(mask)
                grid[gi[0][ii] % G, gi[1][jj] % G, gi[2][kk] % G] = eps_rod
            else:
                sub = grid[rr[0][0]:rr[0][1] + 1, rr[1][0]:rr[1][1] + 1, rr[2][0]:rr[2][1] + 1]
                sub[mask] = eps_rod
    return grid


def rod_parameters(rods, grid_size: int, box_size: float, minor_radius: float = 0.2252,
                   aspect_ratio: float = 2.5, periodic: bool = False):
    """Per-rod FP64 parameters of the membership test, computed with the same
    NumPy expressions as `rasterize_penlike` (structure.py:93-127): returns
    (rodf (R,8) float64: p1u, unit axis n, L_u, 0; rodi (R,6) int32: first and
    last voxel index per axis).  Zero-length rods are dropped."""
    rods = np.asarray(rods, dtype=np.float64).reshape(-1, 6)
    G = int(grid_size)
    s = float(aspect_ratio)
    b = float(minor_radius)
    dx = box_size / G
    half = box_size / 2.0
    coords = (np.arange(G, dtype=np.float64) + 0.5) * dx - half
    pad = b * max(1.0, s) + dx
    p1u = rods[:, :3].copy()
    p1u[:, 2] /= s
    p2u = rods[:, 3:].copy()
    p2u[:, 2] /= s
    vu = p2u - p1u
    # the reference takes the 1-D norm of each rod's axis (a BLAS dot); the
    # same call per rod keeps L_u, and with it the mask, bit-identical
    Lu = np.array([float(np.linalg.norm(v)) for v in vu], dtype=np.float64)
    keep = Lu > 0.0
    p1u, vu, Lu = p1u[keep], vu[keep], Lu[keep]
    nu = vu / Lu[:, None]
    lo = np.minimum(rods[keep, :3], rods[keep, 3:]) - pad
    hi = np.maximum(rods[keep, :3], rods[keep, 3:]) + pad
    if periodic:
        r0 = np.floor((lo + half) / dx - 0.5).astype(np.int64)
        r1 = np.ceil((hi + half) / dx - 0.5).astype(np.int64)
    else:
        r0 = np.empty(lo.shape, np.int64)
        r1 = np.empty(lo.shape, np.int64)
        for i in range(lo.shape[0]):
            for a in range(3):
                r0[i, a], r1[i, a] = _index_range_clipped(coords, lo[i, a], hi[i, a], G)
    rodf = np.zeros((len(Lu), 8), np.float64)
    rodf[:, :3] = p1u
    rodf[:, 3:6] = nu
    rodf[:, 6] = Lu
    rodi = np.stack([r0[:, 0], r1[:, 0], r0[:, 1], r1[:, 1], r0[:, 2], r1[:, 2]], axis=1)
    if np.abs(rodi).max(initial=0) >= 2 ** 31 - 1:
        raise ValueError("rod index box exceeds the int32 range")
    return rodf, rodi.astype(np.int32)


def rasterize_penlike_device(rods, grid_size: int, box_size: float, minor_radius: float = 0.2252,
                             aspect_ratio: float = 2.5, eps_rod: float = 2.9275 ** 2,
                             eps_bg: float = 1.0, periodic: bool = False):
    """`rasterize_penlike` with the voxel test on the GPU; returns the ε grid
    as a (G, G, G) float32 device tensor, bit-identical to the host result."""
    import torch
    from . import _lib
    from ._device import device, stream_ptr
    G = int(grid_size)
    rodf, rodi = rod_parameters(rods, G, box_size, minor_radius, aspect_ratio, periodic)
    dev = device()
    eps = torch.empty((G, G, G), dtype=torch.float32, device=dev)
    f = torch.from_numpy(rodf).to(dev)
    i = torch.from_numpy(np.ascontiguousarray(rodi)).to(dev)
    b = float(minor_radius)
    _lib.check(_lib.load().efn_rasterize(G, float(box_size), float(aspect_ratio), b * b,
                                          f.data_ptr(), i.data_ptr(), int(rodf.shape[0]),
                                          1 if periodic else 0, float(np.float32(eps_rod)),
                                          float(np.float32(eps_bg)), eps.data_ptr(), stream_ptr()),
               "rasterize")
    return eps


def filling_fraction(eps: np.ndarray, eps_bg: float = 1.0) -> float:
    return float((np.asarray(eps) != eps_bg).mean())


def synthetic_network(n_points: int, seed: int, k: int = 4, radius_factor: float = 0.33):
    """Random periodic kNN rod network.  Returns (rods (R,6) µm, box_size, radius)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    pts = rng.random((n_points, 3))
    tree = cKDTree(pts, boxsize=1.0)
    _, idx = tree.query(pts, k=k + 1)
    pairs = set()
    for i in range(n_points):
        for j in idx[i, 1:]:
            j = int(j)
            if j != i:
                pairs.add((min(i, j), max(i, j)))
    pairs = np.array(sorted(pairs), dtype=np.int64)
    d = pts[pairs[:, 1]] - pts[pairs[:, 0]]
    d -= np.round(d)                                  # minimum image
    mean_len = float(np.linalg.norm(d, axis=1).mean())
    L = box_size_for_n(n_points)
    radius = radius_factor * mean_len * L
    p1 = (pts[pairs[:, 0]] - 0.5) * L
    p2 = p1 + d * L
    rods = [np.concatenate([p1, p2], axis=1)]
    shift = -L * np.floor((p2 + L / 2) / L)           # image of rods whose far end left the box
    cross = np.any(shift != 0, axis=1)
    if cross.any():
        rods.append(np.concatenate([p1[cross] + shift[cross], p2[cross] + shift[cross]], axis=1))
    return np.concatenate(rods, axis=0), L, radius


def network_eps(n_points: int, seed: int, grid_size: int, radius_factor: float = 0.33,
                eps_rod: float = 8.41, on_device: bool = False):
    """ε grid of a synthetic network (binary mask, scalar ε per voxel).  Returns
    (eps, L): a NumPy array, or with on_device=True a float32 device tensor
    rasterised on the GPU (bit-identical)."""
    rods, L, radius = synthetic_network(n_points, seed, radius_factor=radius_factor)
    raster = rasterize_penlike_device if on_device else rasterize_penlike
    eps = raster(rods, grid_size, L, minor_radius=radius, aspect_ratio=1.0, eps_rod=eps_rod,
                 eps_bg=1.0, periodic=True)
    return eps, L
