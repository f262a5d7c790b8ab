#!/usr/bin/env python
"""Write golden voxel masks with the reference's own rasterizer.

Runs in the build container only (needs the reference checkout): loads
eigenfns/structure.py from the reference by file path (it imports only re,
pathlib and numpy) and rasterises rod sets produced by the build's synthetic
network generator, plus the face-straddling rod of the reference's
tests/test_structure.py:61-78.  Output: tests/golden/raster_golden.npz with,
per case, the packed voxel mask (np.packbits, C order), the parameters, the
rods (small cases) and a SHA-256 of the rods (all cases).

    python tests/golden/make_raster_golden.py [/path/to/reference]
"""
from __future__ import annotations

import hashlib
import importlib.util
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from eigenfns_amd.structure import synthetic_network  # noqa: E402


def load_reference_structure(ref_root: Path):
    path = ref_root / "eigenfns" / "structure.py"
    spec = importlib.util.spec_from_file_location("reference_structure", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


CASES = [
    # name, n_points, seed, radius_factor, grid, rasterize kwargs (besides radius)
    ("n64_s0_G32_periodic", 64, 0, 0.33, 32, dict(aspect_ratio=1.0, eps_rod=8.41, periodic=True)),
    ("n64_s0_G32_montage_s2.5", 64, 0, 0.33, 32, dict(aspect_ratio=2.5, eps_rod=8.41, periodic=False)),
    ("n216_s2_G64_periodic", 216, 2, 0.33, 64, dict(aspect_ratio=1.0, eps_rod=8.41, periodic=True)),
    ("n1000_s1_G16_periodic", 1000, 1, 0.33, 16, dict(aspect_ratio=1.0, eps_rod=8.41, periodic=True)),
    ("n1000_s1_G128_periodic", 1000, 1, 0.33, 128, dict(aspect_ratio=1.0, eps_rod=8.41, periodic=True)),
    ("n1000_s1_G48_montage_s2.5", 1000, 1, 0.33, 48, dict(aspect_ratio=2.5, eps_rod=8.41, periodic=False)),
    ("n10000_s3_G192_periodic", 10000, 3, 0.331836, 192, dict(aspect_ratio=1.0, eps_rod=8.41, periodic=True)),
]


def main(ref_root: str = "/root/reference") -> None:
    ref = load_reference_structure(Path(ref_root))
    out = {}
    index = []
    for name, n, seed, rf, G, kw in CASES:
        rods, L, radius = synthetic_network(n, seed, radius_factor=rf)
        eps = ref.rasterize_penlike(rods, G, L, minor_radius=radius, **kw)
        mask = eps != 1.0
        out[f"{name}/mask"] = np.packbits(mask.ravel())
        out[f"{name}/rods_sha256"] = np.frombuffer(
            hashlib.sha256(np.ascontiguousarray(rods).tobytes()).digest(), np.uint8)
        if rods.shape[0] <= 5000:
            out[f"{name}/rods"] = rods
        meta = dict(name=name, n_points=n, seed=seed, radius_factor=rf, grid=G, box=L,
                    radius=radius, eps_values=sorted(float(v) for v in np.unique(eps)), **kw)
        index.append(meta)
        print(f"{name}: ff={mask.mean():.6f} rods={rods.shape[0]}", flush=True)
    # tests/test_structure.py:61-78 of the reference: one rod straddling the +x face
    rod = np.array([[4.0, 0.0, 0.0, 6.0, 0.0, 0.0]])
    for periodic in (False, True):
        eps = ref.rasterize_penlike(rod, 40, 10.0, minor_radius=0.5, aspect_ratio=1.0,
                                    periodic=periodic)
        name = f"face_rod_G40_{'periodic' if periodic else 'montage'}"
        out[f"{name}/mask"] = np.packbits((eps != 1.0).ravel())
        out[f"{name}/rods"] = rod
        index.append(dict(name=name, grid=40, box=10.0, radius=0.5, aspect_ratio=1.0,
                          eps_rod=2.9275 ** 2, periodic=periodic))
    out["index_json"] = np.frombuffer(json.dumps(index).encode(), np.uint8)
    np.savez_compressed(HERE / "raster_golden.npz", **out)
    print(f"wrote {HERE / 'raster_golden.npz'}")


if __name__ == "__main__":
    main(*sys.argv[1:])
