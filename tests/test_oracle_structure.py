"""Rasterizer parity: oracle and product masks against the reference's own output.

Golden masks: tests/golden/raster_golden.npz, written by
tests/golden/make_raster_golden.py with the reference's eigenfns/structure.py.
Bit-exact equality is required (the voxel masks are integer data).
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from eigenfns_amd import structure as product
from oracle import structure_np as oracle

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "raster_golden.npz")
INDEX = json.loads(bytes(GOLD["index_json"]).decode())
NETWORK_CASES = [c for c in INDEX if "n_points" in c]
SMALL = [c for c in NETWORK_CASES if c["grid"] <= 64 or c["n_points"] <= 1000]


def _mask(name, G):
    return np.unpackbits(GOLD[f"{name}/mask"])[:G ** 3].reshape(G, G, G).astype(bool)


def _rods(case):
    rods, L, radius = product.synthetic_network(case["n_points"], case["seed"],
                                                radius_factor=case["radius_factor"])
    digest = hashlib.sha256(np.ascontiguousarray(rods).tobytes()).digest()
    assert bytes(GOLD[f"{case['name']}/rods_sha256"]) == digest, "generator drifted"
    assert L == case["box"] and radius == case["radius"]
    return rods, L, radius


@pytest.mark.parametrize("case", NETWORK_CASES, ids=lambda c: c["name"])
def test_product_rasterizer_bit_exact(case):
    rods, L, radius = _rods(case)
    eps = product.rasterize_penlike(rods, case["grid"], L, minor_radius=radius,
                                    aspect_ratio=case["aspect_ratio"], eps_rod=case["eps_rod"],
                                    periodic=case["periodic"])
    assert np.array_equal(eps != 1.0, _mask(case["name"], case["grid"]))
    assert sorted(float(v) for v in np.unique(eps)) == case["eps_values"]


@pytest.mark.parametrize("case", SMALL, ids=lambda c: c["name"])
def test_oracle_rasterizer_bit_exact(case):
    rods, L, radius = _rods(case)
    eps = oracle.rasterize_penlike(rods, case["grid"], L, minor_radius=radius,
                                   aspect_ratio=case["aspect_ratio"], eps_rod=case["eps_rod"],
                                   periodic=case["periodic"])
    assert np.array_equal(eps != 1.0, _mask(case["name"], case["grid"]))


@pytest.mark.parametrize("periodic", [False, True])
def test_face_straddling_rod(periodic):
    # reference tests/test_structure.py:61-78
    name = f"face_rod_G40_{'periodic' if periodic else 'montage'}"
    rod = GOLD[f"{name}/rods"]
    for mod in (product, oracle):
        eps = mod.rasterize_penlike(rod, 40, 10.0, minor_radius=0.5, aspect_ratio=1.0,
                                    periodic=periodic)
        assert np.array_equal(eps != 1.0, _mask(name, 40))
    plain, wrapped = _mask("face_rod_G40_montage", 40), _mask("face_rod_G40_periodic", 40)
    assert wrapped.sum() > plain.sum() and wrapped[0].sum() > 0 and plain[0].sum() == 0
    assert np.all(wrapped[plain])


def test_box_size_and_filling_fraction():
    assert abs(product.box_size_for_n(8000) - 2 * 11.44) < 1e-9
    assert abs(product.box_size_for_n(1000) - 11.44) < 1e-12
    eps = np.ones((4, 4, 4), np.float32)
    eps[0, 0, :2] = 8.41
    assert product.filling_fraction(eps) == 2 / 64


def test_synthetic_network_geometry():
    rods, L, radius = product.synthetic_network(64, 0)
    assert abs(L - product.box_size_for_n(64)) < 1e-12
    seg = rods[:, 3:] - rods[:, :3]
    # every rod is a minimum-image bond: shorter than half the box
    assert np.all(np.linalg.norm(seg, axis=1) < L / 2)
    # every rod has at least one endpoint inside the centred box
    inside = np.all(np.abs(rods[:, :3]) <= L / 2, axis=1) | np.all(np.abs(rods[:, 3:]) <= L / 2, axis=1)
    assert inside.all()
    assert 0 < radius < L / 10


def test_workload_definitions():
    """k-path corners and spacing, C3 widths, and the fixture metadata agree
    with the workload definitions they were generated from."""
    import json
    import os
    from eigenfns_amd import workloads as wl
    ks = wl.kpath(64)
    assert np.allclose(ks[0], 0) and np.allclose(ks[-1], 0.5)
    steps = np.linalg.norm(np.diff(ks, axis=0), axis=1)
    # uniform arc length except where a step turns a corner (shorter chord)
    assert np.all(steps <= steps.max() + 1e-12) and np.median(steps) == pytest.approx(steps.max())
    probs = wl.c3_problems()
    assert len(probs) == 64
    assert all(8 <= w <= 48 and nev == w - 4 for _, w, nev in probs)
    path = os.path.join(os.path.dirname(__file__), "golden", "solve_golden.npz")
    if not os.path.exists(path):
        pytest.skip("solve_golden.npz not generated")
    z = np.load(path, allow_pickle=False)
    meta = {k[:-6]: json.loads(str(z[k])) for k in z.files if k.endswith("__meta")}
    if "c1" in meta:
        assert meta["c1"]["nev"] == wl.C1["nev"] == len(z["c1"])
    for key, md in meta.items():
        if key.startswith("srv_"):
            sp = wl.serving_sample()[md["rid"]]
            assert (sp.grid, sp.nev, sp.n_points) == (md["grid"], md["nev"], md["n_points"])
            assert np.allclose(sp.k_frac, md["k"])
        assert np.all(np.diff(z[key]) >= -1e-12) and np.all(z[key] > 0)
