"""GPU rasterizer: bit-exact against the reference's own masks and the host path.

Golden masks: tests/golden/raster_golden.npz, written by
tests/golden/make_raster_golden.py with the reference's eigenfns/structure.py
(every network case: periodic and montage, aspect 1 and 2.5, up to the
10,000-point network at 192³), plus the face-straddling rod of the
reference's tests/test_structure.py:61-78.  Bar: identical masks.
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from eigenfns_amd import structure as product

pytestmark = pytest.mark.gpu

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "raster_golden.npz")
INDEX = json.loads(bytes(GOLD["index_json"]).decode())
NETWORK_CASES = [c for c in INDEX if "n_points" in c]


def _mask(name, G):
    return np.unpackbits(GOLD[f"{name}/mask"])[:G ** 3].reshape(G, G, G).astype(bool)


@pytest.mark.parametrize("case", NETWORK_CASES, ids=lambda c: c["name"])
def test_device_rasterizer_matches_reference_masks(case):
    rods, L, radius = product.synthetic_network(case["n_points"], case["seed"],
                                                radius_factor=case["radius_factor"])
    digest = hashlib.sha256(np.ascontiguousarray(rods).tobytes()).digest()
    assert bytes(GOLD[f"{case['name']}/rods_sha256"]) == digest, "generator drifted"
    eps = product.rasterize_penlike_device(rods, case["grid"], L, minor_radius=radius,
                                           aspect_ratio=case["aspect_ratio"],
                                           eps_rod=case["eps_rod"],
                                           periodic=case["periodic"]).cpu().numpy()
    assert np.array_equal(eps != 1.0, _mask(case["name"], case["grid"]))
    assert sorted(float(v) for v in np.unique(eps)) == case["eps_values"]


@pytest.mark.parametrize("periodic", [False, True])
def test_device_face_straddling_rod(periodic):
    name = f"face_rod_G40_{'periodic' if periodic else 'montage'}"
    eps = product.rasterize_penlike_device(GOLD[f"{name}/rods"], 40, 10.0, minor_radius=0.5,
                                           aspect_ratio=1.0, periodic=periodic).cpu().numpy()
    assert np.array_equal(eps != 1.0, _mask(name, 40))


@pytest.mark.parametrize("n,seed,G,periodic,aspect", [(64, 7, 36, True, 1.0),
                                                      (216, 4, 50, False, 2.5),
                                                      (1000, 1, 96, True, 1.7)])
def test_device_rasterizer_equals_host(n, seed, G, periodic, aspect):
    rods, L, radius = product.synthetic_network(n, seed)
    host = product.rasterize_penlike(rods, G, L, minor_radius=radius, aspect_ratio=aspect,
                                     eps_rod=8.41, eps_bg=1.5, periodic=periodic)
    dev = product.rasterize_penlike_device(rods, G, L, minor_radius=radius, aspect_ratio=aspect,
                                           eps_rod=8.41, eps_bg=1.5, periodic=periodic)
    assert np.array_equal(host, dev.cpu().numpy())


def test_device_network_eps_feeds_operator():
    import torch
    from eigenfns_amd.operator import MaxwellOperator
    eps_h, L = product.network_eps(216, 2, 32)
    eps_d, L2 = product.network_eps(216, 2, 32, on_device=True)
    assert L == L2 and isinstance(eps_d, torch.Tensor) and eps_d.is_cuda
    assert np.array_equal(eps_h, eps_d.cpu().numpy())
    X = MaxwellOperator(eps_h, L).random_block(2, seed=1)
    assert torch.equal(MaxwellOperator(eps_h, L).theta(X), MaxwellOperator(eps_d, L).theta(X))
