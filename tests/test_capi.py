"""The C-ABI library loads, exports every symbol include/eigenfns_hip.h declares, and
rejects bad arguments before touching the GPU (no compute calls here)."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "eigenfns_hip.h"


def _declared():
    text = HEADER.read_text()
    body = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(efn_[a-z0-9_]+)\s*\(", body)))


@pytest.fixture(scope="module")
def lib():
    from eigenfns_amd import _lib
    return _lib.load()


def test_every_declared_symbol_is_exported_and_bound(lib):
    from eigenfns_amd import _lib
    declared = _declared()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_lib.EXPORTED) == declared


def test_pure_host_queries(lib):
    assert lib.efn_version() >= 100
    for G in (6, 8, 16, 32, 48, 64, 96, 128, 192, 256):
        assert lib.efn_grid_supported(G)
    assert not lib.efn_grid_supported(7)
    assert lib.efn_gram_work_bytes(32, 32, 1 << 20) > 0
    assert lib.efn_gram_work_bytes(0, 32, 100) == 0


def test_argument_errors_map_to_value_error(lib):
    from eigenfns_amd import _lib
    out = C.c_void_p()
    rc = lib.efn_ctx_create(8, 1.0, 0.0, 0.0, 0.0, None, 0, C.byref(out))
    assert rc == _lib.EFN_EINVAL
    with pytest.raises(ValueError, match="eps is NULL"):
        _lib.check(rc)
    import numpy as np
    eps = np.ones(7 ** 3, np.float32)
    rc = lib.efn_ctx_create(7, 1.0, 0.0, 0.0, 0.0, eps.ctypes.data, 0, C.byref(out))
    assert rc == _lib.EFN_EUNSUPPORTED
    assert lib.efn_gram(None, 2, None, 2, 10, None, None, 0, None) == _lib.EFN_EINVAL
    assert lib.efn_rowreduce(7, None, None, 1, 1, None, None, 0, None) == _lib.EFN_EINVAL


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from eigenfns_amd import _lib
    monkeypatch.setattr(_lib, "_LIB_PATH", tmp_path / "nope.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.LibraryMissing):
        _lib.load()
