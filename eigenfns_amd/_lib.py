"""ctypes binding of libeigenfns_hip.so (the C ABI in include/eigenfns_hip.h).

The library is the only compute path: if it is missing or cannot be loaded,
every entry point raises — there is no CPU or framework fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

# EIGENFNS_HIP_LIB names another build of the same library (same C ABI)
_LIB_PATH = Path(os.environ.get("EIGENFNS_HIP_LIB") or
                 Path(__file__).resolve().parent / "libeigenfns_hip.so")
_lock = threading.Lock()
_lib = None

EFN_OK = 0
EFN_EINVAL = -1
EFN_EUNSUPPORTED = -2
EFN_EHIP = -3
EFN_ENOMEM = -4

_vp = C.c_void_p
_i = C.c_int
_i64 = C.c_int64
_f = C.c_float
_d = C.c_double
_sz = C.c_size_t

_SIGNATURES = {
    "efn_version": (_i, []),
    "efn_last_error": (C.c_char_p, []),
    "efn_grid_supported": (_i, [_i]),
    "efn_ctx_create": (_i, [_i, _d, _d, _d, _d, _vp, _i, C.POINTER(_vp)]),
    "efn_ctx_destroy": (_i, [_vp]),
    "efn_ctx_query": (_i, [_vp, C.POINTER(_d), C.POINTER(_d), C.POINTER(_i), C.POINTER(_sz)]),
    "efn_ctx_basis": (_i, [_vp, _vp, _vp, _vp]),
    "efn_theta": (_i, [_vp, _vp, _vp, _i, _f, _f, _vp]),
    "efn_precondition": (_i, [_vp, _vp, _vp, _i, _vp]),
    "efn_theta_multi": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _i, _f, _f, _vp]),
    "efn_precondition_multi": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _i, _vp]),
    "efn_theta_ragged": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _f, _f, _vp]),
    "efn_precondition_ragged": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "efn_gram_ragged": (_i, [_i, _vp, _vp, _vp, _vp, _i64, _i, _vp, _vp, _sz, _vp]),
    "efn_combine_ragged": (_i, [_i, _vp, _vp, _vp, _vp, _i64, _vp, _f, _vp, _vp, _vp, _vp]),
    "efn_precondition_simple": (_i, [_vp, _vp, _vp, _i, _f, _vp]),
    "efn_cheb_step": (_i, [_vp, _vp, _vp, _vp, _f, _f, _i, _vp]),
    "efn_kpm_step": (_i, [_vp, _vp, _vp, _vp, _i, _i, _f, _i, _vp]),
    "efn_kpm_count_step": (_i, [_vp, _vp, _vp, _vp, _vp, _d, _f, _i, _vp]),
    "efn_gram_work_bytes": (_sz, [_i, _i, _i64]),
    "efn_gram_engine": (_i, [_i]),
    "efn_theta_engine": (_i, [_i]),
    "efn_theta_passes": (_i, [_i]),
    "efn_ctx_set_passes": (_i, [_vp, _i]),
    "efn_theta_three_pass": (_i, [_i]),
    "efn_combine_engine": (_i, [_i]),
    "efn_gram": (_i, [_vp, _i, _vp, _i, _i64, _vp, _vp, _sz, _vp]),
    "efn_gram_noalloc": (_i, [_vp, _i, _vp, _i, _i64, _vp, _vp]),
    "efn_combine": (_i, [_vp, _i, _i, _vp, _i64, _vp, _f, _vp, _vp]),
    "efn_gram_multi_work_bytes": (_sz, [_i, _i, _i, _i64]),
    "efn_gram_multi": (_i, [_i, _vp, _vp, _i, _i, _i64, _vp, _vp, _sz, _vp]),
    "efn_gram_multi_herm": (_i, [_i, _vp, _vp, _i, _i64, _vp, _vp, _sz, _vp]),
    "efn_combine_multi": (_i, [_i, _vp, _vp, _i, _i, _i64, _vp, _f, _vp, _vp]),
    "efn_combine_multi_split": (_i, [_i, _vp, _vp, _i, _i, _i64, _vp, _vp, _i, _vp]),
    "efn_heev_batched": (_i, [_i, _i, _vp, _vp, _vp, _i, _vp]),
    "efn_heevd_work_bytes": (_sz, [_i, _i]),
    "efn_heevd_batched": (_i, [_i, _i, _vp, _vp, _vp, _vp, _sz, _vp]),
    "efn_heev_ql_batched": (_i, [_i, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "efn_heevx_batched": (_i, [_i, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "efn_cholqr_batched": (_i, [_i, _i, _vp, _vp, _d, _vp, _vp]),
    "efn_rowreduce_work_bytes": (_sz, [_i, _i64]),
    "efn_rowreduce": (_i, [_i, _vp, _vp, _i, _i64, _vp, _vp, _sz, _vp]),
    "efn_row_axpby": (_i, [_vp, _vp, _vp, _vp, _i, _i64, _vp, _vp]),
    "efn_rowreduce_strided": (_i, [_i, _vp, _i64, _vp, _i64, _i, _i, _i64, _vp, _vp, _sz, _vp]),
    "efn_rowstats_work_bytes": (_sz, [_i, _i64]),
    "efn_rowstats_strided": (_i, [_vp, _i64, _vp, _i64, _i, _i, _i64, _vp, _vp, _sz, _vp]),
    "efn_row_axpby_norm_strided": (_i, [_vp, _i64, _vp, _vp, _i64, _vp, _i, _i, _i64, _vp, _i64, _vp,
                                        _vp, _sz, _vp]),
    "efn_row_axpby_strided": (_i, [_vp, _i64, _vp, _vp, _i64, _vp, _i, _i, _i64, _vp, _i64, _vp]),
    "efn_rowstats_sel": (_i, [_vp, _i64, _vp, _i64, _vp, _i, _i, _i, _i64, _vp, _vp, _sz, _vp]),
    "efn_row_axpby_norm_sel": (_i, [_vp, _i64, _vp, _vp, _i64, _vp, _vp, _i, _i, _i, _i64, _vp, _i64,
                                    _vp, _vp, _sz, _vp]),
    "efn_axpby": (_i, [_vp, _f, _vp, _f, _i64, _vp, _vp]),
    "efn_copy_multi": (_i, [_i, _vp, _vp, _i64, _vp]),
    "efn_random_block": (_i, [_vp, _i, C.c_uint64, _i64, _i, _vp, _vp]),
    "efn_rasterize": (_i, [_i, _d, _d, _d, _vp, _vp, _i, _i, _f, _f, _vp, _vp]),
    "efn_realspace": (_i, [_vp, _i, _vp, _vp, _vp, _i, _vp]),
}

EXPORTED = tuple(_SIGNATURES)


class LibraryMissing(RuntimeError):
    pass


def lib_path() -> Path:
    return _LIB_PATH


def load():
    """Load (once) and return the ctypes library handle; raise loudly if absent."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not _LIB_PATH.exists():
            raise LibraryMissing(
                f"{_LIB_PATH} is not built; run `python -m eigenfns_amd.build` "
                "(hipcc, gfx950). eigenfns_amd has no fallback compute path.")
        lib = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc == EFN_OK:
        return
    msg = load().efn_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc in (EFN_EINVAL, EFN_EUNSUPPORTED):
        raise ValueError(text)
    if rc == EFN_ENOMEM:
        raise MemoryError(text)
    raise RuntimeError(text)
