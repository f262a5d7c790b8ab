"""efn_gram_noalloc (the Gram entry point that allocates nothing: split
partials in the library's static device scratch) against efn_gram with a
caller-provided work buffer, through the C ABI.  Bar: bitwise equal (same
kernels, same split count, same fixed-order FP64 sum), on two calls in a row,
the second smaller than the first, so stale scratch contents from the first
call cannot leak into the second; sizes beyond the scratch are rejected on the
host before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gram_with_work(lib, A, B):
    import torch
    from eigenfns_amd._device import stream_ptr
    ma, mb, n = A.shape[0], B.shape[0], A.shape[1]
    out = torch.empty((ma, mb), dtype=torch.complex128, device=A.device)
    wb = lib.efn_gram_work_bytes(ma, mb, n)
    work = torch.empty(max(wb, 16), dtype=torch.uint8, device=A.device)
    rc = lib.efn_gram(A.data_ptr(), ma, B.data_ptr(), mb, n, out.data_ptr(), work.data_ptr(), wb,
                      stream_ptr())
    assert rc == 0, lib.efn_last_error()
    return out


def _gram_noalloc(lib, A, B):
    import torch
    from eigenfns_amd._device import stream_ptr
    ma, mb, n = A.shape[0], B.shape[0], A.shape[1]
    out = torch.empty((ma, mb), dtype=torch.complex128, device=A.device)
    rc = lib.efn_gram_noalloc(A.data_ptr(), ma, B.data_ptr(), mb, n, out.data_ptr(), stream_ptr())
    assert rc == 0, lib.efn_last_error()
    return out


def _rand(m, n, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.randn((m, n), dtype=torch.complex64, device="cuda", generator=g)


def test_noalloc_gram_equals_gram_on_consecutive_calls():
    import torch
    from eigenfns_amd import _lib
    lib = _lib.load()
    # first a large call (96 × 96 over 2·64³ entries: 128 splits), then a smaller one
    for ma, mb, n, seed in [(96, 96, 2 * 64 ** 3, 1), (24, 40, 2 * 32 ** 3 + 6, 2), (5, 3, 1000, 3)]:
        A, B = _rand(ma, n, seed), _rand(mb, n, seed + 100)
        want = _gram_with_work(lib, A, B)
        got = _gram_noalloc(lib, A, B)
        torch.cuda.synchronize()
        assert torch.equal(got, want), (ma, mb, n)
        ref = A.cpu().numpy().astype(np.complex128).conj() @ B.cpu().numpy().astype(np.complex128).T
        assert np.abs(got.cpu().numpy() - ref).max() < 1e-12 * np.abs(ref).max()


def test_noalloc_gram_rejects_sizes_beyond_the_scratch():
    from eigenfns_amd import _lib
    lib = _lib.load()
    n = 2 * 192 ** 3
    assert lib.efn_gram_work_bytes(96, 96, n) <= 40 << 20   # the Rayleigh–Ritz size fits
    assert lib.efn_gram_work_bytes(128, 128, n) > 40 << 20
    out = C.c_void_p(16)   # never written: the size check runs before any launch
    rc = lib.efn_gram_noalloc(C.c_void_p(16), 128, C.c_void_p(16), 128, n, out, None)
    assert rc == _lib.EFN_EINVAL
    assert b"static scratch" in lib.efn_last_error()
