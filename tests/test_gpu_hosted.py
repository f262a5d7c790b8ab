"""Host-resident basis helpers (eigenfns/interior.py:93-138): `_stream_rotate`,
`gram_hosted`, `svqb_hosted` move host row chunks through the device Gram and
combine kernels.  Bars: Gram to 1e-12 of complex128 (FP64 sums of exact
products), rotations to 2e-5 relative (FP32 combine), SVQB output orthonormal
to 1e-5 and spanning the input."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _host_block(m, n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((m, 2, n)) + 1j * rng.standard_normal((m, 2, n))).astype(np.complex64)


def test_stream_rotate_and_gram_hosted_match_complex128():
    from eigenfns_amd.interior import _stream_rotate, gram_hosted
    Xh = _host_block(19, 1000, 1)
    Yh = _host_block(11, 1000, 2)
    C = (np.random.default_rng(3).standard_normal((19, 13))
         + 1j * np.random.default_rng(4).standard_normal((19, 13)))
    out = _stream_rotate(C, Xh, rows=8)
    want = np.tensordot(C.T, Xh.astype(np.complex128), axes=(1, 0))
    assert out.shape == want.shape and out.dtype == np.complex64
    assert np.abs(out - want).max() < 2e-5 * np.abs(want).max()
    Gm = gram_hosted(Xh, Yh, rows=8)
    Gw = Xh.reshape(19, -1).astype(np.complex128).conj() @ Yh.reshape(11, -1).astype(np.complex128).T
    assert np.abs(Gm - Gw).max() < 1e-12 * np.abs(Gw).max()
    assert np.abs(gram_hosted(Xh, rows=5) - gram_hosted(Xh, Xh, rows=8)).max() < 1e-12 * np.abs(Gw).max()


def test_svqb_hosted_orthonormalises_and_drops_dependent_rows():
    from eigenfns_amd.interior import svqb_hosted
    Xh = _host_block(10, 4000, 5)
    Xh[7] = Xh[2] + 1e-9 * Xh[3]
    Q = svqb_hosted(Xh.copy(), rows=4)
    Qf = Q.reshape(10, -1).astype(np.complex128)
    nrm = np.linalg.norm(Qf, axis=1)
    assert (nrm < 1e-3).sum() == 1
    live = Qf[nrm > 0.5]
    assert np.abs(live.conj() @ live.T - np.eye(live.shape[0])).max() < 1e-5
    # the live rows span the input's range
    Xf = Xh.reshape(10, -1).astype(np.complex128)
    resid = Xf - (Xf @ live.conj().T) @ live
    assert np.linalg.norm(resid) < 1e-4 * np.linalg.norm(Xf)
