"""GPU tests of the batched Hermitian eigensolver (efn_heev_batched, FP64
Jacobi) against LAPACK (numpy.linalg.eigh) on the same FP64 matrices.

Bars: eigenvalues within 1e-13·max|λ| of LAPACK's; residual ‖AV − VΛ‖ and
‖VᴴV − I‖ below 1e-12 (relative to max|λ|); the smallest eigenvalues of a
matrix with eigenvalues from 3e-2 to 1e4 (a Rayleigh–Ritz matrix with
dead-row penalties) within 1e-10 relative; every problem bitwise the same
alone and in a batch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _herm(rng, n):
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return X + X.conj().T


def _rr_like(rng, n, dead=()):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    k = n // 3
    ev = np.concatenate([np.linspace(0.03, 1.0, k), np.logspace(0, 3, n - k)])
    A = (Q * ev) @ Q.conj().T
    for d in dead:
        A[d, :] = 0
        A[:, d] = 0
        A[d, d] = 1e4
    return A


def _check(A, w, V):
    wr = np.linalg.eigvalsh(A)
    scale = np.abs(wr).max()
    assert np.abs(w - wr).max() < 1e-13 * scale
    assert np.abs(A @ V - V * w).max() < 1e-12 * scale
    assert np.abs(V.conj().T @ V - np.eye(len(w))).max() < 1e-12
    return wr


@pytest.mark.parametrize("n", [1, 2, 7, 24, 48, 72, 96])
def test_heev_matches_lapack(n):
    import torch
    from eigenfns_amd.linalg import eigh_batched_device
    rng = np.random.default_rng(n)
    mats = np.stack([_herm(rng, n) for _ in range(5)])
    w, V = eigh_batched_device(torch.from_numpy(mats).cuda())
    w, V = w.cpu().numpy(), V.cpu().numpy()
    for j in range(len(mats)):
        _check(mats[j], w[j], V[j])


def test_heev_gram_and_rayleigh_ritz_shapes():
    """Near-dependent Gram matrix with a zero row (SVQB) and a 72×72 PSD matrix
    with eigenvalues 3e-2..1e3 plus three dead rows at 1e4 (Rayleigh–Ritz)."""
    import torch
    from eigenfns_amd.linalg import eigh_batched_device
    rng = np.random.default_rng(5)
    Y = rng.standard_normal((24, 300)) + 1j * rng.standard_normal((24, 300))
    Y[5] = Y[3] + 1e-6 * Y[7]
    Y[9] = 0
    G = Y.conj() @ Y.T
    w, V = eigh_batched_device(torch.from_numpy(G[None]).cuda())
    _check(G, w.cpu().numpy()[0], V.cpu().numpy()[0])
    A = _rr_like(rng, 72, dead=(50, 60, 70))
    w, V = eigh_batched_device(torch.from_numpy(A[None]).cuda())
    w0 = w.cpu().numpy()[0]
    wr = _check(A, w0, V.cpu().numpy()[0])
    assert np.abs(w0[:24] - wr[:24]).max() < 1e-10 * np.abs(wr[:24]).min()
    # lowest subset
    wl, Vl = eigh_batched_device(torch.from_numpy(A[None]).cuda(), lowest=24)
    assert wl.shape == (1, 24) and Vl.shape == (1, 72, 24)
    assert torch.equal(wl, w[:, :24]) and torch.equal(Vl, V[:, :, :24])


def test_heev_batch_invariant_and_limits():
    import torch
    from eigenfns_amd.linalg import eigh_batched_device
    rng = np.random.default_rng(9)
    mats = torch.from_numpy(np.stack([_rr_like(rng, 48) for _ in range(6)])).cuda()
    wb, Vb = eigh_batched_device(mats)
    for j in range(6):
        w1, V1 = eigh_batched_device(mats[j:j + 1])
        assert torch.equal(w1[0], wb[j]) and torch.equal(V1[0], Vb[j])
    with pytest.raises(ValueError):
        eigh_batched_device(torch.zeros((1, 97, 97), dtype=torch.complex128, device="cuda"))
