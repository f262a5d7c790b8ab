"""GPU tests of the batched Hermitian eigensolvers — efn_heev_batched (the
library's FP64 Jacobi kernel), efn_heev_ql_batched (its tridiagonal QL
kernel), efn_heevx_batched (its bisection + inverse-iteration kernel) and
efn_heevd_batched (rocSOLVER's batched divide and conquer) — against LAPACK
(numpy.linalg.eigh) on the same FP64 matrices.

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


METHODS = ["dc", "jacobi", "ql", "heevx"]


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("n", [1, 2, 7, 24, 48, 72, 96])
def test_heev_matches_lapack(n, method):
    import torch
    from eigenfns_amd.linalg import eigh_batched_device
    rng = np.random.default_rng(n)
    mats = np.stack([_herm(rng, n) for _ in range(5)])
    w, V = eigh_batched_device(torch.from_numpy(mats).cuda(), method=method)
    w, V = w.cpu().numpy(), V.cpu().numpy()
    for j in range(len(mats)):
        _check(mats[j], w[j], V[j])


@pytest.mark.parametrize("method", METHODS)
def test_heev_gram_and_rayleigh_ritz_shapes(method):
    """Near-dependent Gram matrix with a zero row (SVQB) and a 72×72 PSD matrix
    with eigenvalues 3e-2..1e3 plus three dead rows at 1e4 (Rayleigh–Ritz)."""
    import torch
    from eigenfns_amd.linalg import eigh_batched_device
    rng = np.random.default_rng(5)
    Y = rng.standard_normal((24, 300)) + 1j * rng.standard_normal((24, 300))
    Y[5] = Y[3] + 1e-6 * Y[7]
    Y[9] = 0
    G = Y.conj() @ Y.T
    w, V = eigh_batched_device(torch.from_numpy(G[None]).cuda(), method=method)
    _check(G, w.cpu().numpy()[0], V.cpu().numpy()[0])
    A = _rr_like(rng, 72, dead=(50, 60, 70))
    w, V = eigh_batched_device(torch.from_numpy(A[None]).cuda(), method=method)
    w0 = w.cpu().numpy()[0]
    wr = _check(A, w0, V.cpu().numpy()[0])
    assert np.abs(w0[:24] - wr[:24]).max() < 1e-10 * np.abs(wr[:24]).min()
    # lowest subset
    wl, Vl = eigh_batched_device(torch.from_numpy(A[None]).cuda(), lowest=24, method=method)
    assert wl.shape == (1, 24) and Vl.shape == (1, 72, 24)
    assert torch.equal(wl, w[:, :24]) and torch.equal(Vl, V[:, :, :24])


@pytest.mark.parametrize("method", METHODS)
def test_heev_batch_invariant_and_limits(method):
    import torch
    from eigenfns_amd.linalg import eigh_batched_device
    rng = np.random.default_rng(9)
    mats = torch.from_numpy(np.stack([_rr_like(rng, 48) for _ in range(6)])).cuda()
    wb, Vb = eigh_batched_device(mats, method=method)
    for j in range(6):
        w1, V1 = eigh_batched_device(mats[j:j + 1], method=method)
        assert torch.equal(w1[0], wb[j]) and torch.equal(V1[0], Vb[j])
    if method in ("jacobi", "ql", "heevx"):
        with pytest.raises(ValueError):
            eigh_batched_device(torch.zeros((1, 97, 97), dtype=torch.complex128, device="cuda"),
                                method=method)


@pytest.mark.parametrize("method", METHODS)
def test_heev_zero_and_diagonal_matrices(method):
    """An all-zero matrix (the Gram matrix of a block without P rows) gives
    w = 0 and V = I; a diagonal one its sorted diagonal; both finite."""
    import torch
    from eigenfns_amd.linalg import eigh_batched_device
    mats = np.zeros((3, 24, 24), np.complex128)
    mats[1] = np.diag(np.arange(24, 0, -1.0))
    mats[2] = _herm(np.random.default_rng(2), 24)
    w, V = eigh_batched_device(torch.from_numpy(mats).cuda(), method=method)
    w, V = w.cpu().numpy(), V.cpu().numpy()
    assert np.isfinite(w).all() and np.isfinite(V).all()
    assert np.array_equal(w[0], np.zeros(24))
    assert np.abs(np.abs(V[0]) - np.eye(24)).max() < 1e-14 or np.abs(V[0].conj().T @ V[0] - np.eye(24)).max() < 1e-12
    assert np.allclose(w[1], np.arange(1, 25.0), rtol=0, atol=1e-12)
    _check(mats[2], w[2], V[2])


@pytest.mark.parametrize("method", METHODS)
def test_heev_decoupled_dead_block(method):
    """A Rayleigh–Ritz matrix whose whole P block (rows 48-71) is dead: exactly
    zero coupling to the rest and a 24-fold penalty eigenvalue.  Eigenpairs
    finite and to the usual bars (the divide and conquer alone returned
    non-finite vectors for this structure)."""
    import torch
    from eigenfns_amd.linalg import eigh_batched_device
    rng = np.random.default_rng(11)
    mats = np.stack([_rr_like(rng, 72, dead=range(48, 72)), _rr_like(rng, 72, dead=range(60, 72)),
                     _rr_like(rng, 72)])
    w, V = eigh_batched_device(torch.from_numpy(mats).cuda(), method=method)
    w, V = w.cpu().numpy(), V.cpu().numpy()
    assert np.isfinite(w).all() and np.isfinite(V).all()
    for j in range(3):
        wr = _check(mats[j], w[j], V[j])
        assert np.abs(w[j][:24] - wr[:24]).max() < 1e-10 * np.abs(wr[:24]).min()


@pytest.mark.parametrize("method", ["ql", "heevx"])
def test_heev_small_ragged_sizes_in_one_launch(method):
    """Problems of sizes 1-96 padded into one (nprob, 96, 96) stack, the lowest
    24 of each in one launch: each problem to the LAPACK bars, +inf / zero
    padding past its own size, and bitwise the same as the problem alone."""
    import torch
    from eigenfns_amd.linalg import eigh_ragged_device
    rng = np.random.default_rng(21)
    sizes = [1, 5, 24, 30, 48, 72, 95, 96, 0]
    K, low = 96, 24
    A = np.zeros((len(sizes), K, K), np.complex128)
    for j, n in enumerate(sizes):
        if n:
            A[j, :n, :n] = _rr_like(rng, n, dead=(n - 1,) if n > 3 else ())
    At = torch.from_numpy(A).cuda()
    w, V = eigh_ragged_device(At, sizes, low, method)
    wn, Vn = w.cpu().numpy(), V.cpu().numpy()
    for j, n in enumerate(sizes):
        k = min(low, n)
        if n:
            wr = np.linalg.eigvalsh(A[j, :n, :n])
            scale = np.abs(wr).max()
            assert np.abs(wn[j, :k] - wr[:k]).max() < 1e-13 * scale
            Vj = Vn[j, :n, :k]
            assert np.abs(A[j, :n, :n] @ Vj - Vj * wn[j, :k]).max() < 1e-12 * scale
            assert np.abs(Vj.conj().T @ Vj - np.eye(k)).max() < 1e-12
        assert np.isinf(wn[j, k:]).all() and not Vn[j, n:].any() and not Vn[j, :, k:].any()
        w1, V1 = eigh_ragged_device(At[j:j + 1], [n], low, method)
        assert torch.equal(w1[0], w[j]) and torch.equal(V1[0], V[j])


@pytest.mark.parametrize("method", ["ql", "heevx"])
def test_heev_small_more_columns_than_the_stack(method):
    """Asking for more eigenpairs than the stack's size (an engine whose
    requests are all narrower than a third of its width) pads w with +inf and
    V with zero columns; the leading columns are the full solve's."""
    import torch
    from eigenfns_amd.linalg import eigh_ragged_device
    rng = np.random.default_rng(31)
    A = torch.from_numpy(np.stack([_herm(rng, 24), _herm(rng, 24)])).cuda()
    w, V = eigh_ragged_device(A, [24, 10], 32, method)
    wf, Vf = eigh_ragged_device(A, [24, 10], 24, method)
    assert w.shape == (2, 32) and V.shape == (2, 24, 32)
    assert torch.equal(w[:, :24], wf) and torch.equal(V[:, :, :24], Vf)
    assert torch.isinf(w[:, 24:]).all() and not V[:, :, 24:].any()
    assert torch.isinf(w[1, 10:]).all()
