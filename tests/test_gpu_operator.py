"""GPU parity of the fused operator kernels against the FP64 oracle.

Tolerance: Θ-apply within 1e-5 relative in L2 per vector; the
same bar for the preconditioner and the fused Chebyshev/KPM steps.
"""
import numpy as np
import pytest

from oracle.operator_np import MaxwellOperatorNP, make_basis
from oracle.rng_np import random_block as np_random_block
from oracle import kpm_np, interior_np

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _eps(G, seed=0, frac=0.3, eps_rod=8.41):
    rng = np.random.default_rng(seed)
    return np.where(rng.random((G, G, G)) < frac, eps_rod, 1.0).astype(np.float32)


def _block(G, m, seed, k_frac=(0, 0, 0)):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((m, 2, G, G, G)) + 1j * rng.standard_normal((m, 2, G, G, G)))
    if np.allclose(k_frac, 0):
        X[:, :, 0, 0, 0] = 0
    return X.astype(np.complex64)


def _rel(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def _ops(G, k_frac=(0, 0, 0), L=None, seed=0):
    from eigenfns_amd.operator import MaxwellOperator
    eps = _eps(G, seed)
    L = L if L is not None else 0.13 * G
    return MaxwellOperator(eps, L, k_frac), MaxwellOperatorNP(eps, L, k_frac)


CASES = [(6, (0.5, 0, 0)), (8, (0, 0, 0)), (12, (0.1, 0.2, 0.3)), (16, (0, 0, 0)),
         (24, (0, 0, 0)), (32, (0.25, 0.25, 0)), (48, (0, 0, 0)), (64, (0, 0, 0)),
         (96, (0.5, 0.5, 0.5)), (128, (0, 0, 0)), (192, (0, 0, 0))]


@pytest.mark.parametrize("G,k", CASES + [(256, (0.1, 0, 0.2))])
def test_theta_matches_fp64_oracle(G, k):
    import torch
    op, ref = _ops(G, k)
    m = 3 if G <= 64 else 1
    X = _block(G, m, seed=G, k_frac=k)
    Y = op.theta(torch.from_numpy(X).cuda()).cpu().numpy()
    Yr = ref.theta(X)
    rel = _rel(Y, Yr)
    print(f"G={G} k={k} theta rel L2 = {rel.max():.2e}")
    assert rel.max() < TOL


@pytest.mark.parametrize("G,k", CASES)
def test_precondition_matches_fp64_oracle(G, k):
    import torch
    op, ref = _ops(G, k)
    m = 2 if G <= 64 else 1
    X = _block(G, m, seed=G + 1, k_frac=k)
    Y = op.precondition(torch.from_numpy(X).cuda()).cpu().numpy()
    Yr = ref.precondition(X)
    rel = _rel(Y, Yr)
    print(f"G={G} k={k} precondition rel L2 = {rel.max():.2e}")
    assert rel.max() < TOL


@pytest.mark.parametrize("G,k", [(8, (0, 0, 0)), (12, (0.1, 0.2, 0.3)), (32, (0.5, 0, 0))])
def test_kernel_frame_matches_reference_frame(G, k):
    op, _ = _ops(G, k)
    kn, t = op.kernel_basis()
    kn_r, t_r, _ = make_basis(G, op.box_size, k)
    assert np.abs(kn.cpu().numpy() - kn_r).max() <= 2e-6 * kn_r.max()
    assert np.abs(t.cpu().numpy() - t_r).max() < 2e-6
    # the basis tensors of the drop-in API are the FP64 frame cast to FP32
    b = op.basis
    assert np.array_equal(b.t.cpu().numpy(), t_r.astype(np.float32))
    assert np.array_equal(b.kn.cpu().numpy(), kn_r.astype(np.float32))


def test_theta_affine_and_chunking():
    import torch
    from eigenfns_amd.operator import MaxwellOperator
    G = 32
    eps = _eps(G, 3)
    op1 = MaxwellOperator(eps, 4.0, max_chunk=1)
    op5 = MaxwellOperator(eps, 4.0, max_chunk=5)
    X = torch.from_numpy(_block(G, 7, 11)).cuda()
    a = op1.theta_affine(X, 0.37, -1.0)
    b = op5.theta_affine(X, 0.37, -1.0)
    # chunking must not change results (batch invariance)
    assert torch.equal(a, b)
    ref = MaxwellOperatorNP(eps, 4.0)
    r = 0.37 * ref.theta(X.cpu().numpy()) - X.cpu().numpy()
    assert _rel(a.cpu().numpy(), r).max() < TOL


def test_random_blocks_match_oracle_stream():
    from eigenfns_amd.operator import MaxwellOperator
    G = 16
    op = MaxwellOperator(_eps(G), 2.0)
    mask = (make_basis(G, 2.0)[0] > 0)
    Zr = op.random_block(5, seed=123, kind="rademacher", row0=7).cpu().numpy()
    assert np.array_equal(Zr, np_random_block("rademacher", 123, 7, 5, G, mask))
    Zn = op.random_block(4, seed=9, kind="normal", row0=3).cpu().numpy()
    ref = np_random_block("normal", 9, 3, 4, G, mask)
    assert np.abs(Zn - ref).max() < 1e-5
    # rows are independent of the chunking of the draw
    Za = op.random_block(6, seed=9, kind="normal", row0=0).cpu().numpy()
    assert np.array_equal(Za[3:], Zn[:3])


def test_cheb_step_matches_oracle():
    import torch
    op, ref = _ops(24)
    alpha, ck = 2.0 / 900.0, 0.173
    T0 = _block(24, 3, 1)
    T1 = _block(24, 3, 2)
    Y = _block(24, 3, 3)
    t0, t1, y = (torch.from_numpy(a).cuda() for a in (T0, T1, Y))
    op.cheb_step(t1, t0, y, alpha, ck)
    T2 = 2 * (alpha * ref.theta(T1) - T1) - T0
    assert _rel(t0.cpu().numpy(), T2).max() < TOL
    assert _rel(y.cpu().numpy(), Y + ck * T2).max() < TOL


def test_kpm_step_matches_doubling_oracle():
    import torch
    from eigenfns_amd.kpm import kpm_moments
    op, ref = _ops(16, seed=4)
    lam_max = 1.05 * float(np.max(ref.kn) ** 2 / 1.0)
    mom = kpm_moments(op, degree=40, n_probe=3, seed=5, lam_max=lam_max)
    Z = np_random_block("rademacher", 5, 0, 3, 16, ref.mask)
    mref = kpm_np.kpm_moments(ref, 40, 3, 5, lam_max, probes=Z)
    assert mom.shape == mref.shape
    scale = np.abs(mref).max()
    assert np.abs(mom - mref).max() < 1e-5 * scale
    assert np.allclose(mom[:, 0], 2 * 16 ** 3 - 2)


def test_gram_combine_rowreduce_match_numpy():
    import torch
    from eigenfns_amd import linalg
    rng = np.random.default_rng(0)
    for ma, mb, n in [(5, 7, 1000), (32, 32, 70000), (96, 96, 8192), (3, 130, 4097)]:
        A = (rng.standard_normal((ma, n)) + 1j * rng.standard_normal((ma, n))).astype(np.complex64)
        B = (rng.standard_normal((mb, n)) + 1j * rng.standard_normal((mb, n))).astype(np.complex64)
        g = linalg.gram(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()).cpu().numpy()
        gr = A.astype(np.complex128).conj() @ B.astype(np.complex128).T
        assert np.abs(g - gr).max() < 1e-12 * np.abs(gr).max() + 1e-9
        Cm = (rng.standard_normal((ma, mb)) + 1j * rng.standard_normal((ma, mb))).astype(np.complex64)
        out = linalg.combine(torch.from_numpy(Cm).cuda(), torch.from_numpy(A).cuda()).cpu().numpy()
        outr = Cm.astype(np.complex128).T @ A.astype(np.complex128)
        assert np.abs(out - outr).max() < 2e-5 * np.abs(outr).max()
        nr = linalg.rownorm2(torch.from_numpy(A).cuda()).cpu().numpy()
        assert np.allclose(nr, (np.abs(A.astype(np.complex128)) ** 2).sum(1), rtol=1e-12)


def test_ortho_and_deflate():
    import torch
    from eigenfns_amd import linalg
    rng = np.random.default_rng(1)
    n = 2 * 16 ** 3
    X = (rng.standard_normal((12, n)) + 1j * rng.standard_normal((12, n))).astype(np.complex64)
    X[5] = X[3] + 1e-9 * X[4]  # rank-deficient direction → one zero row after SVQB
    Q = linalg.ortho_block(torch.from_numpy(X).cuda().reshape(12, 2, 16, 16, 16))
    Qn = Q.reshape(12, -1).cpu().numpy().astype(np.complex128)
    nrm = np.linalg.norm(Qn, axis=1)
    assert (nrm < 1e-3).sum() == 1
    live = Qn[nrm > 0.5]
    assert np.abs(live.conj() @ live.T - np.eye(live.shape[0])).max() < 1e-5
    L = torch.from_numpy(live[:4].astype(np.complex64)).cuda().reshape(4, 2, 16, 16, 16)
    Y = torch.from_numpy(X[:6]).cuda().reshape(6, 2, 16, 16, 16)
    D = linalg.deflate(Y, L)
    leak = linalg.gram(L, D).abs().max().item()
    assert leak < 1e-4 * float(np.abs(X).max() * np.sqrt(n))


@pytest.mark.parametrize("G,k", [(12, (0.0, 0.0, 0.0)), (16, (0.23, -0.4, 0.1)), (32, (0.0, 0.0, 0.0)),
                                 (48, (0.5, 0.5, 0.0))])
def test_realspace_fields_match_fp64_oracle(G, k):
    """H(r), E(r) and ε|E|² (operator.py:174-184, run_modes.py:116-122) from the
    fused passes against the FP64 oracle: rel L2 < 1e-5."""
    import torch
    from eigenfns_amd.operator import MaxwellOperator
    from oracle.operator_np import MaxwellOperatorNP
    rng = np.random.default_rng(G)
    eps = np.where(rng.random((G, G, G)) < 0.35, 8.41, 1.0).astype(np.float32)
    op = MaxwellOperator(eps, 1.7, k)
    ref = MaxwellOperatorNP(eps, 1.7, k)
    X = op.random_block(3, seed=G)
    Xn = X.cpu().numpy().astype(np.complex128)
    lam = np.array([0.5, 2.0, 7.5])
    for got, want in ((op.h_realspace(X), ref.h_realspace(Xn)),
                      (op.e_realspace(X, lam), ref.e_realspace(Xn, lam)),
                      (op.energy_density(X, lam), ref.energy_density(Xn, lam))):
        g = got.cpu().numpy()
        assert g.shape == want.shape
        rel = np.linalg.norm((g - want).ravel()) / np.linalg.norm(want.ravel())
        assert rel < 1e-5, rel


def test_energy_density_parseval_at_scale():
    """At 96³ on a rod network: Σ_r ε|E(r)|² = Re⟨H, ΘH⟩/(G³λ) per vector
    (size-independent identity; both sides from the device)."""
    from eigenfns_amd.linalg import rowdot
    from eigenfns_amd.operator import MaxwellOperator
    from eigenfns_amd.structure import network_eps
    eps, L = network_eps(216, 2, 96)
    op = MaxwellOperator(eps, L, (0.1, 0.2, 0.0))
    X = op.random_block(4, seed=3)
    lam = np.array([0.3, 1.0, 2.0, 4.0])
    ed = op.energy_density(X, lam).double().sum(dim=(1, 2, 3)).cpu().numpy()
    quad = rowdot(X, op.theta(X)).real.cpu().numpy()
    assert np.allclose(ed, quad / (96 ** 3 * lam), rtol=2e-5), (ed, quad / (96 ** 3 * lam))


def test_chunked_passes_equal_single_pass():
    """Blocks larger than the workspace run chunk by chunk (alternating
    workspaces and lanes when the two-lane pipeline is built in): results are
    bitwise those of one pass, for every fused epilogue."""
    import torch
    from eigenfns_amd.operator import MaxwellOperator
    rng = np.random.default_rng(11)
    eps = np.where(rng.random((16, 16, 16)) < 0.3, 8.41, 1.0).astype(np.float32)
    big = MaxwellOperator(eps, 1.3, (0.1, 0.0, 0.2))
    small = MaxwellOperator(eps, 1.3, (0.1, 0.0, 0.2), max_chunk=1)
    X = big.random_block(5, seed=2)
    assert torch.equal(big.theta(X), small.theta(X))
    assert torch.equal(big.precondition(X), small.precondition(X))
    outs = []
    for op in (big, small):
        T0 = X.roll(1, 0).contiguous()
        Y = torch.zeros_like(X)
        op.cheb_step(X, T0, Y, 0.01, 0.5)
        mom = torch.zeros((5, 8), dtype=torch.float64, device=X.device)
        T0k = X.roll(2, 0).contiguous()
        op.kpm_step(X, T0k, mom, 2, 0.01)
        est = torch.zeros(5, dtype=torch.float64, device=X.device)
        T0c = X.roll(3, 0).contiguous()
        op.kpm_count_step(X, T0c, X.roll(4, 0).contiguous(), est, 0.25, 0.01)
        outs.append((T0, Y, T0k, mom, T0c, est))
    for a, b in zip(*outs):
        assert torch.equal(a, b)
