"""Pins the oracle's operator (oracle/operator_np.py) independently of its formula.

Mirrors the reference's tests/test_operator.py:27-65 (analytic homogeneous
spectrum, Hermitian/PSD on disordered ε, Bloch plane waves) and adds checks a
transposed operand, a wrong sign or a dropped term would fail.
"""
import numpy as np
import pytest

from oracle.dense_np import dense_theta
from oracle.operator_np import MaxwellOperatorNP, make_basis, swap_mask


def _rand_block(G, m, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((m, 2, G, G, G)) + 1j * rng.standard_normal((m, 2, G, G, G))


@pytest.mark.parametrize("k", [(0, 0, 0), (0.5, 0, 0), (0.13, -0.27, 0.41)])
def test_frame_is_orthonormal_and_right_handed(k):
    G, L = 8, 1.3
    kn, t, K = make_basis(G, L, k)
    t1, t2 = t
    live = kn > 0
    khat = K / np.where(live, kn, 1.0)
    assert np.allclose((t1 * t1).sum(0)[live], 1, atol=1e-14)
    assert np.allclose((t2 * t2).sum(0)[live], 1, atol=1e-14)
    assert np.allclose((t1 * t2).sum(0)[live], 0, atol=1e-14)
    assert np.allclose((t1 * K).sum(0)[live], 0, atol=1e-12)
    # right-handed (k̂, t̂₁, t̂₂): t̂₂ = k̂ × t̂₁
    assert np.allclose(np.cross(khat, t1, axis=0)[:, live], t2[:, live], atol=1e-14)
    assert np.all(t[:, :, ~live] == 0)


def test_swap_rule_is_the_exact_integer_test_at_gamma():
    # |K_z| > 0.9|K|  ⇔  19·n_z² > 81·(n_x² + n_y²) at k = 0 (integer reciprocal indices)
    G, L = 16, 2.7
    n = np.fft.fftfreq(G, 1.0 / G).round().astype(int)
    NX, NY, NZ = np.meshgrid(n, n, n, indexing="ij")
    exact = 19 * NZ ** 2 > 81 * (NX ** 2 + NY ** 2)
    assert np.array_equal(swap_mask(G, L), exact)


def test_homogeneous_known_answer_per_coefficient():
    # Θ·H = (|k+G|²/ε₀)·H on every transverse coefficient
    G, L, e0 = 8, 1.7, 3.3
    for k in [(0, 0, 0), (0.3, 0.1, -0.2)]:
        op = MaxwellOperatorNP(np.full((G, G, G), e0, np.float32), L, k)
        X = _rand_block(G, 2, 1)
        if k == (0, 0, 0):
            X[:, :, 0, 0, 0] = 0
        Y = op.theta(X)
        expect = (op.kn ** 2 * float(op.inv_eps.flat[0]))[None, None] * X  # FP32 ε⁻¹ table
        assert np.abs(Y - expect).max() < 1e-11 * np.abs(expect).max()


def test_homogeneous_preconditioner_known_answer():
    # K⁻¹ P ε P K⁻¹ with homogeneous ε₀ is (ε₀/knr²)·I on each transverse coefficient
    G, L, e0 = 8, 1.1, 2.25
    op = MaxwellOperatorNP(np.full((G, G, G), e0, np.float32), L, (0.2, 0, 0))
    X = _rand_block(G, 2, 2)
    knr = np.maximum(op.kn, op.kn_floor)
    expect = (e0 / knr ** 2)[None, None] * X
    Y = op.precondition(X)
    assert np.abs(Y - expect).max() < 1e-10 * np.abs(expect).max()


def test_diagonal_preconditioner_known_answer():
    # A one-hot coefficient at reciprocal-lattice point n (Bloch k) comes back
    # scaled by 1/((2π/L)²|n + k|² + s), n counted as in np.fft.fftfreq, with the
    # shift s = target + 1e-6 for a target above 1e-6·⟨1/ε⟩ and
    # (⟨1/ε⟩ + 1)·1e-6 for target 0 (operator.py:98-101, 163-166).
    G, L, k = 6, 1.7, (0.25, -0.5, 0.1)
    rng = np.random.default_rng(2)
    eps = np.where(rng.random((G, G, G)) < 0.4, 6.0, 1.0).astype(np.float32)
    op = MaxwellOperatorNP(eps, L, k)
    for (ix, iy, iz), comp in (((0, 0, 0), 0), ((1, 5, 3), 1), ((4, 2, 0), 0)):
        n = np.array([ix if ix < G // 2 else ix - G, iy if iy < G // 2 else iy - G,
                      iz if iz < G // 2 else iz - G], float)
        q2 = ((2 * np.pi / L) ** 2) * ((n + np.asarray(k)) ** 2).sum()
        X = np.zeros((1, 2, G, G, G), complex)
        X[0, comp, ix, iy, iz] = 1.0 - 2.0j
        for target, s in ((0.0, (op.eps_mean_inv + 1.0) * 1e-6), (3.0, 3.0 + 1e-6)):
            Y = op.precondition_simple(X, target)
            assert np.count_nonzero(Y) == 1
            assert abs(Y[0, comp, ix, iy, iz] - (1.0 - 2.0j) / (q2 + s)) < 1e-6 / (q2 + s)


def test_dense_homogeneous_spectrum_matches_analytic():
    # reference tests/test_operator.py:27-38
    G, L, e0 = 8, 1.0, 4.0
    op = MaxwellOperatorNP(np.full((G, G, G), e0, np.float32), L)
    M = dense_theta(op)
    w = np.linalg.eigvalsh((M + M.conj().T) / 2)
    k = 2 * np.pi * np.fft.fftfreq(G, d=L / G)
    g2 = (np.array(np.meshgrid(k, k, k, indexing="ij")) ** 2).sum(0).ravel()
    ana = np.sort(np.concatenate([g2[g2 > 0] / e0] * 2))
    assert (np.abs(w) < 1e-5).sum() == 2
    assert (np.abs(np.sort(w)[2:] - ana) / ana).max() < 5e-12


def test_dense_disordered_is_hermitian_psd():
    # reference tests/test_operator.py:41-50
    rng = np.random.default_rng(0)
    G, L = 8, 1.0
    eps = np.asarray(1.0 + 7.57 * (rng.random((G, G, G)) > 0.7), np.float32)
    M = dense_theta(MaxwellOperatorNP(eps, L))
    assert np.abs(M - M.conj().T).max() / np.abs(M).max() < 1e-12
    w = np.linalg.eigvalsh((M + M.conj().T) / 2)
    assert w.min() > -1e-10 * abs(w).max()


def test_dense_bloch_shifted_plane_waves():
    # reference tests/test_operator.py:53-65
    G, L, e0 = 6, 1.0, 2.25
    kf = (0.5, 0.0, 0.0)
    op = MaxwellOperatorNP(np.full((G, G, G), e0, np.float32), L, kf)
    w = np.sort(np.linalg.eigvalsh(dense_theta(op)))
    k1 = 2 * np.pi * np.fft.fftfreq(G, d=L / G)
    k0 = 2 * np.pi * np.asarray(kf) / L
    KX, KY, KZ = np.meshgrid(k1 + k0[0], k1 + k0[1], k1 + k0[2], indexing="ij")
    inv = float(op.inv_eps.flat[0])  # the FP32 ε⁻¹ table value
    ana = np.sort(np.concatenate([(KX ** 2 + KY ** 2 + KZ ** 2).ravel() * inv] * 2))
    assert (np.abs(w - ana) / np.maximum(ana, 1e-12)).max() < 5e-12


def test_fft_convention_against_brute_force_dft():
    # Θ written with explicit DFT matrices: ifftn carries 1/G³, fftn is unnormalised
    G, L = 4, 1.0
    rng = np.random.default_rng(5)
    eps = (1.0 + 7.0 * (rng.random((G, G, G)) > 0.5)).astype(np.float32)
    op = MaxwellOperatorNP(eps, L, (0.1, 0.2, 0.3))
    X = _rand_block(G, 1, 3)
    n = np.arange(G)
    F = np.exp(-2j * np.pi * np.outer(n, n) / G)          # forward DFT matrix
    Fi = np.conj(F) / G                                      # inverse with 1/G
    kn, t = op.kn, op.t
    a, b = X[0, 0], X[0, 1]
    Ec = (-1j * kn * b)[None] * t[0] + (1j * kn * a)[None] * t[1]
    E = np.einsum("ai,bj,ck,xijk->xabc", Fi, Fi, Fi, Ec)
    D = (1.0 / eps.astype(np.float64))[None] * E
    Dc = np.einsum("ai,bj,ck,xijk->xabc", F, F, F, D)
    f1, f2 = (t[0] * Dc).sum(0), (t[1] * Dc).sum(0)
    expect = np.stack([-1j * kn * f2, 1j * kn * f1])[None]
    assert np.abs(op.theta(X) - expect).max() < 1e-12 * np.abs(expect).max()


def test_fp32_restatement_tracks_fp64():
    # the complex64 restatement (CPU baseline) agrees with the FP64 oracle to FP32 accuracy
    G, L = 16, 2.0
    rng = np.random.default_rng(7)
    eps = np.where(rng.random((G, G, G)) < 0.3, 8.41, 1.0).astype(np.float32)
    X = _rand_block(G, 2, 4).astype(np.complex64)
    a = MaxwellOperatorNP(eps, L, fft_dtype=np.complex64, out_dtype=np.complex64).theta(X)
    b = MaxwellOperatorNP(eps, L).theta(X)
    rel = np.linalg.norm((a - b).reshape(2, -1), axis=1) / np.linalg.norm(b.reshape(2, -1), axis=1)
    assert rel.max() < 2e-6


def test_non_cubic_eps_raises():
    with pytest.raises(ValueError):
        MaxwellOperatorNP(np.ones((4, 4, 5), np.float32), 1.0)


# ---- real-space fields and energy density (operator.py:169-184, run_modes.py:116-122)

def _plane_wave(G, L, k, idx, pol):
    from oracle.operator_np import MaxwellOperatorNP
    op = MaxwellOperatorNP(np.full((G, G, G), 2.5), L, k)
    H = np.zeros((1, 2, G, G, G), np.complex128)
    H[(0, pol) + idx] = 1.0
    lam = float(op.kn[idx] ** 2 / 2.5)
    return op, H, lam


@pytest.mark.parametrize("k,idx,pol", [((0, 0, 0), (1, 2, 3), 0), ((0.2, -0.1, 0.3), (5, 0, 7), 1),
                                       ((0, 0, 0), (0, 0, 2), 1)])
def test_realspace_plane_wave_closed_forms(k, idx, pol):
    """A unit transverse coefficient in homogeneous ε₀: |H(r)|² = 1/G⁶ and
    ε₀|E(r)|² = 1/G⁶ at every voxel, E ⊥ H along the frame vectors."""
    G = 8
    op, H, lam = _plane_wave(G, 1.3, k, idx, pol)
    Hr = op.h_realspace(H)[0]
    assert np.allclose((np.abs(Hr) ** 2).sum(0), 1.0 / G ** 6, rtol=1e-12, atol=0)
    ed = op.energy_density(H, [lam])[0]
    assert np.allclose(ed, 1.0 / G ** 6, rtol=1e-12, atol=0)
    E = op.e_realspace(H, [lam])[0]
    # E·H* = 0 pointwise (transverse pair t̂₁, t̂₂ of the same wave)
    assert np.abs((E * Hr.conj()).sum(0)).max() < 1e-12 * np.abs(E).max() * np.abs(Hr).max()


@pytest.mark.parametrize("k", [(0.0, 0.0, 0.0), (0.31, 0.0, -0.17)])
def test_realspace_parseval_identities(k):
    """Disordered ε, random H, arbitrary λ: Σ_r|H(r)|² = ‖H‖²/G³ and
    Σ_r ε|E(r)|² = Re⟨H, ΘH⟩/(G³λ)."""
    from oracle.operator_np import MaxwellOperatorNP
    G = 10
    rng = np.random.default_rng(4)
    eps = np.where(rng.random((G, G, G)) < 0.4, 8.41, 1.0)
    op = MaxwellOperatorNP(eps, 2.0, k)
    H = (rng.standard_normal((2, 2, G, G, G)) + 1j * rng.standard_normal((2, 2, G, G, G))) * op.mask
    lam = np.array([0.7, 3.1])
    Hr = op.h_realspace(H)
    n2 = (np.abs(H) ** 2).reshape(2, -1).sum(1)
    assert np.allclose((np.abs(Hr) ** 2).reshape(2, -1).sum(1), n2 / G ** 3, rtol=1e-12)
    ed = op.energy_density(H, lam)
    TH = op.theta(H)
    quad = np.real(np.sum(H.conj() * TH, axis=(1, 2, 3, 4)))
    assert np.allclose(ed.reshape(2, -1).sum(1), quad / (G ** 3 * lam), rtol=1e-11)
    # a dropped ε⁻¹ or a missing 1/λ fails: the identity is specific
    assert not np.allclose(ed.reshape(2, -1).sum(1), quad / G ** 3, rtol=1e-3)
