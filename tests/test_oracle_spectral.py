"""Pins the oracle's spectral methods (Chebyshev coefficients, filter, RR, KPM, Lanczos,
LOBPCG) against dense FP64 eigendecompositions, closed forms and quadrature."""
import numpy as np
import pytest

from oracle import chebyshev_np, interior_np, kpm_np, linalg_np, lobpcg_np
from oracle.dense_np import dense_eigh, dense_theta
from oracle.operator_np import MaxwellOperatorNP
from oracle.rng_np import random_block


@pytest.fixture(scope="module")
def small():
    """G=8 disordered ε (tests/test_solver.py:20-33 of the reference, seed 3)."""
    rng = np.random.default_rng(3)
    G, L = 8, 1.0
    eps = np.asarray(1.0 + 7.57 * (rng.random((G, G, G)) > 0.7), np.float32)
    op = MaxwellOperatorNP(eps, L)
    w, V = dense_eigh(op)
    return op, w, V


# ---- closed forms -------------------------------------------------------------
def test_jackson_closed_form_properties():
    for p in (1, 10, 300):
        g = interior_np.jackson(p)
        assert abs(g[0] - 1.0) < 1e-14
        assert np.all(np.diff(g) <= 1e-14)            # monotone damping
        assert np.all(g > -1e-14)
        # the Jackson kernel Σ g_k T_k(cos θ) is a non-negative density (Weiße et al. 2006)
        th = np.linspace(0, np.pi, 513)
        ker = g[0] + 2 * np.cos(np.outer(np.arange(1, p + 1), th)).T @ g[1:]
        assert ker.min() > -1e-9 * ker.max()


@pytest.mark.parametrize("xb", [-0.7, 0.0, 0.31, 0.95])
def test_step_coeffs_equal_chebyshev_projection_of_indicator(xb):
    # c_k = (2 − δ_k0)/π ∫_{arccos xb}^{π} cos(kθ) dθ  — midpoint quadrature in θ
    p = 40
    tb = np.arccos(xb)
    th = tb + (np.arange(200000) + 0.5) * (np.pi - tb) / 200000
    w = (np.pi - tb) / 200000
    quad = np.array([(2 - (k == 0)) / np.pi * np.cos(k * th).sum() * w for k in range(p + 1)])
    assert np.abs(interior_np.step_coeffs(xb, p) - quad).max() < 1e-9


def test_bandpass_is_difference_of_steps_and_selects_window():
    lam_max, lo, hi, p = 100.0, 20.0, 35.0, 400
    c = interior_np.bandpass_coeffs(lo, hi, lam_max, p)
    lam = np.array([5.0, 27.0, 60.0])
    x = 2 * lam / lam_max - 1
    val = np.polynomial.chebyshev.chebval(x, c)
    assert val[1] > 0.99 and abs(val[0]) < 1e-3 and abs(val[2]) < 1e-3


# ---- filter, RR, KPM against the dense spectrum ---------------------------------
def test_apply_cheb_poly_equals_dense_polynomial(small):
    op, w, V = small
    lam_max = 1.05 * w.max()
    coef = interior_np.bandpass_coeffs(w[5], w[40], lam_max, 60)
    X = random_block("normal", 1, 0, 2, 8, op.mask).astype(np.complex128)
    Y = interior_np.apply_cheb_poly(op, X, coef, lam_max)
    x = 2 * w / lam_max - 1
    p = np.polynomial.chebyshev.chebval(x, coef)
    Xf = X.reshape(2, -1)
    Yd = (V @ (p[:, None] * (V.conj().T @ Xf.T))).T
    Yd /= np.linalg.norm(Yd, axis=1, keepdims=True)
    assert np.abs(Y.reshape(2, -1) - Yd).max() < 1e-10


def test_rr_extract_on_invariant_subspace(small):
    # reference tests/test_interior.py:80-92: rr on a span of true eigenvectors
    op, w, V = small
    Rot = np.linalg.qr(np.random.default_rng(0).standard_normal((6, 6)))[0]
    Q = Rot @ V[:, :6].T
    lam, Xr, rn = interior_np.rr_extract(op, Q.reshape(6, 2, 8, 8, 8))
    assert rn.max() < 1e-10
    assert np.allclose(lam, w[:6], rtol=1e-12)


def test_kpm_doubling_equals_direct_and_dense_moments(small):
    op, w, V = small
    lam_max = 1.05 * w.max()
    Z = random_block("rademacher", 4, 0, 2, 8, op.mask)
    dbl = kpm_np.kpm_moments(op, 30, 2, 4, lam_max, probes=Z)
    direct = kpm_np.kpm_moments_direct(op, 31, Z, lam_max)
    assert np.abs(dbl - direct).max() < 1e-9 * np.abs(direct).max()
    # μ_k = Σ_i |⟨v_i, z⟩|² T_k(x_i)
    x = 2 * w / lam_max - 1
    wts = np.abs(V.conj().T @ Z.reshape(2, -1).T.astype(np.complex128)) ** 2
    T = np.cos(np.outer(np.arange(32), np.arccos(x)))
    dense_mom = (T @ wts).T
    assert np.abs(dbl - dense_mom).max() < 1e-9 * np.abs(dense_mom).max()
    assert np.allclose(dbl[:, 0], 2 * 8 ** 3 - 2)


def test_kpm_counting_at_top_counts_every_state(small):
    op, w, V = small
    lam_max = 1.05 * w.max()
    mom = kpm_np.kpm_moments(op, 40, 3, 9, lam_max)
    mean, se, _ = kpm_np.counting(mom, lam_max, [lam_max, 0.999 * lam_max])
    assert abs(mean[0] - (2 * 8 ** 3 - 2)) < 1e-9 * 1022   # the step is ≡ 1 at λ_max
    assert abs(mean[1] - (2 * 8 ** 3 - 2)) < 0.01          # Jackson smearing near the edge


def test_kpm_count_below_tracks_exact_count(small):
    op, w, V = small
    lam_max = 1.05 * w.max()
    lam_b = 0.5 * (w[300] + w[301])
    mean, se, _ = chebyshev_np.kpm_count_below(op, lam_b, lam_max, degree=200, n_probe=6, seed=2)
    assert abs(mean - 301) < 6 * se + 10


def test_lanczos_lambda_max_bounds(small):
    op, w, V = small
    est = chebyshev_np.lanczos_lambda_max(op)
    bound = float((op.kn ** 2).max() * op.inv_eps.max())
    assert est > 0.9 * w.max()
    assert est < 1.3 * bound


# ---- LOBPCG against the dense spectrum -----------------------------------------
def test_lobpcg_matches_dense(small):
    # reference tests/test_solver.py:36-48 tightened to 1e-8 (FP64 oracle)
    op, w, V = small
    vals, vecs, st = lobpcg_np.lobpcg_blocks(op, 24, m=16, guard=8, tol=1e-8, maxit=400)
    assert (np.abs(vals - w[:24]) / w[:24]).max() < 1e-10
    assert (np.diff(vals) > -1e-12).all()
    Vf = vecs.reshape(24, -1)
    assert np.abs(Vf.conj() @ Vf.T - np.eye(24)).max() < 1e-8


def test_lobpcg_homogeneous_analytic():
    G, L, e0 = 8, 1.0, 2.0
    op = MaxwellOperatorNP(np.full((G, G, G), e0, np.float32), L)
    vals, _, _ = lobpcg_np.lobpcg_blocks(op, 10, m=16, guard=6, tol=1e-9)
    k = 2 * np.pi * np.fft.fftfreq(G, d=L / G)
    g2 = (np.array(np.meshgrid(k, k, k, indexing="ij")) ** 2).sum(0).ravel()
    inv = float(op.inv_eps.flat[0])
    ana = np.sort(np.concatenate([g2[g2 > 0] * inv] * 2))[:10]
    assert np.abs(vals - ana).max() < 1e-9 * ana.max()
