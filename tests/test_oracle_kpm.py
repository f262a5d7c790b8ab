"""Pins the oracle's KPM analysis (scripts/exp/exp_kpm_analyze.py:39-70):
``dos_curve``, ``smearing_width`` and their consistency with ``counting``,
against identities that do not restate the formulas — Chebyshev orthogonality
on the Chebyshev–Gauss grid, the exact mean and variance of a Jackson-damped
delta, and the derivative relation between the DOS and the counting function.
The product's host analysis (``eigenfns_amd.kpm``) is checked against the
pinned oracle on the same moments."""
import numpy as np
import pytest

from oracle import kpm_np
from oracle.interior_np import jackson


def _theta_weights(lam, lam_max):
    """dλ per grid point of dos_curve's Chebyshev–Gauss grid (midpoints in θ)."""
    n = lam.size
    x = 2 * lam / lam_max - 1
    return (lam_max / 2) * np.sqrt(1 - x ** 2) * (np.pi / n)


def _delta_moments(a, p):
    """Moments of a single state at x = a: μ_k = T_k(a)."""
    return np.cos(np.arange(p + 1) * np.arccos(a))[None, :]


@pytest.mark.parametrize("p,n_pts", [(10, 200), (300, 4000), (1999, 4000)])
def test_dos_integral_equals_mu0(p, n_pts):
    # Σ_j cos(kθ_j) = 0 on the midpoint grid for 0 < k < 2n, so the quadrature of
    # the returned density over λ is exactly the mean μ₀ whatever the other moments
    rng = np.random.default_rng(p)
    mom = rng.standard_normal((5, p + 1))
    mom[:, 0] = 1000.0 + rng.random(5)
    lam_max = 37.5
    lam, rho = kpm_np.dos_curve(mom, lam_max, n_pts=n_pts)
    assert lam.shape == rho.shape == (n_pts,)
    assert np.all(np.diff(lam) > 0)
    total = float(np.sum(rho * _theta_weights(lam, lam_max)))
    assert abs(total - mom[:, 0].mean()) < 1e-10 * mom[:, 0].mean()


@pytest.mark.parametrize("a", [-0.62, 0.0, 0.35, 0.9])
@pytest.mark.parametrize("p", [40, 400])
def test_dos_of_one_state_is_the_jackson_kernel(a, p):
    # ∫ρ T_j dx = g_j T_j(a) by orthogonality: the mean of the damped delta is
    # g₁·a and its second moment (1 + g₂T₂(a))/2, with g₁ = cos(π/(p+1)) in
    # closed form; the midpoint grid integrates these polynomials exactly
    lam_max = 10.0
    lam, rho = kpm_np.dos_curve(_delta_moments(a, p), lam_max, n_pts=4 * p)
    w = rho * _theta_weights(lam, lam_max)
    x = 2 * lam / lam_max - 1
    g = jackson(p)
    assert abs(g[1] - np.cos(np.pi / (p + 1))) < 1e-13
    assert abs(np.sum(w) - 1.0) < 1e-12
    assert abs(np.sum(w * x) - np.cos(np.pi / (p + 1)) * a) < 1e-12
    assert abs(np.sum(w * x ** 2) - (1 + g[2] * (2 * a * a - 1)) / 2) < 1e-12
    # the density peaks at the state
    assert abs(x[np.argmax(rho)] - a) < 4 * np.pi / p


@pytest.mark.parametrize("a", [-0.7, 0.0, 0.5])
@pytest.mark.parametrize("p", [200, 2000, 12000])
def test_smearing_width_is_the_kernel_standard_deviation(a, p):
    # exact variance of the damped delta from the two moments above, in λ units;
    # the Jackson kernel's width is (π/p)·√(1 − a²) to leading order in 1/p
    lam_max = 2.0
    g = jackson(p)
    var_x = (1 + g[2] * (2 * a * a - 1)) / 2 - (g[1] * a) ** 2
    sigma = np.sqrt(var_x) * lam_max / 2
    lam_a = (a + 1) * lam_max / 2
    w = kpm_np.smearing_width(lam_a, lam_max, p)
    assert abs(w / sigma - 1) < 4.0 / p
    # in θ the width does not depend on the position (π/p)
    th_w = w / (lam_max / 2) / np.sqrt(1 - a * a)
    assert abs(th_w - np.pi / p) < 1e-12


def test_dos_integrates_to_the_counting_function():
    # N(λ_b) = Σ g_k μ_k c_k(x_b) is the integral of the damped density below λ_b
    rng = np.random.default_rng(0)
    p, lam_max = 60, 5.0
    states = rng.uniform(0.02, 0.98, 40)
    mom = np.stack([np.cos(np.arange(p + 1) * np.arccos(2 * s - 1)) for s in states]).sum(0)[None]
    mom = np.repeat(mom, 2, axis=0)
    lam, rho = kpm_np.dos_curve(mom, lam_max, n_pts=200000)
    w = rho * _theta_weights(lam, lam_max)
    for lb in (0.7, 2.5, 4.1):
        mean, _, _ = kpm_np.counting(mom, lam_max, [lb])
        assert abs(np.sum(w[lam < lb]) - mean[0]) < 2e-3


def test_product_analysis_matches_oracle():
    from eigenfns_amd import kpm
    rng = np.random.default_rng(5)
    p = 120
    mom = rng.standard_normal((4, p + 2))
    mom[:, 0] = 500.0
    lam_max = 12.0
    lams = [0.5, 3.0, 11.9]
    for a, b in zip(kpm.counting(mom, lam_max, lams), kpm_np.counting(mom, lam_max, lams)):
        assert np.allclose(a, b, rtol=1e-13, atol=1e-10)
    la, ra = kpm.dos_curve(mom, lam_max, n_pts=1000)
    lb, rb = kpm_np.dos_curve(mom, lam_max, n_pts=1000)
    assert np.array_equal(la, lb)
    assert np.allclose(ra, rb, rtol=1e-12, atol=1e-12 * np.abs(rb).max())
    lam = np.linspace(0.1, 11.9, 7)
    assert np.allclose(kpm.smearing_width(lam, lam_max, p), kpm_np.smearing_width(lam, lam_max, p),
                       rtol=1e-14)
