"""The completeness-audit oracle (oracle/audit_np.py) pinned by identities that
do not restate it: a dense FP64 eigendecomposition for the per-probe
estimator, and the flat density of states for the edge-leakage integral."""
import numpy as np
import pytest
from scipy.integrate import quad

from oracle.audit_np import audit_verdict, bandpass_rho, bandpass_trace, edge_leakage
from oracle.dense_np import dense_eigh
from oracle.operator_np import MaxwellOperatorNP
from oracle.rng_np import random_block


@pytest.fixture(scope="module")
def dense8():
    G = 8
    rng = np.random.default_rng(2)
    eps = np.where(rng.random((G, G, G)) < 0.3, 8.41, 1.0)
    op = MaxwellOperatorNP(eps, 1.0)
    w, V = dense_eigh(op)
    return op, w, V


def _window(w):
    # edges in the two widest gaps among the states 100..400
    gaps = np.diff(w[100:400])
    i, j = sorted(np.argsort(gaps)[-2:] + 100)
    return 0.5 * (w[i] + w[i + 1]), 0.5 * (w[j] + w[j + 1]), i + 1, j + 1


def test_bandpass_trace_equals_dense_projection(dense8):
    op, w, V = dense8
    lam_max = 1.02 * w[-1]
    lo, hi, i0, i1 = _window(w)
    Z = random_block("rademacher", 5, 0, 3, 8, op.mask)
    est = bandpass_trace(op, Z, lo, hi, lam_max, 80)
    rho = bandpass_rho(w, lo, hi, lam_max, 80)
    proj = np.abs(V.conj().T @ Z.reshape(3, -1).T) ** 2          # |⟨v_i, z⟩|²
    assert np.allclose(est, proj.T @ rho, rtol=1e-10)
    # deflation against the window's eigenvectors removes exactly their terms
    locked = V[:, i0:i1].T.reshape(-1, 2, 8, 8, 8)
    est_d = bandpass_trace(op, Z, lo, hi, lam_max, 80, locked=locked)
    keep = np.ones(len(w), bool)
    keep[i0:i1] = False
    assert np.allclose(est_d, proj[keep].T @ rho[keep], rtol=1e-8, atol=1e-9)
    assert est.mean() > est_d.mean() + 0.5 * (i1 - i0)


def test_edge_leakage_flat_density_closed_form():
    """Moments of a flat density (N states uniform in λ ∈ [0, λmax]) give
    L = (N/λmax)·∫_outside ρ(λ) dλ over the two outside pieces."""
    N, lam_max, p, degree = 5000.0, 10.0, 400, 300
    k = np.arange(p + 1)
    mu = np.where(k % 2 == 0, N / np.where(k == 1, 1, 1 - k.astype(float) ** 2), 0.0)
    mom = np.vstack([mu, mu])
    lo, hi = 4.0, 4.8
    L, L_se = edge_leakage(mom, lam_max, lo, hi, lam_max, degree, n_grid=40000)
    f = lambda x: float(bandpass_rho([x], lo, hi, lam_max, degree)[0])
    exact = N / lam_max * (quad(f, lo - 0.6, lo, limit=400)[0] + quad(f, hi, hi + 0.6, limit=400)[0])
    assert L_se == 0.0
    assert abs(L - exact) < 2e-3 * abs(exact) + 1e-3, (L, exact)
    # the filter is ~1 inside: the window's own weight is N/λmax·(hi − lo)
    inside = N / lam_max * quad(f, lo, hi, limit=400)[0]
    assert abs(inside - N / lam_max * (hi - lo)) < 0.05 * N / lam_max * (hi - lo)


def test_audit_verdict_rule():
    r = {"deflated_estimate": 3.2, "se": 0.2, "predicted_edge_leakage": 2.9, "leakage_se": 0.1}
    v = audit_verdict(r)
    assert v["pass"] and abs(v["missed_window"] - 0.3) < 1e-12
    r2 = dict(r, deflated_estimate=4.2)
    assert not audit_verdict(r2)["pass"]
    assert not audit_verdict(r, sub=dict(r, deflated_estimate=3.6))["pass"]
