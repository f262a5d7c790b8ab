"""Parity of the fused Chebyshev, KPM-moment and KPM-count steps on the
interior/KPM workload's own structure — the 10,000-point rod network (seed 3,
rod radius 0.331836 × mean bond length, ε = 8.41/1) — at 64³, 192³ and 256³,
where the z passes use the wave-row layout and the workspace holds one vector
(so multi-vector blocks run chunk by chunk through the FP64 partial reduce).

Bars, against the FP64 oracle on the same device-drawn inputs:
  vectors (T_{k+1}, Y)        relative L2 per vector < 1e-5
  moments μ_{2k}, μ_{2k+1}    |Δ| < 1e-5 · 2‖T_k‖‖T_{k+1}‖ (the scale of the dot)
  count estimate increment    |Δ| < 1e-5 · |c_k|·‖z‖‖T_{k+1}‖
  μ₀ of Rademacher probes     exactly 2G³ − 2
Reference: scripts/run_interior.py:154-216 (the filter steps of
interior.py:336-354) and scripts/exp/exp_kpm_dos.py:125-163.
"""
import numpy as np
import pytest

from oracle import kpm_np
from oracle.operator_np import MaxwellOperatorNP

pytestmark = pytest.mark.gpu

TOL = 1e-5
C4_NET = dict(n_points=10000, seed=3, radius_factor=0.331836)


@pytest.fixture(scope="module", params=[64, 192, 256])
def c4(request):
    import torch
    from eigenfns_amd.operator import MaxwellOperator
    from eigenfns_amd.structure import network_eps
    G = request.param
    eps_dev, L = network_eps(grid_size=G, on_device=True, **C4_NET)
    eps = eps_dev.cpu().numpy()
    op = MaxwellOperator(eps_dev, L)
    ref = MaxwellOperatorNP(eps, L)
    lam_max = 1.05 * float((ref.kn ** 2).max() * ref.inv_eps.max())
    print(f"C4 network at {G}³: L = {L:.4f}, filling {np.mean(eps != 1.0):.3f}, "
          f"chunk {op.chunk}, λ_max {lam_max:.4g}", flush=True)
    yield G, op, ref, lam_max
    del op
    torch.cuda.empty_cache()


def _m(G):
    return 2 if G >= 192 else 3


def _rel(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def _norms(A):
    return np.linalg.norm(np.asarray(A).reshape(A.shape[0], -1).astype(np.complex128), axis=1)


def test_cheb_step_at_scale(c4):
    G, op, ref, lam_max = c4
    m = _m(G)
    alpha, ck = 2.0 / lam_max, -0.0731
    T1 = op.random_block(m, seed=41)
    T0 = op.random_block(m, seed=42)
    Y = op.random_block(m, seed=43)
    t1, t0, y = (a.cpu().numpy().astype(np.complex128) for a in (T1, T0, Y))
    op.cheb_step(T1, T0, Y, alpha, ck)
    T2 = 2 * (alpha * ref.theta(t1) - t1) - t0
    r1 = _rel(T0.cpu().numpy(), T2)
    r2 = _rel(Y.cpu().numpy(), y + ck * T2)
    print(f"G={G} cheb_step: T_k+1 {r1.max():.2e}, Y {r2.max():.2e}")
    assert r1.max() < TOL and r2.max() < TOL


def test_kpm_step_moments_at_scale(c4):
    import torch
    G, op, ref, lam_max = c4
    m = _m(G)
    alpha = 2.0 / lam_max
    T1 = op.random_block(m, seed=51)
    T0 = op.random_block(m, seed=52)
    t1, t0 = (a.cpu().numpy().astype(np.complex128) for a in (T1, T0))
    k = 3
    mom = torch.zeros((m, 2 * k + 2), dtype=torch.float64, device=T1.device)
    mu01 = np.array([[1234.5, -17.25]] * m)
    mom[:, :2] = torch.from_numpy(mu01).to(T1.device)
    op.kpm_step(T1, T0, mom, k, alpha)
    T2 = 2 * (alpha * ref.theta(t1) - t1) - t0
    flat1, flat2 = t1.reshape(m, -1), T2.reshape(m, -1)
    want_even = 2 * np.real(np.sum(flat1.conj() * flat1, axis=1)) - mu01[:, 0]
    want_odd = 2 * np.real(np.sum(flat2.conj() * flat1, axis=1)) - mu01[:, 1]
    got = mom.cpu().numpy()
    scale = 2 * _norms(t1) * _norms(T2)
    e_even = np.abs(got[:, 2 * k] - want_even) / scale
    e_odd = np.abs(got[:, 2 * k + 1] - want_odd) / scale
    rv = _rel(T0.cpu().numpy(), T2)
    print(f"G={G} kpm_step: T_k+1 {rv.max():.2e}, μ_2k {e_even.max():.2e}, μ_2k+1 {e_odd.max():.2e}")
    assert rv.max() < TOL and e_even.max() < TOL and e_odd.max() < TOL


def test_kpm_count_step_at_scale(c4):
    import torch
    G, op, ref, lam_max = c4
    m = _m(G)
    alpha, ck = 2.0 / lam_max, 0.3177
    T1 = op.random_block(m, seed=61)
    T0 = op.random_block(m, seed=62)
    Z = op.random_block(m, seed=63, kind="rademacher")
    t1, t0, z = (a.cpu().numpy().astype(np.complex128) for a in (T1, T0, Z))
    est0 = np.array([0.5 * i - 3.0 for i in range(m)])
    est = torch.from_numpy(est0.copy()).to(T1.device)
    op.kpm_count_step(T1, T0, Z, est, ck, alpha)
    T2 = 2 * (alpha * ref.theta(t1) - t1) - t0
    want = est0 + ck * np.real(np.sum(z.reshape(m, -1).conj() * T2.reshape(m, -1), axis=1))
    err = np.abs(est.cpu().numpy() - want) / (abs(ck) * _norms(z) * _norms(T2))
    rv = _rel(T0.cpu().numpy(), T2)
    print(f"G={G} kpm_count_step: T_k+1 {rv.max():.2e}, estimate {err.max():.2e}")
    assert rv.max() < TOL and err.max() < TOL


def test_kpm_moments_short_run_at_scale(c4):
    """A short KPM run (degree 16, one Rademacher probe) against the oracle's
    doubling recurrence on the same probe; μ₀ = 2G³ − 2 exactly at Γ."""
    from eigenfns_amd.kpm import kpm_moments
    G, op, ref, lam_max = c4
    if G == 64:
        pytest.skip("the short run is the 256³ / 192³ case; 64³ is covered by the step tests")
    deg = 16
    mom = kpm_moments(op, degree=deg, n_probe=1, seed=9, lam_max=lam_max)
    Z = op.random_block(1, seed=9, kind="rademacher").cpu().numpy()
    mref = kpm_np.kpm_moments(ref, deg, 1, 9, lam_max, probes=Z)
    assert mom.shape == mref.shape == (1, deg + 2)
    assert mom[0, 0] == 2 * G ** 3 - 2
    err = np.abs(mom - mref).max() / np.abs(mref).max()
    print(f"G={G} KPM degree {deg}: max moment error {err:.2e} of max|μ|")
    assert err < TOL
