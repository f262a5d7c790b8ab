"""GPU completeness audit (exp_i2_v2.py) against the oracle and a dense solve.

Bars: per-probe bandpass traces equal to the FP64 oracle's on the same
counter-based probes (with and without deflation) to 1e-5 of the largest
estimate; the end-to-end gate at G = 8 passes when every window state is
locked and fails when one is withheld.
"""
import numpy as np
import pytest

from oracle.audit_np import bandpass_trace as np_bandpass_trace
from oracle.dense_np import dense_eigh
from oracle.operator_np import MaxwellOperatorNP
from oracle.rng_np import random_block

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    G = 8
    rng = np.random.default_rng(2)
    eps = np.where(rng.random((G, G, G)) < 0.3, 8.41, 1.0).astype(np.float32)
    ref = MaxwellOperatorNP(eps, 1.0)
    w, V = dense_eigh(ref)
    gaps = np.diff(w[100:400])
    i, j = sorted(np.argsort(gaps)[-2:] + 100)
    lo, hi = 0.5 * (w[i] + w[i + 1]), 0.5 * (w[j] + w[j + 1])
    return eps, ref, w, V, lo, hi, i + 1, j + 1


def test_bandpass_trace_matches_oracle(case):
    import torch
    from eigenfns_amd.audit import bandpass_trace
    from eigenfns_amd.operator import MaxwellOperator
    eps, ref, w, V, lo, hi, i0, i1 = case
    op = MaxwellOperator(eps, 1.0)
    lam_max = 1.02 * w[-1]
    Z = random_block("rademacher", 9, 0, 6, 8, ref.mask)
    locked_np = V[:, i0:i1].T.reshape(-1, 2, 8, 8, 8)
    locked = torch.from_numpy(locked_np.astype(np.complex64)).cuda()
    for lk, lkn in ((None, None), (locked, locked_np.astype(np.complex64).astype(np.complex128))):
        got = bandpass_trace(op, lo, hi, lam_max, 120, 6, seed=9, locked=lk)
        want = np_bandpass_trace(ref, Z, lo, hi, lam_max, 120, locked=lkn)
        assert np.abs(got - want).max() < 1e-5 * np.abs(want).max() + 1e-6, (got, want)


def test_completeness_gate_detects_a_missing_state(case):
    import torch
    from eigenfns_amd.audit import completeness_audit
    from eigenfns_amd.operator import MaxwellOperator
    eps, ref, w, V, lo, hi, i0, i1 = case
    op = MaxwellOperator(eps, 1.0)
    lam_max = 1.02 * w[-1]
    # exact KPM moments of the dense spectrum (2 identical rows: no sampling
    # error), resolved 4× finer than the filter; the leakage margin covers the
    # filter's transition tails at this λ scale (≈ (π/300)·λmax/2 ≈ 8)
    p = 1200
    x = 2 * w / lam_max - 1
    mu = np.cos(np.outer(np.arange(p + 1), np.arccos(x))).sum(1)
    kpm = (np.vstack([mu, mu]), lam_max)
    locked = torch.from_numpy(V[:, i0:i1].T.reshape(-1, 2, 8, 8, 8).astype(np.complex64)).cuda()
    full = completeness_audit(op, locked, w[i0:i1], lo, hi, lam_max, degree=300, n_probe=64,
                              seed=3, kpm=kpm, chunk=32, leak_margin=50.0)
    print("all locked:", full["missed_window"], full["missed_window_se"])
    assert full["pass"]
    mid = (i0 + i1) // 2
    keep = [r for r in range(i1 - i0) if r != mid - i0]
    short = completeness_audit(op, locked[keep], np.delete(w[i0:i1], mid - i0), lo, hi, lam_max,
                               degree=300, n_probe=64, seed=3, kpm=kpm, chunk=32,
                               leak_margin=50.0)
    print("one withheld:", short["missed_window"], short["missed_window_se"])
    assert not short["pass"]
    assert abs(short["missed_window"] - 1.0) < 4 * short["missed_window_se"] + 0.1
