"""GPU tests of the batched (multi-operator, continuously admitted) LOBPCG engine.

Bars: eigenvalues within 1e-6 relative of the FP64 dense solve for every
request in a mixed batch (different ε and Bloch k); a request's eigenvalues
bitwise identical alone and in a mix; multi-operator Θ identical to the
per-operator Θ.
"""
import numpy as np
import pytest

from oracle.dense_np import dense_eigh
from oracle.operator_np import MaxwellOperatorNP

pytestmark = pytest.mark.gpu


def _eps(seed, G=8):
    rng = np.random.default_rng(seed)
    return np.where(rng.random((G, G, G)) < 0.3, 8.41, 1.0).astype(np.float32)


PROBLEMS = [  # (eps seed, box, k, nev)
    (1, 1.0, (0.0, 0.0, 0.0), 20),
    (2, 1.3, (0.25, 0.0, 0.0), 16),
    (3, 0.9, (0.1, 0.2, 0.3), 24),
    (4, 1.1, (0.5, 0.5, 0.0), 12),
    (5, 1.0, (0.0, 0.0, 0.0), 8),
]


@pytest.fixture(scope="module")
def dense_refs():
    out = []
    for seed, L, k, nev in PROBLEMS:
        w, _ = dense_eigh(MaxwellOperatorNP(_eps(seed), L, k))
        out.append(w)
    return out


def _requests(idx=None):
    from eigenfns_amd.operator import MaxwellOperator
    from eigenfns_amd.batched import Request
    reqs = []
    for i, (seed, L, k, nev) in enumerate(PROBLEMS):
        if idx is not None and i not in idx:
            continue
        reqs.append(Request(MaxwellOperator(_eps(seed), L, k), nev, seed=100 + i, rid=i))
    return reqs


def test_theta_multi_equals_per_operator():
    import torch
    from eigenfns_amd.operator import MaxwellOperator, theta_multi, precondition_multi
    ops = [MaxwellOperator(_eps(s, 16), 1.0 + 0.1 * s, (0.1 * s, 0, 0)) for s in range(3)]
    ws = ops[0]
    rng = np.random.default_rng(0)
    X = torch.from_numpy((rng.standard_normal((12, 2, 16, 16, 16))
                          + 1j * rng.standard_normal((12, 2, 16, 16, 16))).astype(np.complex64)).cuda()
    Y = torch.zeros_like(X)
    # slot i: rows 4i..4i+3 -> output rows in reverse slot order
    theta_multi(ws, ops, X, [0, 4, 8], Y, [8, 4, 0], 4)
    for i, op in enumerate(ops):
        ref = op.theta(X[4 * i:4 * i + 4])
        assert torch.equal(Y[8 - 4 * i:12 - 4 * i], ref)
    precondition_multi(ws, ops, X, [0, 4, 8], Y, [0, 4, 8], 4)
    for i, op in enumerate(ops):
        assert torch.equal(Y[4 * i:4 * i + 4], op.precondition(X[4 * i:4 * i + 4]))


@pytest.mark.parametrize("ortho", ["cholqr", "svqb"])
def test_batched_mix_matches_dense(dense_refs, ortho):
    from eigenfns_amd.batched import solve_batched
    done = solve_batched(_requests(), 8, m=16, guard=6, nslots=3, tol=1e-6, maxit=300,
                         verify_dead_rows=True, ortho=ortho)
    assert len(done) == len(PROBLEMS)
    for req in done:
        w = dense_refs[req.rid]
        rel = np.abs(req.vals - w[:req.nev]) / w[:req.nev]
        print(f"request {req.rid}: nev {req.nev}, iterations {req.iterations}, max rel {rel.max():.2e}")
        assert rel.max() < 1e-6
        assert (np.diff(req.vals) > -1e-8).all()


def test_batch_invariance_alone_vs_mix():
    from eigenfns_amd.batched import solve_batched
    alone = solve_batched(_requests([2]), 8, m=16, guard=6, nslots=1, max_nev=24, tol=1e-5)
    mixed = solve_batched(_requests(), 8, m=16, guard=6, nslots=4, max_nev=24, tol=1e-5)
    a = alone[0].vals
    b = next(r for r in mixed if r.rid == 2).vals
    assert np.array_equal(a, b), (a - b)


def test_batched_agrees_with_single_problem_solver(dense_refs):
    from eigenfns_amd.batched import solve_batched
    from eigenfns_amd.solver import lobpcg_blocks
    req = _requests([0])[0]
    single, _, _ = lobpcg_blocks(req.op, req.nev, m=16, guard=6, tol=1e-6, verbose=False)
    batched = solve_batched([req], 8, m=16, guard=6, nslots=2, tol=1e-6)[0].vals
    assert np.abs(single - batched).max() < 1e-6 * dense_refs[0][req.nev - 1]


def test_batched_keeps_vectors_orthonormal():
    import torch
    from eigenfns_amd.batched import solve_batched
    done = solve_batched(_requests([1, 3]), 8, m=16, guard=6, nslots=2, tol=1e-6, keep_vectors=True)
    for req in done:
        V = req.vecs.reshape(req.nev, -1).cpu().numpy().astype(np.complex128)
        assert np.abs(V.conj() @ V.T - np.eye(req.nev)).max() < 5e-5
        # residuals with the request's own operator
        HV = req.op.theta(req.vecs).reshape(req.nev, -1).cpu().numpy()
        lam = np.real(np.sum(V.conj() * HV, axis=1))
        r = np.linalg.norm(HV - lam[:, None] * V, axis=1) / np.linalg.norm(HV, axis=1)
        assert r.max() < 1e-4


def test_fused_row_statistics_match_separate_reductions():
    """rowstats / row_axpby_norm (one pass each) against the separate row
    reductions and axpby over the same strided rows."""
    import torch
    from eigenfns_amd.linalg import (row_axpby_norm_strided, row_axpby_strided, rowreduce_strided,
                                     rowstats_strided)
    rng = np.random.default_rng(9)
    m, nslot, n = 5, 3, 2 * 12 ** 3
    S = torch.from_numpy((rng.standard_normal((nslot * 3 * m, n))
                          + 1j * rng.standard_normal((nslot * 3 * m, n))).astype(np.complex64)).cuda()
    H = torch.from_numpy((rng.standard_normal((nslot * 3 * m, n))
                          + 1j * rng.standard_normal((nslot * 3 * m, n))).astype(np.complex64)).cuda()
    M = nslot * m
    st = rowstats_strided(S, H, m, 3 * m, 3 * m, M, n)
    dot = rowreduce_strided(1, S, H, m, 3 * m, 3 * m, M, n)
    assert torch.allclose(st[:, 0], dot.real, rtol=1e-13) and torch.allclose(st[:, 1], dot.imag, rtol=1e-12)
    assert torch.allclose(st[:, 2], rowreduce_strided(0, S, None, m, 3 * m, 3 * m, M, n).real, rtol=1e-13)
    assert torch.allclose(st[:, 3], rowreduce_strided(0, H, None, m, 3 * m, 3 * m, M, n).real, rtol=1e-13)
    lam = st[:, 0] / st[:, 2]
    R1 = torch.empty((M, n), dtype=torch.complex64, device="cuda")
    R2 = torch.empty_like(R1)
    nr = row_axpby_norm_strided(H, 3 * m, torch.ones_like(lam), S, 3 * m, -lam, m, M, n, R1, m)
    row_axpby_strided(H, 3 * m, torch.ones_like(lam), S, 3 * m, -lam, m, M, n, R2, m)
    assert torch.equal(R1, R2)
    assert torch.allclose(nr, rowreduce_strided(0, R2, None, m, m, m, M, n).real, rtol=1e-13)


@pytest.mark.parametrize("small_dense", ["device", "hybrid"])
def test_device_small_eigensolves_match_dense(dense_refs, small_dense):
    """The engine with its SVQB / whitening / Rayleigh–Ritz eigensolves on the
    device (efn_heev_batched) meets the same 1e-6 bar against the FP64 dense
    spectrum, with the dead-row bookkeeping verified against row norms, and is
    batch-invariant too."""
    from eigenfns_amd.batched import solve_batched
    done = solve_batched(_requests(), 8, m=16, guard=6, nslots=3, tol=1e-6, maxit=300,
                         verify_dead_rows=True, small_dense=small_dense, ortho="svqb")
    assert len(done) == len(PROBLEMS)
    for req in done:
        w = dense_refs[req.rid]
        assert (np.abs(req.vals - w[:req.nev]) / w[:req.nev]).max() < 1e-6
    alone = solve_batched(_requests([2]), 8, m=16, guard=6, nslots=1, max_nev=24, tol=1e-5,
                          small_dense=small_dense, ortho="svqb")
    mixed = solve_batched(_requests(), 8, m=16, guard=6, nslots=4, max_nev=24, tol=1e-5,
                          small_dense=small_dense, ortho="svqb")
    assert np.array_equal(alone[0].vals, next(r for r in mixed if r.rid == 2).vals)


@pytest.mark.parametrize("small_dense", ["host", "device"])
def test_cholqr_engine_matches_dense(dense_refs, small_dense):
    """The engine with Cholesky-QR orthonormalisation (efn_cholqr_batched for
    the SVQB steps and the P whitening): the same 1e-6 bar against the FP64
    dense spectrum, its keep masks (rows below each problem's rank) verified
    against measured row norms, and batch invariance."""
    from eigenfns_amd.batched import solve_batched
    done = solve_batched(_requests(), 8, m=16, guard=6, nslots=3, tol=1e-6, maxit=300,
                         verify_dead_rows=True, small_dense=small_dense, ortho="cholqr")
    assert len(done) == len(PROBLEMS)
    for req in done:
        w = dense_refs[req.rid]
        assert (np.abs(req.vals - w[:req.nev]) / w[:req.nev]).max() < 1e-6
    alone = solve_batched(_requests([2]), 8, m=16, guard=6, nslots=1, max_nev=24, tol=1e-5,
                          small_dense=small_dense, ortho="cholqr")
    mixed = solve_batched(_requests(), 8, m=16, guard=6, nslots=4, max_nev=24, tol=1e-5,
                          small_dense=small_dense, ortho="cholqr")
    assert np.array_equal(alone[0].vals, next(r for r in mixed if r.rid == 2).vals)
