"""Ragged multi-problem launches (one launch over problems of different block
widths) and the batched engine with per-request widths.

Bars: ragged Θ / preconditioner bitwise equal to the per-operator calls;
ragged Gram and combine within 1e-12 / 1e-6 of per-problem products (zero
outside each problem's own block); a mixed-width engine's eigenvalues within
1e-6 of the FP64 dense solve, and bitwise the same for a request alone and
in a mix of widths."""
import numpy as np
import pytest

from oracle.dense_np import dense_eigh
from oracle.operator_np import MaxwellOperatorNP

pytestmark = pytest.mark.gpu


def _eps(seed, G=8):
    rng = np.random.default_rng(seed)
    return np.where(rng.random((G, G, G)) < 0.3, 8.41, 1.0).astype(np.float32)


def _block(rows, G, seed):
    import torch
    rng = np.random.default_rng(seed)
    shape = (rows, 2, G, G, G)
    return torch.from_numpy((rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
                            .astype(np.complex64)).cuda()


def test_theta_and_precondition_ragged_equal_per_operator():
    import torch
    from eigenfns_amd.operator import MaxwellOperator, precondition_ragged, theta_ragged
    G = 16
    ops = [MaxwellOperator(_eps(s, G), 1.0 + 0.1 * s, (0.1 * s, 0.05, 0)) for s in range(4)]
    rows = [3, 0, 5, 1]
    X = _block(sum(rows), G, 0)
    Y = torch.zeros_like(X)
    off = np.concatenate([[0], np.cumsum(rows)[:-1]]).tolist()
    out = [sum(rows) - o - r for o, r in zip(off, rows)]   # reversed placement
    theta_ragged(ops[0], ops, X, off, Y, out, rows)
    for op, o, q, r in zip(ops, off, out, rows):
        if r:
            assert torch.equal(Y[q:q + r], op.theta(X[o:o + r]))
    precondition_ragged(ops[0], ops, X, off, Y, off, rows)
    for op, o, r in zip(ops, off, rows):
        if r:
            assert torch.equal(Y[o:o + r], op.precondition(X[o:o + r]))


def test_gram_and_combine_ragged_match_per_problem():
    import torch
    from eigenfns_amd.linalg import combine_ragged, gram_ragged
    G = 8
    ma, mb = [5, 17, 33, 1], [9, 3, 40, 2]
    As = [_block(m, G, 10 + i) for i, m in enumerate(ma)]
    Bs = [_block(m, G, 20 + i) for i, m in enumerate(mb)]
    out = gram_ragged(As, Bs).cpu().numpy()
    for p, (A, B) in enumerate(zip(As, Bs)):
        a = A.reshape(A.shape[0], -1).cpu().numpy().astype(np.complex128)
        b = B.reshape(B.shape[0], -1).cpu().numpy().astype(np.complex128)
        ref = a.conj() @ b.T
        assert np.abs(out[p, :ma[p], :mb[p]] - ref).max() < 1e-12 * np.abs(ref).max()
        pad = out[p].copy()
        pad[:ma[p], :mb[p]] = 0
        assert not pad.any()
    # Hermitian form on square problems
    H = gram_ragged(As, As, hermitian=True).cpu().numpy()
    for p, A in enumerate(As):
        a = A.reshape(A.shape[0], -1).cpu().numpy().astype(np.complex128)
        ref = a.conj() @ a.T
        assert np.abs(H[p, :ma[p], :ma[p]] - ref).max() < 1e-12 * np.abs(ref).max()
    # combine: out_p = C_pᵀ·S_p with each problem's own m_in × m_out, split outputs
    m_in, m_out, split = ma, [7, 40, 12, 2], [3, 40, 0, 1]
    rng = np.random.default_rng(3)
    Cf = (rng.standard_normal((4, max(m_in), max(m_out)))
          + 1j * rng.standard_normal((4, max(m_in), max(m_out)))).astype(np.complex64)
    O1 = [_block(max(sp, 1), G, 30 + i) * 0 for i, sp in enumerate(split)]
    O2 = [_block(max(mo - sp, 1), G, 40 + i) * 0 for i, (mo, sp) in enumerate(zip(m_out, split))]
    combine_ragged(torch.from_numpy(Cf).cuda(), As, m_in, m_out, O1, O2, split)
    for p, A in enumerate(As):
        a = A.reshape(A.shape[0], -1).cpu().numpy().astype(np.complex128)
        ref = Cf[p, :m_in[p], :m_out[p]].astype(np.complex128).T @ a
        got = np.concatenate([O1[p].reshape(O1[p].shape[0], -1).cpu().numpy()[:split[p]],
                              O2[p].reshape(O2[p].shape[0], -1).cpu().numpy()[:m_out[p] - split[p]]])
        assert np.abs(got - ref).max() < 1e-5 * np.abs(ref).max()


PROBLEMS = [  # (eps seed, box, k, nev, width, guard)
    (1, 1.0, (0.0, 0.0, 0.0), 12, 16, 4),
    (2, 1.3, (0.25, 0.0, 0.0), 4, 8, 4),
    (3, 0.9, (0.1, 0.2, 0.3), 20, 24, 4),
    (4, 1.1, (0.5, 0.5, 0.0), 9, 12, 3),
]


def _requests(idx=None):
    from eigenfns_amd.batched import Request
    from eigenfns_amd.operator import MaxwellOperator
    return [Request(MaxwellOperator(_eps(seed), L, k), nev, seed=50 + i, rid=i, m=w, guard=g)
            for i, (seed, L, k, nev, w, g) in enumerate(PROBLEMS) if idx is None or i in idx]


@pytest.mark.parametrize("small_dense,engine_m", [("host", 24), ("device", 24), ("device", 32)])
def test_mixed_width_engine_matches_dense(small_dense, engine_m):
    """engine_m 32: every request narrower than the engine, so the Rayleigh–Ritz
    stacks are padded to three times the widest active request, not 3·m."""
    from eigenfns_amd.batched import solve_batched
    done = solve_batched(_requests(), 8, m=engine_m, guard=4, nslots=4, tol=1e-6, maxit=300,
                         verify_dead_rows=True, small_dense=small_dense)
    assert len(done) == len(PROBLEMS)
    for req in done:
        seed, L, k, nev, w, g = PROBLEMS[req.rid]
        ref, _ = dense_eigh(MaxwellOperatorNP(_eps(seed), L, k))
        rel = np.abs(req.vals - ref[:nev]) / ref[:nev]
        print(f"request {req.rid}: width {w}, nev {nev}, iterations {req.iterations}, max rel {rel.max():.2e}")
        assert rel.max() < 1e-6


def test_mixed_width_alone_equals_in_mix():
    from eigenfns_amd.batched import solve_batched
    alone = solve_batched(_requests([2]), 8, m=24, guard=4, nslots=1, max_nev=20, tol=1e-5)[0].vals
    mixed = solve_batched(_requests(), 8, m=24, guard=4, nslots=4, max_nev=20, tol=1e-5)
    assert np.array_equal(alone, next(r for r in mixed if r.rid == 2).vals)


def test_mixed_width_request_matches_plain_solver():
    """A request of width w in a wider engine solves as lobpcg_blocks with
    block w and the same guard (same random rows, same steps)."""
    from eigenfns_amd.batched import solve_batched
    from eigenfns_amd.solver import lobpcg_blocks
    req = _requests([0])[0]
    single, _, _ = lobpcg_blocks(req.op, req.nev, m=16, guard=4, tol=1e-6, seed=req.seed, verbose=False)
    batched = solve_batched([req], 8, m=24, guard=4, nslots=2, tol=1e-6)[0].vals
    assert np.abs(single - batched).max() < 1e-6 * single[-1]


def test_selected_row_statistics_equal_strided():
    """efn_rowstats_sel / efn_row_axpby_norm_sel give each selected row the
    strided entry points' sums bit for bit; an entry outside nrow × nslot
    reads and writes nothing and gives zero sums."""
    import torch
    from eigenfns_amd.linalg import (row_axpby_norm_sel, row_axpby_norm_strided, rowstats_sel,
                                     rowstats_strided)
    G, nslot, M = 8, 4, 6
    S3 = 3 * M
    n = 2 * G ** 3
    S = _block(nslot * S3, G, 60)
    HS = _block(nslot * S3, G, 61)
    pairs = [(0, 0), (0, 1), (0, 2), (2, 0), (2, 3), (2, 5), (3, 0), (nslot, 0), (1, M)]
    sel = torch.tensor([(s << 16) | r for s, r in pairs], dtype=torch.int32, device=S.device)
    full = rowstats_strided(S, HS, M, S3, S3, nslot * M, n)
    st = rowstats_sel(S, HS, sel, M, nslot, S3, S3, n)
    for i, (s, r) in enumerate(pairs):
        want = full[s * M + r] if s < nslot and r < M else torch.zeros(4, dtype=torch.float64, device=S.device)
        assert torch.equal(st[i], want)
    lam_full = full[:, 0] / full[:, 2]
    R_full = torch.full((nslot * M, 2, G, G, G), 7.0 + 0j, dtype=torch.complex64, device=S.device)
    R_sel = R_full.clone()
    nf = row_axpby_norm_strided(HS, S3, torch.ones_like(lam_full), S, S3, -lam_full, M, nslot * M, n,
                                R_full, M)
    lam_sel = torch.stack([lam_full[s * M + r] if s < nslot and r < M else lam_full[0] for s, r in pairs])
    ns = row_axpby_norm_sel(HS, S3, torch.ones_like(lam_sel), S, S3, -lam_sel, sel, M, nslot, n, R_sel, M)
    touched = set()
    for i, (s, r) in enumerate(pairs):
        if s < nslot and r < M:
            touched.add(s * M + r)
            assert torch.equal(ns[i], nf[s * M + r])
            assert torch.equal(R_sel[s * M + r], R_full[s * M + r])
        else:
            assert ns[i].item() == 0.0
    for row in range(nslot * M):
        if row not in touched:
            assert bool((R_sel[row] == 7.0).all())
