"""GPU tests of the three Gram engines (vector-ALU FP64, FP64 MFMA, FP32 MFMA
with FP64 flushes) against the complex128 product of the same FP32 inputs.

Bars: the FP64 engines within 1e-12 of max|AᴴB| (products exact, FP64 sums);
the FP32 engine within 2e-6 of max_ij Σ_k |A_ik||B_jk| (FP32 partial sums of at
most 64 products); every engine bitwise reproducible and bitwise identical
alone and in a multi-problem launch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ENGINES = ["valu_f64", "mfma_f64", "mfma_f32"]
SHAPES = [(5, 7, 1000), (32, 32, 70000), (96, 96, 8192), (33, 65, 2 * 16 ** 3), (3, 130, 4098),
          (96, 96, 2 * 48 ** 3)]


def _cblock(rng, m, n):
    return (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(np.complex64)


@pytest.fixture
def engine_guard():
    from eigenfns_amd.linalg import gram_engine
    prev = gram_engine()
    yield gram_engine
    gram_engine(prev)


@pytest.mark.parametrize("engine", ENGINES)
def test_gram_engine_matches_complex128(engine, engine_guard):
    import torch
    from eigenfns_amd.linalg import gram
    engine_guard(engine)
    rng = np.random.default_rng(7)
    for ma, mb, n in SHAPES:
        A, B = _cblock(rng, ma, n), _cblock(rng, mb, n)
        g = gram(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()).cpu().numpy()
        A2, B2 = A.astype(np.complex128), B.astype(np.complex128)
        gr = A2.conj() @ B2.T
        err = np.abs(g - gr).max()
        if engine == "mfma_f32":
            scale = (np.abs(A2) @ np.abs(B2).T).max()
            assert err < 2e-6 * scale, (ma, mb, n, err / scale)
        else:
            assert err < 1e-12 * np.abs(gr).max(), (ma, mb, n, err)
        print(f"{engine} {ma}x{mb} n={n}: max err {err:.3e}")


def test_gram_engines_agree_on_orthonormal_rows(engine_guard):
    """Near-orthonormal rows (the solver's case): every engine's identity defect
    is the FP32 rounding of the rows themselves, not the engine's."""
    import torch
    from eigenfns_amd.linalg import gram
    rng = np.random.default_rng(3)
    n = 2 * 32 ** 3
    Q, _ = np.linalg.qr(_cblock(rng, n, 48).astype(np.complex128))
    X = torch.from_numpy(Q.T.astype(np.complex64)).cuda()
    Xd = X.cpu().numpy().astype(np.complex128)
    exact = Xd.conj() @ Xd.T
    for engine in ENGINES:
        engine_guard(engine)
        err = np.abs(gram(X, X).cpu().numpy() - exact).max()
        print(f"{engine}: identity defect error {err:.2e}")
        # FP64 engines: summation rounding over n = 65536 terms (numpy's too)
        assert err < (1e-6 if engine == "mfma_f32" else 1e-13), (engine, err)


@pytest.mark.parametrize("engine", ENGINES)
def test_gram_multi_bitwise_equals_single(engine, engine_guard):
    import torch
    from eigenfns_amd.linalg import gram, gram_multi
    engine_guard(engine)
    rng = np.random.default_rng(11)
    n = 2 * 24 ** 3
    As = [torch.from_numpy(_cblock(rng, 40, n)).cuda() for _ in range(3)]
    Bs = [torch.from_numpy(_cblock(rng, 40, n)).cuda() for _ in range(3)]
    multi = gram_multi(As, Bs)
    for p in range(3):
        single = gram(As[p], Bs[p])
        assert torch.equal(multi[p], single)
        assert torch.equal(single, gram(As[p], Bs[p]))


def test_gram_unaligned_rows_use_vector_alu(engine_guard):
    """Rows that start off a 16-byte boundary (odd n) take the vector-ALU path
    under every engine and stay exact."""
    import torch
    from eigenfns_amd.linalg import gram
    rng = np.random.default_rng(5)
    n = 4097
    A, B = _cblock(rng, 9, n), _cblock(rng, 6, n)
    gr = A.astype(np.complex128).conj() @ B.astype(np.complex128).T
    for engine in ENGINES:
        engine_guard(engine)
        g = gram(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()).cpu().numpy()
        assert np.abs(g - gr).max() < 1e-12 * np.abs(gr).max()


@pytest.mark.parametrize("engine", ENGINES)
def test_gram_hermitian_upper_tiles(engine, engine_guard):
    """The Hermitian form is exactly Hermitian: its strict upper triangle equals
    the full product bitwise, the lower triangle is its conjugate transpose and
    the diagonal is the full product's real part."""
    import torch
    from eigenfns_amd.linalg import gram, gram_multi
    engine_guard(engine)
    rng = np.random.default_rng(13)
    for m, n in [(40, 6000), (96, 2 * 16 ** 3), (7, 1000), (130, 4096)]:
        A = torch.from_numpy(_cblock(rng, m, n)).cuda()
        B = torch.from_numpy(_cblock(rng, m, n)).cuda()
        for X, Y in ((A, A), (A, B)):
            h = gram_multi([X], [Y], hermitian=True)[0].cpu().numpy()
            full = gram(X, Y).cpu().numpy()
            iu = np.triu_indices(m, 1)
            assert np.array_equal(h[iu], full[iu])
            assert np.array_equal(h, h.conj().T)
            assert np.array_equal(np.diag(h), np.diag(full).real.astype(h.dtype))
        hA = gram_multi([A], [A], hermitian=True)[0].cpu().numpy()
        fA = gram(A, A).cpu().numpy()
        tol = 2e-6 if engine == "mfma_f32" else 1e-12   # bars of the full-product test
        assert np.abs(hA - fA).max() < tol * np.abs(fA).max()


@pytest.mark.parametrize("engine", ["valu", "mfma"])
def test_combine_engines_match_complex128(engine):
    """out = alpha·X + Cᵀ·S on both engines: ragged sizes (m_in not a multiple
    of 4, m_out across the 8/16/32 tile edges, n not a multiple of 256), X
    aliasing out, and the multi-problem form; FP32 accumulation bar 2e-5."""
    import torch
    from eigenfns_amd.linalg import combine_engine, combine_into, combine_multi
    prev = combine_engine(engine)
    try:
        rng = np.random.default_rng(17)
        for m_in, m_out, n in [(3, 5, 700), (96, 12, 5000), (97, 32, 8193), (40, 40, 4096),
                               (96, 96, 2 * 16 ** 3), (72, 48, 2 * 16 ** 3 + 6), (72, 47, 1030),
                               (64, 64, 4099), (90, 61, 2 * 12 ** 3)]:
            S = _cblock(rng, m_in, n)
            Cm = (rng.standard_normal((m_in, m_out)) + 1j * rng.standard_normal((m_in, m_out))).astype(np.complex64)
            X = _cblock(rng, m_out, n)
            ref = 0.5 * X.astype(np.complex128) + Cm.astype(np.complex128).T @ S.astype(np.complex128)
            scale = np.abs(ref).max()
            Sd, Cd = torch.from_numpy(S).cuda(), torch.from_numpy(Cm).cuda()
            out = torch.from_numpy(X).cuda()
            combine_into(Cd, Sd, out, X=out, alpha=0.5)          # X aliases out
            assert np.abs(out.cpu().numpy() - ref).max() < 2e-5 * scale, (m_in, m_out, n)
            outs = [torch.empty((m_out, n), dtype=torch.complex64, device="cuda") for _ in range(2)]
            combine_multi(torch.stack([Cd, Cd]), [Sd, Sd], outs)
            plain = Cm.astype(np.complex128).T @ S.astype(np.complex128)
            for o in outs:
                assert np.abs(o.cpu().numpy() - plain).max() < 2e-5 * np.abs(plain).max()
            assert torch.equal(outs[0], outs[1])
    finally:
        combine_engine(prev)


@pytest.mark.parametrize("engine", ["valu", "mfma"])
def test_combine_in_place_and_unaligned_rows(engine):
    """out may be S itself for m_out <= 64 (one row tile: every column is read
    in full before it is written) — same values as into a fresh buffer; larger
    m_out in place is refused.  Rows starting 8 bytes off a 16-byte boundary
    take the per-element path with the same results."""
    import torch
    from eigenfns_amd.linalg import combine_engine, combine_into
    prev = combine_engine(engine)
    try:
        rng = np.random.default_rng(23)
        for m, n in [(24, 2 * 16 ** 3), (32, 4098), (13, 777), (32, 2 * 8 ** 3 + 1), (40, 3000),
                     (64, 2 * 12 ** 3)]:
            S = _cblock(rng, m, n)
            Cm = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))).astype(np.complex64)
            Sd, Cd = torch.from_numpy(S).cuda(), torch.from_numpy(Cm).cuda()
            fresh = torch.empty_like(Sd)
            combine_into(Cd, Sd, fresh)
            combine_into(Cd, Sd, Sd)
            assert torch.equal(Sd, fresh), (m, n)
            ref = Cm.astype(np.complex128).T @ S.astype(np.complex128)
            assert np.abs(fresh.cpu().numpy() - ref).max() < 2e-5 * np.abs(ref).max()
            # the same rows at an 8-byte offset (no 16-byte access possible)
            flat = torch.zeros(m * n + 1, dtype=torch.complex64, device="cuda")
            Su = flat[1:].view(m, n)
            Su.copy_(torch.from_numpy(S).cuda())
            ou = torch.zeros(m * n + 1, dtype=torch.complex64, device="cuda")[1:].view(m, n)
            combine_into(Cd, Su, ou)
            assert torch.equal(ou, fresh), (m, n)
        S = torch.from_numpy(_cblock(rng, 72, 1000)).cuda()
        C = torch.zeros((72, 72), dtype=torch.complex64, device="cuda")
        with pytest.raises(ValueError):
            combine_into(C, S, S)
    finally:
        combine_engine(prev)


@pytest.mark.parametrize("engine", ["valu", "mfma"])
def test_combine_split_into_own_rows(engine):
    """efn_combine_multi_split: rows < split to one block, the rest to another;
    both blocks may be rows of S itself (the Rayleigh–Ritz update writes X and
    P into S's X and P rows) — bitwise the same as separate combines into
    fresh buffers; up to 64 outputs in place, more refused."""
    import torch
    from eigenfns_amd.linalg import combine_engine, combine_multi, combine_multi_split
    prev = combine_engine(engine)
    try:
        rng = np.random.default_rng(29)
        for m, n in [(24, 2 * 16 ** 3), (32, 4098), (7, 1001), (16, 2 * 12 ** 3)]:
            S0 = torch.from_numpy(_cblock(rng, 3 * m, n)).cuda()
            C = torch.from_numpy((rng.standard_normal((2, 3 * m, 2 * m))
                                  + 1j * rng.standard_normal((2, 3 * m, 2 * m))).astype(np.complex64)).cuda()
            ref = [torch.empty((2 * m, n), dtype=torch.complex64, device="cuda") for _ in range(2)]
            Ss = [S0.clone(), S0.clone()]
            combine_multi(C, [S0, S0], ref)
            combine_multi_split(C, Ss, [S[:m] for S in Ss], [S[2 * m:] for S in Ss], m)
            for S, r in zip(Ss, ref):
                assert torch.equal(S[:m], r[:m]), (m, n)
                assert torch.equal(S[2 * m:], r[m:]), (m, n)
                assert torch.equal(S[m:2 * m], S0[m:2 * m])       # W rows untouched
            plain = C[0].cpu().numpy().astype(np.complex128).T @ S0.cpu().numpy().astype(np.complex128)
            assert np.abs(ref[0].cpu().numpy() - plain).max() < 2e-5 * np.abs(plain).max()
        S = torch.from_numpy(_cblock(rng, 96, 1000)).cuda()
        C = torch.zeros((1, 96, 96), dtype=torch.complex64, device="cuda")
        with pytest.raises(ValueError):
            combine_multi_split(C, [S], [S[:48]], [S], 48)
        with pytest.raises(ValueError):       # an output starting inside S, not at S
            combine_multi_split(C, [S], [S[:48]], [S[48:]], 48)
    finally:
        combine_engine(prev)
