"""Pins the oracle's block algebra (oracle/linalg_np.py) by identities."""
import numpy as np

from oracle import linalg_np as la


def _cplx(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_gram_is_vdot_entrywise():
    rng = np.random.default_rng(0)
    A, B = _cplx(rng, 3, 2, 2, 2, 2), _cplx(rng, 4, 2, 2, 2, 2)
    Gm = la.gram(A, B)
    for i in range(3):
        for j in range(4):
            assert abs(Gm[i, j] - np.vdot(A[i].ravel(), B[j].ravel())) < 1e-12


def test_combine_is_explicit_sum():
    rng = np.random.default_rng(1)
    S1, S2 = _cplx(rng, 2, 5), _cplx(rng, 3, 5)
    C = _cplx(rng, 5, 4)
    out = la.combine(C, S1, S2)
    S = np.concatenate([S1, S2])
    for i in range(4):
        assert np.allclose(out[i], sum(C[j, i] * S[j] for j in range(5)))


def test_svqb_orthonormal_and_drops_dependent_rows():
    rng = np.random.default_rng(2)
    X = _cplx(rng, 6, 300)
    X[4] = 2 * X[1] - X[0]
    Q = la.ortho_block(X)
    nrm = la.rownorm(Q)
    assert (nrm < 1e-8).sum() == 1 and np.all(nrm[nrm > 0.5] > 0.5)
    live = Q[nrm > 0.5]
    assert np.abs(live.conj() @ live.T - np.eye(5)).max() < 1e-12
    # span preserved: every original row lies in span(Q)
    P = live.conj().T @ live
    assert np.abs(X @ P - X).max() < 1e-10 * np.abs(X).max()


def test_deflate_is_an_idempotent_projector():
    rng = np.random.default_rng(3)
    L = la.ortho_block(_cplx(rng, 4, 200))
    X = _cplx(rng, 3, 200)
    D = la.deflate(X, L, chunk_rows=3)  # chunking over locked rows
    assert np.abs(la.gram(L, D)).max() < 1e-12 * np.abs(X).max()
    assert np.abs(la.deflate(D, L) - D).max() < 1e-12 * np.abs(X).max()
    # the removed part lies in span(L)
    R = X - D
    assert np.abs(la.deflate(R, L)).max() < 1e-12 * np.abs(X).max()


def test_rowdot_and_rownorm():
    rng = np.random.default_rng(4)
    A, B = _cplx(rng, 3, 50), _cplx(rng, 3, 50)
    assert np.allclose(la.rowdot_re(A, B), [np.vdot(a, b).real for a, b in zip(A, B)])
    assert np.allclose(la.rownorm(A), np.linalg.norm(A, axis=1))
