"""The plain solver with its steady-state iteration captured in a HIP graph
(lobpcg_blocks(graph=True)): every replay runs between host residual
read-backs, and the solve must equal the uncaptured loop (device
eigensolves, same kernels, same order) bitwise — eigenvalues, iteration
counts and locked vectors — and the FP64 dense spectrum within 1e-6."""
import numpy as np
import pytest

from oracle.dense_np import dense_eigh
from oracle.operator_np import MaxwellOperatorNP

pytestmark = pytest.mark.gpu


def _eps(seed, G):
    rng = np.random.default_rng(seed)
    return np.where(rng.random((G, G, G)) < 0.3, 8.41, 1.0).astype(np.float32)


@pytest.mark.parametrize("G,m,guard,nev,k", [(8, 16, 6, 20, (0.0, 0.0, 0.0)),
                                              (8, 24, 8, 30, (0.1, 0.2, 0.3)),
                                              (16, 40, 8, 48, (0.25, 0.0, 0.0))])
def test_graph_replayed_iterations_equal_eager(monkeypatch, G, m, guard, nev, k):
    import torch
    from eigenfns_amd.operator import MaxwellOperator
    from eigenfns_amd.solver import lobpcg_blocks
    monkeypatch.setenv("EFN_SMALL_DENSE", "device")   # the eager loop's eigensolves on the device too
    eps = _eps(G + m, G)
    op = MaxwellOperator(eps, 1.1, k)
    v0, x0, s0 = lobpcg_blocks(op, nev, m=m, guard=guard, tol=1e-6, seed=3, verbose=False)
    v1, x1, s1 = lobpcg_blocks(op, nev, m=m, guard=guard, tol=1e-6, seed=3, verbose=False, graph=True)
    print(f"G={G} m={m}: {s1.iterations} iterations ({len(s1.rounds)} blocks), eager {s0.iterations}")
    assert s0.iterations == s1.iterations
    assert np.array_equal(v0, v1)
    assert torch.equal(x0, x1)
    if G == 8:
        ref, _ = dense_eigh(MaxwellOperatorNP(eps, 1.1, k))
        assert (np.abs(v1 - ref[:nev]) / ref[:nev]).max() < 1e-6


def test_graph_needs_cholesky_qr():
    from eigenfns_amd.operator import MaxwellOperator
    from eigenfns_amd.solver import lobpcg_blocks
    op = MaxwellOperator(_eps(1, 8), 1.0)
    with pytest.raises(ValueError):
        lobpcg_blocks(op, 8, m=16, guard=6, ortho="svqb", graph=True, verbose=False)
