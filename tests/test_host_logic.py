"""Host-side bookkeeping of the batched engine (no GPU needed)."""
import pytest


def test_plan_slots_budget_and_work_cap():
    from eigenfns_amd.serving import _slot_bytes, plan_slots
    grids = [32, 48, 64, 96, 128]
    plan = plan_slots(grids, 24, 8, 64, 32, 200e9)
    assert sum(_slot_bytes(g, 24, 8, 64) * plan[g] for g in grids) <= 200e9
    assert all(1 <= plan[g] <= 32 for g in grids)
    assert plan[32] >= plan[64] >= plan[128]
    capped = plan_slots(grids, 24, 8, 64, 32, 200e9, engine_voxels=8 * 64 ** 3)
    assert capped[128] == 1 and capped[96] == 2 and capped[64] == 8
    assert all(capped[g] * g ** 3 <= 8 * 64 ** 3 or capped[g] == 1 for g in grids)
    # a cap below one slot still leaves one slot per grid
    assert plan_slots(grids, 24, 8, 64, 32, 200e9, engine_voxels=1)[32] == 1


def test_blas_thread_limit_is_scoped_and_restored():
    """The host eigensolve pool limits BLAS to one thread only while it runs
    (nested and concurrent holders share one limit); afterwards the caller's
    setting is back."""
    import threading
    import numpy as np
    pytest.importorskip("threadpoolctl")
    from threadpoolctl import threadpool_info
    from eigenfns_amd import linalg

    def blas_threads():
        return sorted(i["num_threads"] for i in threadpool_info() if i["user_api"] == "blas")

    before = blas_threads()
    a = np.random.default_rng(0).standard_normal((6, 16, 16)) + 0j
    linalg.eigh_stack_host(a)
    assert blas_threads() == before
    with linalg.one_blas_thread():
        assert all(n == 1 for n in blas_threads())
        ths = [threading.Thread(target=linalg.eigh_stack_host, args=(a,)) for _ in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        assert all(n == 1 for n in blas_threads())
    assert blas_threads() == before


def test_tuning_table_lookup_and_user_override(tmp_path, monkeypatch):
    """The committed tuning table is read for a grid, the user table wins over
    it, and grids in neither table give no entry unless first-use tuning is on."""
    import json
    from eigenfns_amd import autotune
    user = tmp_path / "t.json"
    monkeypatch.setenv("EFN_TUNING_CACHE", str(user))
    monkeypatch.delenv("EFN_AUTOTUNE", raising=False)
    autotune.reload()
    pkg = json.loads(autotune.PACKAGE_TABLE.read_text())[autotune.ARCH] if autotune.PACKAGE_TABLE.exists() else {}
    for G, e in pkg.items():
        assert e["passes"] in (3, 5) and e["chunk"] >= 1
        assert autotune.choice(int(G)) == e
    user.write_text(json.dumps({autotune.ARCH: {"64": {"passes": 5, "chunk": 7}}}))
    autotune.reload()
    assert autotune.choice(64) == {"passes": 5, "chunk": 7}
    assert autotune.choice(4) is None or "4" in pkg
    autotune.reload()
    monkeypatch.delenv("EFN_TUNING_CACHE")
    autotune.reload()


def test_tuning_candidates_respect_the_workspace_cap():
    from eigenfns_amd import autotune
    for G in (32, 64, 128, 192, 256):
        c = autotune.candidates(G)
        assert all(ch * 3 * G ** 3 * 8 <= 1.1e9 for _, ch in c)
        assert (5, autotune.default_chunk(G)) in c
    assert {p for p, _ in autotune.candidates(128)} == {3, 5}
    assert {p for p, _ in autotune.candidates(192)} == {3, 5}   # half-plane three passes
    assert {p for p, _ in autotune.candidates(256)} == {5}


def test_rr_dead_rows_with_masks_narrower_than_the_engine():
    """Dead rows of padded Rayleigh–Ritz stacks when the keep masks are only as
    wide as the widest active problem (a ragged launch's padding), narrower
    than the engine's block: every row matches a per-problem loop, padding
    rows are not dead, and a mask narrower than the widest problem is refused
    (torch's CPU gather checks its indices, so an index past a mask raises)."""
    import numpy as np
    import torch
    from eigenfns_amd.batched import rr_dead_rows
    rng = np.random.default_rng(4)
    widths, M = [3, 5, 2], 8
    wa = max(widths)
    K = 3 * wa
    xn = torch.from_numpy(rng.uniform(0.0, 1.0, (3, M)))
    keep_w = torch.from_numpy(rng.random((3, wa)) < 0.7)
    keepp = torch.from_numpy(rng.random((3, wa)) < 0.7)
    dead = rr_dead_rows(widths, K, xn, keep_w, keepp).numpy()
    for j, w in enumerate(widths):
        ref = np.concatenate([xn[j, :w].numpy() < 0.5, ~keep_w[j, :w].numpy(), ~keepp[j, :w].numpy(),
                              np.zeros(K - 3 * w, dtype=bool)])
        assert np.array_equal(dead[j], ref)
    with pytest.raises(ValueError):
        rr_dead_rows(widths, K, xn, keep_w[:, :wa - 1], keepp)
    with pytest.raises(ValueError):
        rr_dead_rows(widths, K - 1, xn, keep_w, keepp)


def test_engines_per_grid_policy():
    """Two engines for the largest grid by default, one elsewhere; an int for
    every grid; a dict (string keys accepted) for the grids it names."""
    from eigenfns_amd.serving import engines_for
    assert engines_for([48, 32, 64]) == {32: 1, 48: 1, 64: 2}
    assert engines_for([32]) == {32: 2}
    assert engines_for([]) == {}
    assert engines_for([32, 64], 3) == {32: 3, 64: 3}
    assert engines_for([32, 64], 0) == {32: 1, 64: 1}
    assert engines_for([32, 48, 64], {"48": 2}) == {32: 1, 48: 2, 64: 1}
