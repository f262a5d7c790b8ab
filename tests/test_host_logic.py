"""Host-side bookkeeping of the batched engine (no GPU needed)."""
from eigenfns_amd.batched import _runs


def test_slot_runs_cover_each_slot_once_in_order():
    assert _runs([]) == []
    assert _runs([5]) == [(5, 1)]
    assert _runs([3, 1, 2, 7, 8, 0]) == [(0, 4), (7, 2)]
    idx = [0, 2, 3, 4, 9, 10, 12]
    runs = _runs(idx)
    assert runs == [(0, 1), (2, 3), (9, 2), (12, 1)]
    covered = [s for s0, c in runs for s in range(s0, s0 + c)]
    assert covered == sorted(idx)
