"""Host-side bookkeeping of the batched engine (no GPU needed)."""


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
