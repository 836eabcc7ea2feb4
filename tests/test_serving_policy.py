"""The serving scheduler's tick selection (shortest remaining work first) on the host."""

from deconlib_amd.serving import pick_shortest


def test_pick_shortest_orders_by_remaining_work_and_fills_the_tick():
    rem = [50.0, 10.0, 30.0, 10.0, 5.0]
    vox = [100, 40, 60, 40, 20]
    # 20 (job 4) + 40 (job 1) + 40 (job 3) = 100 >= 90: three jobs, returned in list order
    assert pick_shortest(rem, vox, 90) == [1, 3, 4]
    # a budget above every job's voxels takes them all
    assert pick_shortest(rem, vox, 10 ** 9) == [0, 1, 2, 3, 4]
    # at least one job, even when the first alone exceeds the budget
    assert pick_shortest([3.0, 1.0], [500, 800], 10) == [1]
    # ties keep list order
    assert pick_shortest([1.0, 1.0, 1.0], [5, 5, 5], 8) == [0, 1]
    assert pick_shortest([], [], 10) == []
