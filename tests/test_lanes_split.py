"""Lane assignment (host logic, no GPU)."""
from nonnewton_amd.lanes import split_by_cost


def test_split_by_cost_is_longest_first_to_least_loaded():
    costs = [10, 1, 7, 3, 3, 8, 2]
    lanes = split_by_cost(costs, 3)
    assert sorted(k for l in lanes for k in l) == list(range(len(costs)))
    # by hand: 10->L0, 8->L1, 7->L2, 3->L2 (10), 3->L1 (11), 2->L0 (12), 1->L2 (11)
    assert lanes == [[0, 6], [5, 4], [2, 3, 1]]
    loads = [sum(costs[k] for k in l) for l in lanes]
    assert max(loads) - min(loads) <= max(costs)


def test_split_by_cost_single_lane_keeps_everything():
    assert split_by_cost([3, 2, 1], 1) == [[0, 1, 2]]
