"""Tape planning of the checkpointed adjoint (samoa_amd/adjoint.py); host only."""
import math

import pytest

from samoa_amd.adjoint import GaugeInversion

plan = GaugeInversion.plan_tape


def test_store_everything_when_it_fits():
    n, N = 2 ** 21, 200
    assert plan(n, N, "auto", hbm_budget=N * 25 * n) == (True, 1)


def test_sqrt_checkpoints_when_tight():
    n, N = 2 ** 21, 200
    store_all, K = plan(n, N, "auto", hbm_budget=60 * 25 * n)
    assert not store_all and K == math.ceil(math.sqrt(N)) == 15


def test_budget_too_small_is_an_error():
    with pytest.raises(MemoryError):
        plan(2 ** 21, 200, "auto", hbm_budget=10 * 25 * 2 ** 21)


def test_explicit_choices():
    assert plan(100, 50, "all") == (True, 1)
    assert plan(100, 50, 7) == (False, 7)
    with pytest.raises(ValueError):
        plan(100, 50, 0)


# ---- per-instance checkpoint pools (plan_pools) ------------------------------

from samoa_amd.adjoint import plan_pools, tape_bytes  # noqa: E402


def test_pools_all_store_when_everything_fits():
    cells, N = [2 ** 17, 2 ** 19, 2 ** 20], 200
    budget = sum(tape_bytes(c, N, True) for c in cells)
    assert plan_pools(cells, N, budget) == [[(True, 1, [0, 1, 2])]]


def test_pools_largest_instances_switch_to_checkpoints_first():
    cells, N = [2 ** 17, 2 ** 21, 2 ** 18, 2 ** 20], 200
    K = math.ceil(math.sqrt(N))
    # room for every checkpointed tape plus the two smallest stored whole
    budget = (sum(tape_bytes(c, N, False, K) for c in cells)
              + sum(tape_bytes(c, N, True) - tape_bytes(c, N, False, K) for c in cells[::2]))
    rounds = plan_pools(cells, N, budget)
    assert rounds == [[(True, 1, [0, 2]), (False, K, [1, 3])]]
    used = sum(tape_bytes(cells[s], N, sa, k) for sa, k, m in rounds[0] for s in m)
    assert used <= budget


def test_pools_admission_rounds_when_the_sum_exceeds_the_budget():
    """The summed checkpointed tapes exceed the budget while each fits: the
    instances are admitted in rounds that each fit."""
    cells, N = [2 ** 20] * 3 + [2 ** 16] * 2, 200
    K = math.ceil(math.sqrt(N))
    one = tape_bytes(2 ** 20, N, False, K)
    budget = 1.5 * one
    rounds = plan_pools(cells, N, budget)
    assert len(rounds) == 3
    seen = sorted(s for r in rounds for _, _, m in r for s in m)
    assert seen == [0, 1, 2, 3, 4]
    for r in rounds:
        used = sum(tape_bytes(cells[s], N, sa, k) for sa, k, m in r for s in m)
        assert used <= budget


def test_pools_instance_beyond_the_budget_is_an_error():
    with pytest.raises(MemoryError):
        plan_pools([2 ** 21, 2 ** 10], 200, 10 * 25 * 2 ** 21)


def test_pools_explicit_modes():
    assert plan_pools([10, 20], 50, 0, "all") == [[(True, 1, [0, 1])]]
    assert plan_pools([10, 20], 50, 0, 7) == [[(False, 7, [0, 1])]]
