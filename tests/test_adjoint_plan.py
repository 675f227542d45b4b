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
    store_all, K = plan(n, N, "auto", hbm_budget=40 * 25 * n)
    assert not store_all and K == math.ceil(math.sqrt(N)) == 15


def test_budget_too_small_is_an_error():
    with pytest.raises(MemoryError):
        plan(2 ** 21, 200, "auto", hbm_budget=10 * 25 * 2 ** 21)


def test_explicit_choices():
    assert plan(100, 50, "all") == (True, 1)
    assert plan(100, 50, 7) == (False, 7)
    with pytest.raises(ValueError):
        plan(100, 50, 0)
