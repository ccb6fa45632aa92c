"""The batch autotuner's search (host logic, no GPU)."""
from __future__ import annotations

import itertools

from acoustic_agent_amd.autotune import BATCH_SPACE, MULTI_SOURCE_AXES, coordinate_descent


def _grid(space):
    keys = list(space)
    return [dict(zip(keys, v)) for v in itertools.product(*(space[k] for k in keys))]


def test_descent_finds_the_minimum_of_a_separable_cost_with_few_evaluations():
    space = {"a": (0, 1, 2, 3), "b": (5, 6, 7), "c": (0, 1)}
    calls = []

    def cost(p):
        calls.append(dict(p))
        return (p["a"] - 2) ** 2 + 3 * (p["b"] - 7) ** 2 + (0.5 if p["c"] else 0.0)

    best, t = coordinate_descent(space, cost)
    assert best == {"a": 2, "b": 7, "c": 0} and t == 0.0
    keys = [tuple(c.values()) for c in calls]
    assert len(keys) == len(set(keys))  # every point costed once
    assert len(calls) < len(_grid(space))


def test_descent_starts_at_the_first_values_and_keeps_them_when_nothing_is_faster():
    space = {"a": (4, 1, 2), "b": (9, 8)}
    calls = []

    def cost(p):
        calls.append(dict(p))
        return 1.0

    best, _ = coordinate_descent(space, cost)
    assert calls[0] == {"a": 4, "b": 9} and best == {"a": 4, "b": 9}
    assert len(calls) == 1 + 2 + 1  # one sweep: the start, then each other value of each axis


def test_descent_follows_a_coupled_cost_over_several_sweeps():
    # the best b depends on a: (0,0) -> a=1 -> b=1 in the first sweep, then
    # a=2 -> b=2 in the second, and the third finds nothing faster
    space = {"a": (0, 1, 2), "b": (0, 1, 2)}
    calls = []

    def cost(p):
        calls.append(dict(p))
        return (p["a"] - 2) ** 2 + 0.5 * (p["b"] - p["a"]) ** 2

    best, t = coordinate_descent(space, cost)
    assert best == {"a": 2, "b": 2} and t == 0.0
    assert coordinate_descent(space, cost, max_sweeps=1)[0] == {"a": 1, "b": 1}


def test_batch_space_covers_the_tuning_fields():
    from acoustic_agent_amd.batch import BatchTuning

    fields = set(BatchTuning.__dataclass_fields__)
    assert set(BATCH_SPACE) <= fields
    assert {"block_size", "packet_rays", "workgroups_per_cu", "sources_per_lane"} <= set(BATCH_SPACE)
    assert set(MULTI_SOURCE_AXES) <= set(BATCH_SPACE)
    for k, values in BATCH_SPACE.items():
        if k == "packet_rays":
            # held at the guided sizes (0): a sample cannot show the queue
            # drains they remove from the full batch
            assert tuple(values) == (0,), k
            continue
        assert len(values) == len(set(values)) >= 2, k
    assert BATCH_SPACE["block_size"][0] == BatchTuning().block_size
    assert BATCH_SPACE["packet_rays"][0] == BatchTuning().packet_rays
