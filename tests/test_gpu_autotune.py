"""The batch autotuner's descent against an exhaustive sweep of the same
space, on samples of the batched mic-array and moving-source workloads."""
from __future__ import annotations

import pytest

from acoustic_agent_amd import workloads
from acoustic_agent_amd.autotune import BATCH_SPACE, MULTI_SOURCE_AXES, Autotuner

pytestmark = pytest.mark.gpu

SAMPLES = {
    # 4 scenes x 4-mic ring: 16 single-source traces of 20k-200k rays
    "mic_arrays": lambda: workloads.mic_array_batch(num_scenes=4, seed=11, num_bounces=200),
    # 2 furnished plans x 64 keyframes: 2 traces of 64 sources, 50k rays
    "keyframes": lambda: workloads.keyframe_batch(num_scenes=2, seed=12, num_rays=50_000, num_bounces=200),
}


def _key(fields: dict) -> tuple:
    return tuple(sorted(fields.items()))


@pytest.mark.parametrize("kind", sorted(SAMPLES))
def test_descent_choice_is_within_5_percent_of_the_exhaustive_best(kind):
    reqs = SAMPLES[kind]()
    multi = any(r.num_sources > 1 for r in reqs)
    descent = Autotuner(0)
    choice = descent.tune_batch(reqs, sample=len(reqs), reps=3, force=True)
    sweep = Autotuner(0)
    sweep.tune_batch(reqs, sample=len(reqs), reps=3, force=True, exhaustive=True)
    axes = [k for k in BATCH_SPACE if multi or k not in MULTI_SOURCE_AXES]
    grid = 1
    for k in axes:
        grid *= len(BATCH_SPACE[k])
    times = {_key(c): t for _, c, t in sweep.log}
    assert len(times) == grid
    assert len(descent.log) < grid
    chosen = times[_key({k: getattr(choice, k) for k in axes})]
    best = min(times.values())
    print(f"{kind}: descent {len(descent.log)} of {grid} points, chosen {chosen:.2f} ms, "
          f"exhaustive best {best:.2f} ms ({chosen / best - 1:+.1%})")
    assert chosen <= 1.05 * best
