"""aa_batch_submit_static, the batch entry point that allocates nothing,
against aa_batch_submit on the same traces, through the C ABI: two calls in a
row, the second with a smaller batch (the static table keeps the first call's
contents).  Integer outputs must be identical and echograms within relative
L1 1e-5 (only the order of float atomic sums differs)."""
from __future__ import annotations

import numpy as np
import pytest

from acoustic_agent_amd import workloads
from acoustic_agent_amd.batch import BatchEngine

from helpers import rel_l1

pytestmark = pytest.mark.gpu


def _same(a, b):
    for k in ("hit_counts", "contrib_counts", "actual_bounces", "active_count"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("bounce_queries", "shadow_queries", "contributions", "escaped_rays"):
        assert a["counters"][k] == b["counters"][k], k
    assert rel_l1(a["echogram"], b["echogram"]) <= 1e-5
    assert rel_l1(a["surface_energy"], b["surface_energy"]) <= 1e-5


def test_static_batch_equals_batch_submit_on_two_calls():
    big = workloads.mic_array_batch(num_scenes=12, seed=3, num_bounces=80, rays=(3000, 12000))
    moving = workloads.keyframe_batch(num_scenes=1, seed=4, keyframes=16, num_rays=4000, num_bounces=60)
    small = workloads.mic_array_batch(num_scenes=2, seed=5, num_bounces=40, rays=(1000, 3000), num_mics=2)
    engine = BatchEngine(0, max_traces=64)
    for reqs in (big + moving, small):
        assert len(reqs) <= 64
        ref = {r["tag"]: r for r in engine.submit(reqs).wait()}
        got = {r["tag"]: r for r in engine.submit_static(reqs).wait()}
        assert set(ref) == set(got) and len(got) == len(reqs)
        for tag, r in ref.items():
            _same(got[tag], r)
    engine.close()
