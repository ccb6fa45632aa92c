"""GPU: the serving loop and the autotuner.

Each served result must be the same alone (one persistent single-trace launch
per trace) and in the continuous-batching mix: counts exactly, histograms to
rel L1 <= 1e-5.  The HBM budget charges resident scenes and output arenas.
"""
from __future__ import annotations

import numpy as np
import pytest

from acoustic_agent_amd import workloads
from acoustic_agent_amd.autotune import BATCH_SPACE, Autotuner
from acoustic_agent_amd.batch import trace_sequential
from acoustic_agent_amd.serving import SceneCache, ServingLoop, _host_scenes, request_bytes, serve_sequential

from helpers import rel_l1

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stream():
    return workloads.served_stream(rate_hz=150.0, duration_s=0.25, seed=11, num_scenes=4, max_rays=20_000)


def test_served_results_equal_alone(stream):
    loop = ServingLoop(0, slots=2, max_batch_traces=32)
    done, summary = loop.run(stream, realtime=True)
    loop.close()
    assert len(done) == len(stream)
    assert summary["latency_all"]["count"] == len(stream)
    assert summary["latency_all"]["p99_ms"] >= summary["latency_all"]["p50_ms"] > 0
    for comp in done:
        alone, _ = trace_sequential(comp.request.traces)
        assert len(alone) == len(comp.results)
        for a, m in zip(alone, comp.results):
            assert rel_l1(m["echogram"], a["echogram"]) <= 1e-5
            np.testing.assert_array_equal(m["hit_counts"], a["hit_counts"])
            np.testing.assert_array_equal(m["active_count"], a["active_count"])
            assert m["counters"]["bounce_queries"] == a["counters"]["bounce_queries"]
        assert comp.rt60 is not None and all(v >= 0 for v in comp.rt60.values())


def test_hbm_budget_limits_admission(stream):
    # the largest request with its own scenes (upper bound) is the smallest budget that serves the stream:
    # scenes are evicted between batches to make room, and results do not change
    first = stream[:6]
    alone = max(request_bytes(r) + SceneCache().pending_bytes(_host_scenes(r)) for r in first)
    loop = ServingLoop(0, slots=2, max_batch_traces=32, hbm_budget_bytes=alone)
    done, summary = loop.run(first, realtime=False)
    loop.close()
    assert len(done) == 6 and summary["peak_charged_bytes"] <= alone
    distinct = len({h.content_hash for r in first for h in _host_scenes(r)})
    assert summary["scene_cache"]["misses"] >= distinct
    roomy = ServingLoop(0, slots=2, max_batch_traces=32)
    done2, summary2 = roomy.run(first, realtime=False)
    roomy.close()
    assert summary2["scene_cache"]["misses"] == distinct and summary2["scene_cache"]["evictions"] == 0
    by_arrival = {c.request.arrival_s: c for c in done2}
    for c in done:
        for a, b in zip(c.results, by_arrival[c.request.arrival_s].results):
            np.testing.assert_array_equal(a["hit_counts"], b["hit_counts"])
            assert rel_l1(a["echogram"], b["echogram"]) <= 1e-5
    loop = ServingLoop(0, slots=1, max_batch_traces=32, hbm_budget_bytes=alone - 1)
    with pytest.raises(ValueError):
        loop.run(first, realtime=False)
    loop.close()


def test_scene_bound_covers_the_uploaded_scene(stream):
    for r in stream[:8]:
        for h in _host_scenes(r):
            ds = h.upload(0)
            assert ds.device_bytes <= h.bound_device_bytes and ds.info["content_hash"] == h.content_hash
            ds.close()


def test_sequential_baseline_reports_latency(stream):
    done, summary = serve_sequential(stream[:4], realtime=False)
    assert len(done) == 4 and summary["traces"] == sum(len(r.traces) for r in stream[:4])


def test_autotuner_picks_from_space_and_caches(stream, tmp_path):
    from acoustic_agent_amd.serving import _resolve

    cache = SceneCache(0)
    traces = [t for r in stream[:6] for t in r.traces]
    reqs = _resolve(traces, cache.acquire(t.scene for t in traces))
    path = tmp_path / "tune.json"
    tuner = Autotuner(0, cache_path=path)
    space = {"packet_rays": (0, 1024, 4096), "block_size": (256, 128), "workgroups_per_cu": (0,)}
    best = tuner.tune_batch(reqs, space=space, sample=8, reps=1)
    assert best.block_size in space["block_size"] and best.packet_rays in space["packet_rays"]
    assert 1 + 2 + 1 <= len(tuner.log) <= 6  # at least one full sweep, never more than the grid
    again = Autotuner(0, cache_path=path).tune_batch(reqs, space=space, sample=8, reps=1)
    assert again.key() == best.key()
    assert set(BATCH_SPACE) >= set(space)
    cache.close()
