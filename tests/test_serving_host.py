"""CPU checks of the serving loop's host logic: the scene cache (content-hash
LRU, pins, byte accounting) with a stand-in upload, and the served stream's
request mix (no device needed: scene hashes are computed on the host)."""
from __future__ import annotations

import numpy as np
import pytest

from acoustic_agent_amd import workloads
from acoustic_agent_amd.serving import SceneCache, _host_scenes, request_bytes


class _Host:
    def __init__(self, key, nbytes):
        self.content_hash = key
        self.bound_device_bytes = nbytes + 10


class _Resident:
    closed = 0

    def __init__(self, host):
        self.device_bytes = host.bound_device_bytes - 10
        self.key = host.content_hash

    def close(self):
        _Resident.closed += 1


def test_scene_cache_is_a_pinned_lru():
    cache = SceneCache(upload=lambda h, dev: _Resident(h))
    a, b, c = _Host(1, 100), _Host(2, 200), _Host(3, 300)
    assert cache.pending_bytes([a, b, a]) == 110 + 210  # one bound per distinct content
    got = cache.acquire([a, b])
    assert set(got) == {1, 2} and cache.bytes == 300 and cache.stats["misses"] == 2
    assert cache.pending_bytes([a, b]) == 0 and cache.pending_bytes([c]) == 310
    assert cache.evictable_bytes() == 0  # both pinned
    assert cache.evict(1000) == 0 and len(cache) == 2
    cache.release([1, 2])
    cache.acquire([a])  # a becomes most recently used
    cache.release([1])
    assert cache.evictable_bytes() == 300 and cache.evictable_bytes(keep={2}) == 100
    # oldest first: b (200) goes before a
    assert cache.evict(150) == 200 and len(cache) == 1 and cache.resident(a) and not cache.resident(b)
    assert cache.stats == {"hits": 1, "misses": 2, "evictions": 1}
    before = _Resident.closed
    cache.close()
    assert _Resident.closed == before + 1 and cache.bytes == 0


def test_served_stream_draws_the_four_workload_kinds():
    stream = workloads.served_stream(50.0, 1.5, seed=50, num_scenes=4, num_keyframe_scenes=1)
    assert 40 <= len(stream) <= 120
    arrivals = [r.arrival_s for r in stream]
    assert arrivals == sorted(arrivals) and arrivals[-1] <= 1.5
    kinds = {r.kind for r in stream}
    assert kinds == {"shoebox", "single", "mic_array", "keyframes"}
    for r in stream:
        tri = r.traces[0].scene.info["num_triangles"]
        if r.kind == "shoebox":
            (t,) = r.traces
            assert tri == 12 and t.num_rays == 10_000 and t.num_bounces == 50 and t.num_bins == 1000
            np.testing.assert_allclose(t.listener, (3.6, 2.7, 1.4))
        elif r.kind == "keyframes":
            (t,) = r.traces
            assert t.num_sources == 64 and t.num_rays == 100_000 and t.num_bins == 1000
            steps = np.linalg.norm(np.diff(np.asarray(t.sources), axis=0), axis=1) * 32.0
            assert 0.5 - 1e-3 <= float(np.median(steps)) <= 1.5 + 1e-3
        else:
            assert 1 <= len(r.traces) <= (8 if r.kind == "mic_array" else 1)
            assert all(10_000 <= t.num_rays <= 200_000 and t.num_bins == 2000 for t in r.traces)
            assert 12 <= tri <= 60_000
        assert request_bytes(r) > 0
    # requests share scenes by content: few distinct scenes, many requests
    keys = {h.content_hash for r in stream for h in _host_scenes(r)}
    assert len(keys) <= 4 + 1 + 1


@pytest.mark.parametrize("seed", [1, 2])
def test_served_stream_is_seeded(seed):
    a = workloads.served_stream(30.0, 0.5, seed=seed, num_scenes=2, num_keyframe_scenes=1)
    b = workloads.served_stream(30.0, 0.5, seed=seed, num_scenes=2, num_keyframe_scenes=1)
    assert [(r.arrival_s, r.kind, len(r.traces)) for r in a] == [(r.arrival_s, r.kind, len(r.traces)) for r in b]
