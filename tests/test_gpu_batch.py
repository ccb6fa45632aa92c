"""GPU: the persistent batch against single traces of the same requests.

A trace must give the same result alone (``aa_trace``) and in a mixed batch:
per-surface counts, active rays, bounce counts and query counts exactly;
histograms to rel L1 <= 1e-5 (the order of float atomic sums differs).
"""
from __future__ import annotations

import numpy as np
import pytest

from acoustic_agent_amd import hip_backend, scenes
from acoustic_agent_amd.batch import BatchEngine, BatchTuning, TraceRequest, trace_sequential

from helpers import rel_l1

pytestmark = pytest.mark.gpu


def _scene(w):
    a = w.scene.kernel_arrays()
    return hip_backend.DeviceScene(w.mesh, np.asarray(a["reflection"]), np.asarray(a["scattering"]), device=0)


@pytest.fixture(scope="module")
def requests():
    box = scenes.shoebox_workload(0)
    plan = scenes.l_plan_workload(0, target_triangles=20000)
    small = scenes.random_scene(5, 3000)
    s_box, s_plan, s_small = _scene(box), _scene(plan), _scene(small)
    rng = np.random.default_rng(3)
    reqs = []
    for i in range(6):
        for w, s in ((box, s_box), (plan, s_plan), (small, s_small)):
            ks = 1 if i % 3 else 3
            srcs = np.stack([w.source + rng.uniform(-0.3, 0.3, 3) * (1, 1, 0.2) for _ in range(ks)])
            reqs.append(TraceRequest(scene=s, listener=w.listener + rng.uniform(-0.05, 0.05, 3), sources=srcs,
                                     num_rays=int(rng.integers(500, 6000)), num_bounces=int(rng.integers(20, 120)),
                                     duration_s=float(rng.choice([0.5, 1.0, 2.0])), seed=int(rng.integers(0, 1 << 40)),
                                     tag=len(reqs)))
    return reqs


def _check_same(a, b):
    assert rel_l1(a["echogram"], b["echogram"]) <= 1e-5
    np.testing.assert_array_equal(a["hit_counts"], b["hit_counts"])
    np.testing.assert_array_equal(a["contrib_counts"], b["contrib_counts"])
    np.testing.assert_array_equal(a["active_count"], b["active_count"])
    np.testing.assert_array_equal(a["actual_bounces"], b["actual_bounces"])
    for key in ("bounce_queries", "shadow_queries", "contributions"):
        assert a["counters"][key] == b["counters"][key], key


def test_batch_equals_single_traces(requests):
    single, _ = trace_sequential(requests)
    engine = BatchEngine(0, max_traces=64)
    got = {r["tag"]: r for r in engine.submit(requests).wait()}
    for s in single:
        _check_same(got[s["tag"]], s)
    engine.close()


@pytest.mark.parametrize("tuning", [
    BatchTuning(block_size=64, packet_rays=256),
    BatchTuning(block_size=128, packet_rays=4096, workgroups_per_cu=2),
    BatchTuning(block_size=512, packet_rays=1024, workgroups_per_cu=1),
])
def test_batch_tunings_agree(requests, tuning):
    engine = BatchEngine(0, max_traces=64)
    base = {r["tag"]: r for r in engine.submit(requests).wait()}
    other = {r["tag"]: r for r in engine.submit(requests, tuning).wait()}
    for tag, r in base.items():
        _check_same(other[tag], r)
    engine.close()


def test_guided_packets_equal_fixed_packets():
    """A batch large enough for guided packets of several sizes: 64 traces of
    400k rays over two scenes (2 of the 8 XCD queues), 25.6 M lane tasks, so
    on 3072 resident waves the first traces of each queue get 4096-ray packets
    (25.6 M / (2 x 3072) = 4167 > 4096) and the last ones 1024.  Every trace's
    result equals the fixed 1024-ray packets' (order of float sums aside)."""
    box = scenes.shoebox_workload(0)
    plan = scenes.l_plan_workload(1, target_triangles=2000)
    s_box, s_plan = _scene(box), _scene(plan)
    rng = np.random.default_rng(11)
    reqs = []
    for i in range(64):
        w, s = (box, s_box) if i % 2 else (plan, s_plan)
        reqs.append(TraceRequest(scene=s, listener=w.listener + rng.uniform(-0.05, 0.05, 3), sources=w.source[None, :],
                                 num_rays=400_000, num_bounces=3, duration_s=0.5,
                                 seed=int(rng.integers(0, 1 << 40)), tag=i))
    engine = BatchEngine(0, max_traces=64)
    guided = {r["tag"]: r for r in engine.submit(reqs, BatchTuning(packet_rays=0)).wait()}
    fixed = {r["tag"]: r for r in engine.submit(reqs, BatchTuning(packet_rays=1024)).wait()}
    engine.close()
    assert sorted(guided) == list(range(64))
    for tag, r in fixed.items():
        _check_same(guided[tag], r)
        assert r["counters"]["bounce_queries"] > 400_000


def test_graph_replays_follow_new_contents(requests):
    import torch

    engine = BatchEngine(0, max_traces=64)
    tuning = BatchTuning(use_graph=True)
    single = {r["tag"]: r for r in trace_sequential(requests)[0]}
    subsets = [requests, requests[:7], requests[5:], requests]
    for rep, subset in enumerate(subsets):
        res = engine.submit(subset, tuning).wait()
        assert len(res) == len(subset)
        for r in res:
            _check_same(r, single[r["tag"]])
        # eager work between replays
        x = torch.randn(1 << 16, device="cuda:0")
        torch.cuda.synchronize()
        assert torch.isfinite(x.sum()).item()
    engine.close()


def test_batch_rejects_oversize_and_empty(requests):
    engine = BatchEngine(0, max_traces=4)
    with pytest.raises(ValueError):
        engine.submit([])
    with pytest.raises(ValueError):
        engine.submit(requests[:5])
    engine.close()


def test_batch_ambisonics_equal_single_traces():
    # FOA echograms through the batch path (per-trace render_ambisonics)
    box = scenes.shoebox_workload(1)
    s_box = _scene(box)
    reqs = [TraceRequest(scene=s_box, listener=box.listener, sources=box.source[None, :], num_rays=3000,
                         num_bounces=40, duration_s=1.0, seed=77 + i, render_ambisonics=bool(i % 2), tag=i)
            for i in range(4)]
    single = {r["tag"]: r for r in trace_sequential(reqs)[0]}
    engine = BatchEngine(0, max_traces=8)
    got = {r["tag"]: r for r in engine.submit(reqs).wait()}
    engine.close()
    for i in range(4):
        _check_same(got[i], single[i])
        if i % 2:
            assert got[i]["ambisonic"].shape == (1, 6, 4, 1000)
            assert rel_l1(got[i]["ambisonic"], single[i]["ambisonic"]) <= 1e-5


def test_planned_motion_equals_frame_by_frame_traces():
    # the motion planner's multi-source trace equals one trace per frame
    from acoustic_agent_amd.mic import microphone_array
    from acoustic_agent_amd.models import SimConfig
    from acoustic_agent_amd.plan import motion_traces

    w = scenes.l_plan_workload(0, target_triangles=20000)
    ds = _scene(w)
    cfg = SimConfig(rt_num_rays=3000, rt_num_bounces=60, rt_duration_s=1.0, duration_s=1.0, rt_bin_duration_s=1e-3,
                    seed=21)
    path = scenes.moving_source_path(4, keyframes=5)
    frames = [{"source": p, "receiver": w.listener} for p in path]
    plan = motion_traces(ds, frames, cfg, microphone_array("linear", count=2, spacing_m=0.1))
    assert len(plan.traces) == 2
    engine = BatchEngine(0, max_traces=8)
    got = engine.submit(plan.traces).wait()
    engine.close()
    for f, frame in enumerate(frames):
        for t, k in plan.frames[f]:
            tr = plan.traces[t]
            single = TraceRequest(scene=ds, listener=tr.listener, sources=np.asarray(frame["source"])[None, :],
                                  num_rays=tr.num_rays, num_bounces=tr.num_bounces, duration_s=tr.duration_s,
                                  bin_dur=tr.bin_dur, seed=tr.seed)
            (one,), _ = trace_sequential([single])
            assert rel_l1(got[t]["echogram"][k], one["echogram"][0]) <= 1e-5
            np.testing.assert_array_equal(got[t]["hit_counts"][k], one["hit_counts"][0])
            assert int(got[t]["active_count"][k]) == int(one["active_count"][0])


def test_stream_query_polls_a_submission(requests):
    # aa_stream_query / aa_stream_sync: the poll and wait of a submitted batch
    import ctypes

    import torch

    from acoustic_agent_amd import _lib

    lib = _lib.load()
    engine = BatchEngine(0, max_traces=64)
    s = torch.cuda.Stream(0)
    res = engine.submit(requests[:6], stream=s)
    done = ctypes.c_int32(-1)
    _lib.check(lib.aa_stream_query(ctypes.c_void_p(s.cuda_stream), ctypes.byref(done)))
    assert done.value in (0, 1)
    _lib.check(lib.aa_stream_sync(ctypes.c_void_p(s.cuda_stream)))
    _lib.check(lib.aa_stream_query(ctypes.c_void_p(s.cuda_stream), ctypes.byref(done)))
    assert done.value == 1
    assert len(res.wait()) == 6
    engine.close()
