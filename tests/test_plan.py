"""Receiver models and motion planning (host logic, no GPU).

``microphone_array`` / ``channel_positions`` against their closed forms and
the clamping of acoustic_agent/mic.py:11-80; ``channel_traces`` and
``motion_traces`` against the per-channel / per-frame solves they replace
(engine.py:64-73, api.py:669-713).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from acoustic_agent_amd.mic import channel_positions, microphone_array
from acoustic_agent_amd.models import SimConfig
from acoustic_agent_amd.plan import MAX_SOURCES_PER_TRACE, channel_traces, motion_traces


def test_circular_array_positions():
    m = microphone_array("circular", count=4, radius_m=0.05)
    assert m["type"] == "circular_array" and [c["id"] for c in m["channels"]] == ["M1", "M2", "M3", "M4"]
    pos = np.asarray(channel_positions((1.0, 2.0, 1.5), m))
    want = np.array([[1.05, 2.0, 1.5], [1.0, 2.05, 1.5], [0.95, 2.0, 1.5], [1.0, 1.95, 1.5]])
    np.testing.assert_allclose(pos, want, atol=1e-12)
    m = microphone_array("circular_array", count=6, radius_m=0.1, orientation_deg=30.0)
    pos = np.asarray(channel_positions((0.0, 0.0, 0.0), m))
    for i, p in enumerate(pos):
        a = math.radians(30.0) + 2 * math.pi * i / 6
        assert p == pytest.approx((round(0.1 * math.cos(a), 6), round(0.1 * math.sin(a), 6), 0.0))


def test_linear_array_positions_and_clamping():
    m = microphone_array("linear", count=3, spacing_m=0.1, orientation_deg=90.0)
    pos = np.asarray(channel_positions((1.0, 1.0, 1.0), m))
    np.testing.assert_allclose(pos, [[1.0, 0.9, 1.0], [1.0, 1.0, 1.0], [1.0, 1.1, 1.0]], atol=1e-12)
    assert len(microphone_array("linear", count=1)["channels"]) == 2
    assert len(microphone_array("linear", count=500)["channels"]) == 64
    assert len(microphone_array("circular", count=2)["channels"]) == 3
    assert len(microphone_array("circular")["channels"]) == 8
    big = microphone_array("circular", count=3, radius_m=9.0)
    assert max(abs(c["offset"][0]) for c in big["channels"]) == pytest.approx(5.0)


def test_mono_and_hrtf_are_one_position():
    assert channel_positions((1, 2, 3), None) == [(1.0, 2.0, 3.0)]
    assert channel_positions((1, 2, 3), microphone_array("hrtf")) == [(1.0, 2.0, 3.0)]
    assert microphone_array("MONO")["type"] == "mono"
    with pytest.raises(ValueError):
        microphone_array("tetrahedral")


def test_channel_traces_one_per_channel():
    cfg = SimConfig(rt_num_rays=1000, rt_num_bounces=20, duration_s=0.5, rt_duration_s=1.0, seed=7)
    model = microphone_array("circular", count=5, radius_m=0.05)
    reqs = channel_traces("scene", (1, 1, 1), (2.0, 2.0, 1.5), model, cfg)
    assert len(reqs) == 5
    for r, p in zip(reqs, channel_positions((2.0, 2.0, 1.5), model)):
        np.testing.assert_allclose(r.listener, p)
        assert r.num_sources == 1 and r.seed == 7 and r.num_rays == 1000 and r.num_bounces == 20
        assert r.duration_s == 1.0 and r.num_bins == 100  # max(rt_duration, duration) at rt_bin_duration_s


def test_motion_frames_group_by_receiver():
    cfg = SimConfig(rt_num_rays=100, rt_num_bounces=5)
    rng = np.random.default_rng(0)
    fixed = (3.0, 2.0, 1.4)
    frames = [{"source": rng.uniform(0, 4, 3), "receiver": fixed} for _ in range(70)]
    frames.insert(10, {"source": (1.0, 1.0, 1.0), "receiver": (0.5, 0.5, 1.2)})
    plan = motion_traces("scene", frames, cfg)
    # 70 frames at the fixed receiver: 64 + 6 sources; one frame elsewhere
    assert sorted(t.num_sources for t in plan.traces) == [1, 6, MAX_SOURCES_PER_TRACE]
    for f, frame in enumerate(frames):
        (t, k), = plan.frames[f]
        tr = plan.traces[t]
        np.testing.assert_allclose(tr.sources[k], frame["source"])
        np.testing.assert_allclose(tr.listener, frame["receiver"])
        assert tr.seed == cfg.seed


def test_motion_frames_with_an_array_receiver():
    cfg = SimConfig(rt_num_rays=100, rt_num_bounces=5)
    model = microphone_array("linear", count=2, spacing_m=0.2)
    frames = [{"source": (float(i), 1.0, 1.0), "receiver": (2.0, 2.0, 1.5)} for i in range(3)]
    plan = motion_traces("scene", frames, cfg, model)
    assert len(plan.traces) == 2 and all(t.num_sources == 3 for t in plan.traces)
    for f in range(3):
        assert [tk for tk in plan.frames[f]] == [(0, f), (1, f)]
    with pytest.raises(ValueError):
        motion_traces("scene", [], cfg)
    with pytest.raises(ValueError):
        motion_traces("scene", [{"source": (0, 0, 0)}], cfg)


def test_planned_traces_carry_the_whole_configuration():
    from acoustic_agent_amd.steam_rt import histogram_bins

    cfg = SimConfig(rt_num_rays=500, rt_num_bounces=7, duration_s=0.5, rt_duration_s=0.9, rt_bin_duration_s=0.03,
                    rt_num_diffuse_samples=64, rt_specular_exponent=50.0, rt_source_radius=0.2, seed=3)
    model = {"type": "weighted_dipole", "dipole_weight": 0.37, "dipole_power": 2.0, "orientation_deg": 90.0}
    reqs = channel_traces("scene", (1, 1, 1), (2, 2, 1.5), microphone_array("linear", count=2), cfg,
                          source_model=model)
    # 0.9 / 0.03 is 30.000000000000004 in binary64: 31 bins, as trace_energy_field gives
    assert histogram_bins(0.9, 0.03) == 31
    for r in reqs:
        assert r.num_bins == 31 and r.num_diffuse_samples == 64 and r.render_ambisonics
        assert r.dipole_weight == 0.37 and r.dipole_power == 2.0
        np.testing.assert_allclose(r.source_forward, (0.0, 1.0, 0.0), atol=1e-15)
        assert r.specular_exponent == 50.0 and r.source_radius == 0.2
        p = r.params()
        assert p.num_bins == 31 and p.num_diffuse_samples == 64 and p.render_ambisonics == 1
        assert abs(p.dipole_weight - 0.37) < 1e-7 and p.dipole_power == 2.0
    plan = motion_traces("scene", [{"source": (1, 1, 1), "receiver": (2, 2, 1.5)}], cfg, source_model="cardioid",
                         render_ambisonics=False)
    assert plan.traces[0].dipole_weight == 0.5 and not plan.traces[0].render_ambisonics
    with pytest.raises(ValueError, match="counter"):
        channel_traces("scene", (1, 1, 1), (2, 2, 1.5), None, SimConfig(rt_rng="reference"))
