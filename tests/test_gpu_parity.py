"""GPU parity: the HIP tracer against the oracle on the same inputs.

Tolerances:
  * vs the binary32 oracle (same arithmetic as the kernel): identical hit
    triangle for every logged segment and identical per-surface hit counts on
    the shoebox; echogram relative L1 <= 1e-5 (only the order of the float
    atomic sums differs).
  * vs the binary64 oracle (the reference Numba semantics): per-band total
    energy within 1e-4 relative on the shoebox, band RT60 and EDT within 2%.
"""
from __future__ import annotations

import numpy as np
import pytest

from acoustic_agent_amd import scenes
from acoustic_agent_amd.models import SimConfig
from acoustic_agent_amd.steam_rt import estimate_early_decay_times, estimate_reverb_times

from helpers import Case, band_rel, rel_l1, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def shoebox():
    return scenes.shoebox_workload(seed=0)


@pytest.fixture(scope="module")
def lplan():
    return scenes.l_plan_workload(seed=0, target_triangles=20000)


def _field(echogram, bin_dur):
    return {"echogram": np.asarray(echogram, dtype=np.float64), "bin_duration_s": bin_dur,
            "num_bins": np.asarray(echogram).shape[-1]}


@pytest.mark.parametrize("rng", ["reference", "counter"])
def test_shoebox_matches_fp32_oracle_ray_for_ray(shoebox, rng):
    case = Case(shoebox, num_rays=4096, num_bounces=50, num_bins=1000, rng=rng, seed=5, log_bounces=50)
    gpu = run_gpu(case, log_rays=4096)
    ref = run_oracle(case, "float32", log_rays=4096)
    assert np.array_equal(gpu["log_triangle"], ref["log_prim"])
    assert np.array_equal(gpu["hit_counts"], ref["hit_counts"])
    assert np.array_equal(gpu["contrib_counts"], ref["contrib_counts"])
    assert gpu["active_count"] == ref["active_count"]
    assert gpu["actual_bounces"] == ref["actual_bounces"]
    assert rel_l1(gpu["echogram"], ref["echogram"]) <= 1e-5
    assert gpu["counters"]["bounce_queries"] == ref["counters"]["bounce_queries"]
    assert gpu["counters"]["shadow_queries"] == ref["counters"]["shadow_queries"]


def test_shoebox_matches_fp64_oracle(shoebox):
    case = Case(shoebox, num_rays=10000, num_bounces=50, num_bins=1000, seed=1)
    gpu = run_gpu(case)
    ref = run_oracle(case, "float64")
    assert np.all(band_rel(gpu["echogram"], ref["echogram"]) <= 1e-4)
    cfg = SimConfig()
    for est in (estimate_reverb_times, estimate_early_decay_times):
        a = est(_field(gpu["echogram"], case.bin_dur), cfg)
        b = est(_field(ref["echogram"], case.bin_dur), cfg)
        for band in a:
            assert b[band] > 0
            assert abs(a[band] - b[band]) / b[band] <= 0.02, (est.__name__, band, a[band], b[band])


@pytest.mark.parametrize("rng", ["reference", "counter"])
def test_lplan_matches_fp32_oracle(lplan, rng):
    # the furnished plan (19,952 triangles), every ray for 100 bounces: the
    # same hit triangle and the same binary32 segment length, bit for bit,
    # at every bounce of every ray
    case = Case(lplan, num_rays=4096, num_bounces=100, num_bins=2000, seed=2, log_bounces=100, rng=rng)
    gpu = run_gpu(case, log_rays=4096)
    ref = run_oracle(case, "float32", log_rays=4096)
    assert np.array_equal(gpu["log_triangle"], ref["log_prim"])
    hit = gpu["log_triangle"] >= 0
    assert np.array_equal(gpu["log_t"][hit].view(np.int32), np.asarray(ref["log_t"], np.float32)[hit].view(np.int32))
    assert np.array_equal(gpu["hit_counts"], ref["hit_counts"])
    assert np.array_equal(gpu["contrib_counts"], ref["contrib_counts"])
    assert gpu["active_count"] == ref["active_count"] and gpu["actual_bounces"] == ref["actual_bounces"]
    for key in ("bounce_queries", "shadow_queries"):
        assert gpu["counters"][key] == ref["counters"][key], key
    assert rel_l1(gpu["echogram"], ref["echogram"]) <= 1e-5
    assert np.all(band_rel(gpu["echogram"], ref["echogram"]) <= 1e-5)


def test_lplan_rt60_matches_fp64_oracle(lplan):
    case = Case(lplan, num_rays=20000, num_bounces=200, num_bins=2000, seed=3)
    gpu = run_gpu(case)
    ref = run_oracle(case, "float64")
    cfg = SimConfig()
    for est in (estimate_reverb_times, estimate_early_decay_times):
        a = est(_field(gpu["echogram"], case.bin_dur), cfg)
        b = est(_field(ref["echogram"], case.bin_dur), cfg)
        for band in a:
            assert b[band] > 0
            assert abs(a[band] - b[band]) / b[band] <= 0.02, (est.__name__, band, a[band], b[band])


def test_bvh_equals_linear_on_gpu(lplan):
    case = Case(lplan, num_rays=512, num_bounces=20, num_bins=2000, seed=4, log_bounces=20)
    a = run_gpu(case, use_bvh=True, log_rays=512)
    b = run_gpu(case, use_bvh=False, log_rays=512)
    assert np.array_equal(a["log_triangle"], b["log_triangle"])
    assert np.array_equal(a["hit_counts"], b["hit_counts"])
    assert rel_l1(a["echogram"], b["echogram"]) <= 1e-5


@pytest.mark.parametrize("tuning", [
    {"block_size": 64, "lds_histogram": 0},
    {"block_size": 128, "lds_histogram": 1, "workgroups_per_cu": 2},
    {"block_size": 512, "lds_histogram": 1, "workgroups_per_cu": 1},
])
def test_launch_shapes_agree(lplan, tuning):
    case = Case(lplan, num_rays=4096, num_bounces=60, num_bins=2000, seed=6, rng="counter")
    base = run_gpu(case)
    other = run_gpu(case, tuning=tuning)
    assert np.array_equal(base["hit_counts"], other["hit_counts"])
    assert rel_l1(other["echogram"], base["echogram"]) <= 1e-5


@pytest.mark.parametrize("sources_per_lane", [0, 2])
def test_multi_source_equals_separate_traces(lplan, sources_per_lane):
    rng = np.random.default_rng(9)
    srcs = np.stack([np.asarray(scenes.L_SOURCE) + rng.uniform(-0.5, 0.5, 3) * (1, 1, 0.2) for _ in range(5)])
    case = Case(lplan, num_rays=2048, num_bounces=80, num_bins=1000, seed=7, rng="counter", sources=srcs)
    multi = run_gpu(case, tuning={"sources_per_lane": sources_per_lane})
    for k in range(srcs.shape[0]):
        single = run_gpu(case, source=srcs[k])
        assert rel_l1(multi["echogram"][k], single["echogram"]) <= 1e-5
        assert np.array_equal(multi["hit_counts"][k], single["hit_counts"])
        assert int(multi["active_count"][k]) == single["active_count"]
        assert int(multi["actual_bounces"][k]) == single["actual_bounces"]


def test_ambisonic_channel_ratio_is_initial_direction(shoebox):
    # tests/test_steam_audio_parity.py:329-357: one ray, FOA/W equals its direction
    case = Case(shoebox, num_rays=1, num_bounces=12, num_bins=1000, seed=0, ambisonics=True)
    gpu = run_gpu(case)
    from acoustic_agent_amd import sampling

    d = sampling.sphere_samples(1, 0)[0].astype(np.float32)
    w = gpu["ambisonic"][0, 0]
    valid = w > 0
    assert np.count_nonzero(valid) > 1
    for ch in range(3):
        np.testing.assert_allclose(gpu["ambisonic"][0, ch + 1, valid] / w[valid], d[ch], rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("rng", ["reference", "counter"])
def test_plain_kernel_matches_fp32_oracle_ray_for_ray(shoebox, rng):
    # the one-thread-per-ray baseline kernel gives the same trajectories
    case = Case(shoebox, num_rays=4096, num_bounces=50, num_bins=1000, rng=rng, seed=5, log_bounces=50)
    gpu = run_gpu(case, log_rays=4096, plain=True)
    ref = run_oracle(case, "float32", log_rays=4096)
    assert np.array_equal(gpu["log_triangle"], ref["log_prim"])
    assert np.array_equal(gpu["hit_counts"], ref["hit_counts"])
    assert np.array_equal(gpu["contrib_counts"], ref["contrib_counts"])
    assert gpu["active_count"] == ref["active_count"]
    assert gpu["actual_bounces"] == ref["actual_bounces"]
    assert rel_l1(gpu["echogram"], ref["echogram"]) <= 1e-5
    assert gpu["counters"]["bounce_queries"] == ref["counters"]["bounce_queries"]
    assert gpu["counters"]["shadow_queries"] == ref["counters"]["shadow_queries"]


@pytest.mark.parametrize("use_bvh", [True, False])
def test_plain_kernel_equals_persistent_kernel(lplan, use_bvh):
    n = 4096 if use_bvh else 256
    case = Case(lplan, num_rays=n, num_bounces=60, num_bins=2000, seed=11, rng="counter", log_bounces=60)
    a = run_gpu(case, use_bvh=use_bvh, log_rays=n)
    b = run_gpu(case, use_bvh=use_bvh, log_rays=n, plain=True)
    assert np.array_equal(a["log_triangle"], b["log_triangle"])
    assert np.array_equal(a["hit_counts"], b["hit_counts"])
    assert np.array_equal(a["contrib_counts"], b["contrib_counts"])
    assert a["counters"]["shadow_queries"] == b["counters"]["shadow_queries"]
    assert rel_l1(b["echogram"], a["echogram"]) <= 1e-5


def test_plain_kernel_multi_source(lplan):
    rng = np.random.default_rng(12)
    srcs = np.stack([np.asarray(scenes.L_SOURCE) + rng.uniform(-0.5, 0.5, 3) * (1, 1, 0.2) for _ in range(3)])
    case = Case(lplan, num_rays=2048, num_bounces=60, num_bins=1000, seed=13, rng="counter", sources=srcs)
    a = run_gpu(case)
    b = run_gpu(case, plain=True)
    assert np.array_equal(a["hit_counts"], b["hit_counts"])
    assert np.array_equal(a["active_count"], b["active_count"])
    assert np.array_equal(a["actual_bounces"], b["actual_bounces"])
    assert rel_l1(b["echogram"], a["echogram"]) <= 1e-5
