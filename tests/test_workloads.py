"""CPU checks of the benchmark workload generators."""
from __future__ import annotations

import math

import numpy as np

from acoustic_agent_amd import scenes
from acoustic_agent_amd.scene import point_in_polygon


def test_circular_array_matches_reference_layout():
    mics = scenes.circular_array((1.0, 2.0, 1.5), 4, 0.05)
    np.testing.assert_allclose(mics[0], (1.05, 2.0, 1.5))
    np.testing.assert_allclose(mics[1], (1.0, 2.05, 1.5), atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(mics[:, :2] - (1.0, 2.0), axis=1), 0.05, atol=1e-6)


def test_moving_source_path_speed_and_room():
    for seed in range(5):
        path = scenes.moving_source_path(seed, keyframes=64, fps=32.0)
        steps = np.linalg.norm(np.diff(path, axis=0), axis=1)
        # keyframes are equally spaced along the curve; chords only shorten across sharp turns
        speed = float(np.median(steps)) * 32.0
        assert 0.5 - 1e-3 <= speed <= 1.5 + 1e-3
        assert np.mean(np.abs(steps - np.median(steps)) <= 0.02 * np.median(steps)) >= 0.9
        assert np.all(steps <= np.median(steps) * 1.02)
        assert np.all((path[:, 0] >= 0.5) & (path[:, 0] <= 6.5) & (path[:, 1] >= 0.5) & (path[:, 1] <= 4.5))


def test_keyframe_scene_keeps_furniture_off_the_path():
    w, path = scenes.keyframe_workload(7, target_triangles=3000)
    arrays = w.scene.kernel_arrays()
    for si in np.flatnonzero(arrays["kinds"] == 2):
        c, hw, hd = arrays["box_center"][si], *arrays["box_half"][si]
        u, v = arrays["box_axis_u"][si], arrays["box_axis_v"][si]
        for p in path:
            rel = p[:2] - c
            inside = abs(rel @ u) <= hw and abs(rel @ v) <= hd
            assert not inside


def test_log_uniform_triangle_targets():
    from acoustic_agent_amd.workloads import log_uniform_triangles

    rng = np.random.default_rng(0)
    t = np.asarray([log_uniform_triangles(rng) for _ in range(2000)])
    assert t.min() >= 12 and t.max() <= 50_000
    med = np.median(np.log(t))
    assert abs(med - 0.5 * (math.log(12) + math.log(50_000))) < 0.3


def test_l_plan_source_and_listener_are_in_free_space():
    for seed in range(4):
        room = scenes.l_plan(seed)
        a = room.kernel_arrays()
        assert point_in_polygon(*scenes.L_SOURCE[:2], np.asarray(scenes.L_OUTLINE))
        for si in np.flatnonzero(a["kinds"] == 2):
            c, hw, hd = a["box_center"][si], *a["box_half"][si]
            for p in (scenes.L_SOURCE, scenes.L_LISTENER):
                rel = np.asarray(p[:2]) - c
                assert not (abs(rel @ a["box_axis_u"][si]) <= hw and abs(rel @ a["box_axis_v"][si]) <= hd)
