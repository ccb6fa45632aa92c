"""The hit-log classifier used by the ray-for-ray GPU tests, checked on the
CPU with the oracle: binary32 against binary64 traces of the same rays part
ways only at grazing hits, end tests or after drifting apart, and wrong hits
planted in one log are reported as real disagreements."""
from __future__ import annotations

import numpy as np
import pytest

from acoustic_agent_amd import scenes

from helpers import Case, edge_distance, hit_log_divergences, run_oracle


@pytest.fixture(scope="module")
def logs():
    w = scenes.l_plan_workload(seed=2, target_triangles=3000)
    case = Case(w, num_rays=3000, num_bounces=100, num_bins=2000, seed=4, rng="counter", log_bounces=100)
    a = run_oracle(case, "float32", log_rays=3000)
    b = run_oracle(case, "float64", log_rays=3000)
    return w, a, b


def test_binary32_and_binary64_part_only_at_grazing_hits(logs):
    w, a, b = logs
    div = hit_log_divergences(a["log_prim"], a["log_point"], b["log_prim"], b["log_point"], w.mesh, w.listener)
    assert not div["other"], div["other"]
    assert div["same"] >= 0.97 * div["rays"]
    assert div["drift"] + div["termination"] <= 0.02 * div["rays"]


def test_planted_wrong_hits_are_reported(logs):
    w, a, b = logs
    v = w.mesh.vertices.reshape(-1, 3, 3).astype(np.float64)
    tri, pts = a["log_prim"].copy(), a["log_point"].copy()
    same = np.flatnonzero(np.all(tri == b["log_prim"], axis=1) & (tri[:, 30] >= 0))
    j, planted = 20, []
    for ray in same:
        prev, hit = int(tri[ray, j - 1]), int(tri[ray, j])
        if edge_distance(b["log_point"][ray, j - 1], v[prev]) < 1e-3:
            continue
        # a triangle far from the true hit, hit at its centroid
        far = int(np.argmax(np.linalg.norm(v.mean(axis=1) - v[hit].mean(axis=0), axis=1)))
        tri[ray, j:] = far
        pts[ray, j:] = v[far].mean(axis=0)
        planted.append(int(ray))
        if len(planted) == 20:
            break
    assert len(planted) == 20
    div = hit_log_divergences(tri, pts, b["log_prim"], b["log_point"], w.mesh, w.listener)
    flagged = {ray for ray, *_ in div["other"]}
    assert set(planted) <= flagged
    # a wrong triangle index whose point is not on it is reported too
    tri2 = a["log_prim"].copy()
    tri2[planted[0], j] = (int(tri2[planted[0], j]) + 1) % v.shape[0]
    div2 = hit_log_divergences(tri2, a["log_point"], b["log_prim"], b["log_point"], w.mesh, w.listener)
    assert planted[0] in {ray for ray, *_ in div2["other"]}
