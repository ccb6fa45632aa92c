"""The batched and served paths on their own full-size workloads, against the
oracle.

* C0 and C1 hit logs against the binary64 oracle, ray for ray: every ray whose
  triangle sequence differs must part ways where the geometry explains it
  (helpers.hit_log_divergences): at a grazing hit (within a micrometre per
  bounce of an edge or of the other trace's triangle, while the traces agree
  up to rounding), after the binary32 trace drifted measurably away from the
  binary64 one (the hit then lies within that drift of an edge), or at an end
  test on an identical segment; nothing unexplained, and few rays that part
  at all.  The first-divergent-bounce histogram is printed.
* C2, 256 scenes (12 to 50k triangles) x 4-microphone ring = 1024 traces in
  one persistent launch: two submissions give identical integer outputs; every
  trace's tallies agree with each other; 8 traces spanning the triangle range,
  logged ray for ray inside the full batch, match the binary32 oracle (the
  reference CUDA path's arithmetic); two traces' band RT60 / EDT are within 2 %
  of the binary64 oracle.
* C3, 32 furnished plans x 64 keyframes = 2048 RIRs in one hipGraph-replayed
  batch: replays agree; sampled keyframes equal their single-source traces,
  which match the binary32 oracle.
* C4, the 50 requests/s Poisson stream over 1.5 s through the serving loop:
  every result equals the same trace run alone.

Tolerances: integer outputs exact; echograms relative L1 <= 1e-5 where the
rays agree (only the order of float atomic sums differs); rays that part ways
at a grazing hit may move their energy, bounded by their share of the rays.
"""
from __future__ import annotations

import numpy as np
import pytest

from acoustic_agent_amd import scenes, workloads
from acoustic_agent_amd.batch import BatchEngine, BatchTuning, trace_sequential
from acoustic_agent_amd.models import SimConfig
from acoustic_agent_amd.steam_rt import estimate_early_decay_times, estimate_reverb_times

from helpers import Case, hit_log_divergences, oracle_for_request, rel_l1, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

INT_KEYS = ("hit_counts", "contrib_counts", "actual_bounces", "active_count")


def _report(name, div):
    print(f"{name}: {div['same']}/{div['rays']} rays identical; parted: {div['grazing']} grazing, "
          f"{div['drift']} after drifting apart (up to {div['max_drift_m']:.2e} m), {div['termination']} end "
          f"tests, {len(div['other'])} unexplained; first divergent bounce {div['first_bounce_hist']}; worst "
          f"explained distance / tolerance {div['worst_grazing_ratio']:.3f}", flush=True)


def _check_divergences(div, max_fraction):
    assert not div["other"], div["other"][:10]
    assert div["same"] >= (1.0 - max_fraction) * div["rays"], div
    assert div["drift"] + div["termination"] <= max_fraction * div["rays"], div


def _field(echogram, bin_dur):
    e = np.asarray(echogram, dtype=np.float64)
    return {"echogram": e, "bin_duration_s": bin_dur, "num_bins": e.shape[-1]}


def _rt_errors(a, b, bin_dur):
    cfg = SimConfig()
    worst = 0.0
    for est in (estimate_reverb_times, estimate_early_decay_times):
        ra, rb = est(_field(a, bin_dur), cfg), est(_field(b, bin_dur), cfg)
        for band in rb:
            assert rb[band] > 0, (est.__name__, band)
            worst = max(worst, abs(ra[band] - rb[band]) / rb[band])
    return worst


def _check_tallies(res, req):
    """Tallies of one trace (every source) agree with each other."""
    c = res["counters"]
    assert int(res["contrib_counts"].sum()) == c["contributions"]
    e = res["echogram"].astype(np.float64)
    # binary32 atomic sums: a surface's energy adds up ~10^6 contributions in
    # one address, the echogram spreads them over thousands of bins
    np.testing.assert_allclose(res["surface_energy"].astype(np.float64).sum(axis=1), e.sum(axis=(1, 2)), rtol=1e-3)
    assert np.all(res["contrib_counts"] <= res["hit_counts"])
    assert np.all(res["active_count"] <= req.num_rays)
    assert np.all((res["actual_bounces"] >= 1) & (res["actual_bounces"] <= req.num_bounces))
    # one bounce query per ray start and per continued hit; hits are counted per source
    assert c["bounce_queries"] >= req.num_rays
    assert int(res["hit_counts"].sum()) <= c["bounce_queries"] * req.num_sources
    assert np.all(np.isfinite(e)) and np.all(e >= 0)


# ---------------------------------------------------------------------------
# C0 / C1 against the binary64 oracle, ray for ray
# ---------------------------------------------------------------------------


def test_c0_hit_log_matches_fp64_oracle():
    w = scenes.shoebox_workload(seed=0)
    case = Case(w, num_rays=10000, num_bounces=50, num_bins=1000, seed=1729, log_bounces=50)
    gpu = run_gpu(case, log_rays=10000)
    ref = run_oracle(case, "float64", log_rays=10000)
    div = hit_log_divergences(gpu["log_triangle"], gpu["log_point"], ref["log_prim"], ref["log_point"],
                              w.mesh, w.listener)
    _report("C0 vs binary64", div)
    _check_divergences(div, 0.01)
    assert _rt_errors(gpu["echogram"], ref["echogram"], case.bin_dur) <= 0.02


def test_c1_hit_log_sample_matches_fp64_oracle():
    w = scenes.l_plan_workload(seed=0, target_triangles=20000)
    n_log = 20000
    case = Case(w, num_rays=200000, num_bounces=200, num_bins=2000, seed=0, rng="counter", log_bounces=200)
    gpu = run_gpu(case, log_rays=n_log)
    ref = run_oracle(case, "float64", log_rays=n_log, ray_range=(0, n_log))
    div = hit_log_divergences(gpu["log_triangle"], gpu["log_point"], ref["log_prim"][:n_log],
                              ref["log_point"][:n_log], w.mesh, w.listener)
    _report("C1 (first 20k rays) vs binary64", div)
    _check_divergences(div, 0.02)


# ---------------------------------------------------------------------------
# C2: 256 scenes x 4-microphone ring in one persistent launch
# ---------------------------------------------------------------------------

C2_LOGGED = 8


@pytest.fixture(scope="module")
def c2():
    reqs = workloads.mic_array_batch(num_scenes=256, seed=0, num_bounces=200)
    tris = np.asarray([r.scene.info["num_triangles"] for r in reqs])
    # one trace per triangle-count octile (the fewest rays of its octile), logged ray for ray
    order = np.argsort(tris, kind="stable")
    picks = []
    for chunk in np.array_split(order, C2_LOGGED):
        picks.append(int(min(chunk, key=lambda i: reqs[i].num_rays)))
    for i in picks:
        reqs[i].log_rays, reqs[i].log_bounces = reqs[i].num_rays, reqs[i].num_bounces
    engine = BatchEngine(0, max_traces=len(reqs))
    first = engine.submit(reqs).wait()
    for i in picks:
        reqs[i].log_rays = reqs[i].log_bounces = 0
    second = engine.submit(reqs).wait()
    engine.close()
    by_tag = [{r["tag"]: r for r in res} for res in (first, second)]
    yield reqs, picks, by_tag[0], by_tag[1]


def test_c2_full_batch_reproducible_and_consistent(c2):
    reqs, _, a, b = c2
    assert len(reqs) == 1024
    for r in reqs:
        x, y = a[r.tag], b[r.tag]
        for k in INT_KEYS:
            assert np.array_equal(x[k], y[k]), (r.tag, k)
        for k in ("bounce_queries", "shadow_queries", "contributions", "escaped_rays", "nodes_bounce",
                  "nodes_shadow"):
            assert x["counters"][k] == y["counters"][k], (r.tag, k)
        assert rel_l1(x["echogram"], y["echogram"]) <= 1e-5, r.tag
        _check_tallies(x, r)


def test_c2_logged_traces_match_fp32_oracle(c2):
    reqs, picks, a, _ = c2
    tris = sorted(reqs[i].scene.info["num_triangles"] for i in picks)
    assert tris[0] <= 100 and tris[-1] >= 20000, tris
    for i in picks:
        r = reqs[i]
        gpu = a[r.tag]
        ref = oracle_for_request(r, precision="float32", log=True)
        div = hit_log_divergences(gpu["log_triangle"], gpu["log_point"], ref["log_prim"], ref["log_point"],
                                  r.scene.workload.mesh, r.listener)
        _report(f"C2 trace {r.tag} ({r.num_rays} rays) vs binary32", div)
        # binary32 against binary32: every ray of the trace hits the same
        # triangles (the same arithmetic in the same order)
        assert div["same"] == div["rays"], div
        assert np.array_equal(gpu["hit_counts"][0], ref["hit_counts"])
        assert np.array_equal(gpu["contrib_counts"][0], ref["contrib_counts"])
        assert int(gpu["active_count"][0]) == ref["active_count"]
        assert int(gpu["actual_bounces"][0]) == ref["actual_bounces"]
        for key in ("bounce_queries", "shadow_queries"):
            assert gpu["counters"][key] == ref["counters"][key], (r.tag, key)
        assert rel_l1(gpu["echogram"][0], ref["echogram"]) <= 1e-5
        assert gpu["counters"]["contributions"] == int(gpu["contrib_counts"].sum())


def test_c2_band_rt60_matches_fp64_oracle(c2):
    reqs, picks, a, _ = c2
    # the largest and a mid-sized logged scene, full ray sets
    for i in (picks[-1], picks[len(picks) // 2]):
        r = reqs[i]
        ref = oracle_for_request(r, precision="float64")
        err = _rt_errors(a[r.tag]["echogram"][0], ref["echogram"], r.bin_dur)
        print(f"C2 trace {r.tag}: band RT60/EDT max rel. error vs binary64 {err:.2e}", flush=True)
        assert err <= 0.02


# ---------------------------------------------------------------------------
# C3: 32 furnished plans x 64 keyframes, hipGraph replay
# ---------------------------------------------------------------------------


@pytest.fixture(scope="module")
def c3():
    reqs = workloads.keyframe_batch(num_scenes=32, seed=0, num_bounces=200)
    engine = BatchEngine(0, max_traces=len(reqs))
    tuning = BatchTuning(use_graph=True)
    runs = [{r["tag"]: r for r in engine.submit(reqs, tuning).wait()} for _ in range(2)]
    engine.close()
    yield reqs, runs[0], runs[1]


def test_c3_graph_replays_agree_and_tallies_hold(c3):
    reqs, a, b = c3
    assert len(reqs) == 32 and sum(r.num_sources for r in reqs) == 2048
    for r in reqs:
        x, y = a[r.tag], b[r.tag]
        for k in INT_KEYS:
            assert np.array_equal(x[k], y[k]), (r.tag, k)
        assert x["counters"]["contributions"] == y["counters"]["contributions"]
        assert rel_l1(x["echogram"], y["echogram"]) <= 1e-5
        _check_tallies(x, r)


def test_c3_sampled_keyframes_match_single_traces_and_fp32_oracle(c3):
    import dataclasses

    reqs, a, _ = c3
    for scene_index, key in ((0, 0), (7, 31), (19, 63)):
        r = reqs[scene_index]
        batch = a[r.tag]
        one = dataclasses.replace(r, sources=np.asarray(r.sources)[key:key + 1], log_rays=r.num_rays,
                                  log_bounces=r.num_bounces, tag=None)
        (alone,), _ = trace_sequential([one])
        for k in INT_KEYS:
            assert np.array_equal(batch[k][key], alone[k][0]), (scene_index, key, k)
        assert rel_l1(batch["echogram"][key], alone["echogram"][0]) <= 1e-5
        ref = oracle_for_request(one, precision="float32", log=True)
        div = hit_log_divergences(alone["log_triangle"], alone["log_point"], ref["log_prim"], ref["log_point"],
                                  r.scene.workload.mesh, r.listener)
        _report(f"C3 scene {scene_index} keyframe {key} vs binary32", div)
        assert div["same"] == div["rays"], div  # every ray, as for the mic arrays
        assert rel_l1(alone["echogram"][0], ref["echogram"]) <= 1e-5
        assert np.array_equal(alone["hit_counts"][0], ref["hit_counts"])
        assert np.array_equal(alone["contrib_counts"][0], ref["contrib_counts"])


# ---------------------------------------------------------------------------
# C4: the served stream, each result against its traces alone
# ---------------------------------------------------------------------------


def test_c4_served_results_equal_traces_alone():
    from acoustic_agent_amd.serving import ServingLoop

    stream = workloads.served_stream(50.0, 1.5, seed=50)
    kinds = {req.kind for req in stream}
    assert {"shoebox", "single", "mic_array", "keyframes"} <= kinds, kinds
    loop = ServingLoop(0, slots=3, max_batch_traces=256)
    done, summary = loop.run(stream)
    loop.close()
    assert summary["requests"] == len(stream)
    print(f"C4 served: {summary['latency_all']} rirs/s {summary['rirs_per_s']}", flush=True)
    for c in done:
        alone, _ = trace_sequential(c.request.traces)
        assert len(alone) == len(c.results)
        for served, solo in zip(c.results, alone):
            for k in INT_KEYS:
                assert np.array_equal(served[k], solo[k]), (c.request.kind, k)
            assert rel_l1(served["echogram"], solo["echogram"]) <= 1e-5
        rt_alone = estimate_reverb_times(_field(alone[0]["echogram"][0], c.request.traces[0].bin_dur), SimConfig())
        for band, v in rt_alone.items():
            assert c.rt60[band] == pytest.approx(v, rel=1e-3, abs=1e-6)


# ---------------------------------------------------------------------------
# planned traces (receivers, motion) carry the trace_energy_field configuration
# ---------------------------------------------------------------------------


def test_planned_dipole_ambisonic_trace_equals_trace_energy_field():
    from acoustic_agent_amd import hip_backend
    from acoustic_agent_amd.plan import channel_traces
    from acoustic_agent_amd.steam_rt import trace_energy_field

    w = scenes.l_plan_workload(seed=3, target_triangles=4000)
    cfg = SimConfig(rt_num_rays=4096, rt_num_bounces=60, rt_bin_duration_s=1e-3, rt_duration_s=0.9,
                    duration_s=0.5, rt_num_diffuse_samples=64, seed=11)
    model = {"type": "weighted_dipole", "dipole_weight": 0.37, "dipole_power": 2.0, "orientation_deg": 120.0,
             "elevation_deg": 10.0}
    field = trace_energy_field(w.scene, w.source, w.listener, cfg, model, render_ambisonics=True)
    scene = hip_backend.cached_scene(w.scene.kernel_arrays(), device=0)
    (req,) = channel_traces(scene, w.source, w.listener, None, cfg, source_model=model)
    engine = BatchEngine(0, max_traces=1)
    (res,) = engine.submit([req]).wait()
    engine.close()
    assert res["echogram"].shape[-1] == field["num_bins"]
    assert rel_l1(res["echogram"][0], field["echogram"]) <= 1e-5
    assert rel_l1(res["ambisonic"][0], field["ambisonic_echogram"]) <= 1e-5
    names = w.scene.kernel_arrays()["names"]
    hits = {names[i]: int(v) for i, v in enumerate(res["hit_counts"][0]) if v > 0}
    assert hits == field["surface_hit_count"]
