"""Benchmark of the MI355X reflection tracer.

Default line (``python bench.py``, one GPU): the batched mic-array workload,
256 seeded scenes (triangle counts log-uniform in 12..50k: shoeboxes and
furnished multi-room plans) x a 4-microphone ring (5 cm) = 1024 traces of
20k-200k rays, 200 bounces, 2.0 s at 1 ms x 6 bands, binary32, counter RNG,
traced in one persistent launch with the autotuned launch shape.  A step is
one submission of the whole batch (the scenes and every buffer resident in
HBM).  ``value`` is RIRs/s (completed traces per second).  The line carries:

* ``roofline`` of ``batch_kernel``: algorithmic bytes per launch (segment
  queries x the canonical binary-BVH node / triangle counts of the binary64
  oracle, x 32 B / 48 B, + 48 B per contribution) over the HIP-event time of
  the submission, against 8 TB/s; next to it the measured HBM traffic and
  issue counters of the same kernel (``profiles/``) and the limiter they show;
* ``cpu_baseline``: the binary64 oracle (plain-C restatement of the reference
  Numba kernel) on a seeded sample of the same traces, on all host cores and
  on one thread, median of 3 after a warm-up, trace only and with host RT60;
* ``accuracy_vs_fp64_oracle``: band RT60 / EDT of full traces against the
  binary64 oracle; the line is marked invalid (exit 3) if band RT60 or EDT
  is off by more than 2 %;
* ``baselines``: the same sample of traces one per launch on the plain
  kernel (the reference GPU backend's execution model) and on the persistent
  single-trace kernel, the single-scene trace (furnished L plan, 200k rays),
  and the served stream (Poisson arrivals at 50 / 100 / 200 / 400 req/s):
  p50/p99 latency per size class and per kind against sequential execution.

Other modes (``--workload``): ``c1`` the single-scene trace alone with its
counters (``--plain``: on the plain kernel), ``c3`` the batched 64-keyframe workload, ``c4`` the served
stream alone, ``--copy-probe`` the PMC calibration copy.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))
from __graft_entry__ import load_package  # noqa: E402

load_package()
from acoustic_agent_amd import _lib, hip_backend, sampling, scenes, workloads  # noqa: E402
from acoustic_agent_amd.models import SimConfig  # noqa: E402
from acoustic_agent_amd.steam_rt import estimate_early_decay_times, estimate_reverb_times  # noqa: E402

C = 343.0
HBM_PEAK_GBS = 8000.0
BYTES_PER_NODE = 32   # FP32 AABB (24 B) + packed start/count/escape (8 B)
BYTES_PER_TRI = 48    # 3 x FP32 vertices + id / padding
BYTES_PER_CONTRIB = 48  # 6 bands x 8 B histogram read-modify-write
METRIC = "RIRs/s + rays/s at <=2% band-RT60 err vs FP64 oracle; p50/p99 served latency"
RT60_TOLERANCE = 0.02


def _progress(line: str) -> None:
    print(line, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _field(echogram, bin_dur):
    e = np.asarray(echogram, dtype=np.float64)
    return {"echogram": e, "bin_duration_s": bin_dur, "num_bins": e.shape[-1]}


def band_errors(gpu, ref, bin_dur: float) -> dict:
    cfg = SimConfig()
    out = {}
    for name, est in (("rt60", estimate_reverb_times), ("edt", estimate_early_decay_times)):
        fa, fb = est(_field(gpu, bin_dur), cfg), est(_field(ref, bin_dur), cfg)
        errs = [abs(fa[k] - fb[k]) / fb[k] for k in fb if fb[k] > 0]
        out[name + "_max_rel_err"] = float(max(errs)) if errs else None
    tot_g, tot_o = np.asarray(gpu, np.float64).sum(axis=-1), np.asarray(ref, np.float64).sum(axis=-1)
    out["band_energy_max_rel_err"] = float(np.max(np.abs(tot_g - tot_o) / np.maximum(tot_o, 1e-300)))
    return out


def _traffic(name: str):
    path = REPO / "profiles" / name
    if path.exists():
        try:
            return json.loads(path.read_text())
        except (OSError, ValueError):
            return None
    return None


# ---------------------------------------------------------------------------
# the binary64 oracle: CPU baseline, canonical traversal counts, accuracy
# ---------------------------------------------------------------------------


def cpu_baseline_leg(requests, results, *, sample: int, seed: int, counts_rays: int, accuracy_traces: int,
                     reps: int = 3) -> dict:
    """Everything that runs the oracle (as the checker; nothing here is the
    product path):

    * the CPU baseline: ``sample`` seeded traces of the workload, each over
      8 evenly spaced ray ranges covering 1/8 of its rays (all cores) or
      1/128 (one thread); one warm-up then the median of ``reps``; reported
      as trace-equivalents per second, trace only and with host RT60;
    * canonical traversal counts per segment query, per scene, from a
      ``counts_rays``-ray binary64 trace of each distinct scene (the reference
      binary BVH walked the reference way: closest hit for bounce queries,
      any hit for shadow queries);
    * accuracy: full binary64 traces of ``accuracy_traces`` traces (the fewest
      rays of each triangle-count class) against the GPU's echograms.
    """
    from oracle import oracle as O

    cores = len(os.sched_getaffinity(0))
    scenes_o: dict[int, object] = {}

    def oscene(r):
        key = id(r.scene)
        if key not in scenes_o:
            w = r.scene.workload
            scenes_o[key] = O.OracleScene(w.scene.kernel_arrays(), w.mesh, "float64")
        return scenes_o[key]

    def kwargs(r, num_rays=None):
        src = np.asarray(r.sources, dtype=np.float64).reshape(-1, 3)[0]
        return dict(source=src, listener=r.listener, num_rays=num_rays or r.num_rays, num_bounces=r.num_bounces,
                    num_bins=r.num_bins, bin_dur=r.bin_dur, c=r.c, max_path_len=r.duration_s * r.c,
                    direct_delay=float(np.linalg.norm(src - np.asarray(r.listener))) / r.c,
                    diffuse_bank=sampling.diffuse_sample_bank(r.num_diffuse_samples), rng="counter", seed=r.seed)

    rng = np.random.default_rng(seed)
    picks = sorted(int(i) for i in rng.choice(len(requests), size=min(sample, len(requests)), replace=False))

    def one_pass(fraction: float, threads: int) -> tuple[float, float, float]:
        """(seconds tracing, seconds of host RT60, trace-equivalents)."""
        t_trace = t_rt = 0.0
        done = 0.0
        for i in picks:
            r = requests[i]
            n = r.num_rays
            ranges = 8
            span = max(1, int(n * fraction / ranges))
            starts = np.linspace(0, n - span, ranges).astype(int)
            acc = None
            t = time.perf_counter()
            for s in starts:
                res = O.trace(oscene(r), ray_range=(int(s), int(s) + span), threads=threads, **kwargs(r))
                acc = res["echogram"] if acc is None else acc + res["echogram"]
            t_trace += time.perf_counter() - t
            t = time.perf_counter()
            estimate_reverb_times(_field(acc, r.bin_dur), SimConfig())
            t_rt += time.perf_counter() - t
            done += ranges * span / n
        return t_trace, t_rt, done

    out = {}
    for label, fraction, threads in (("all_cores", 1.0 / 8.0, cores), ("one_thread", 1.0 / 128.0, 1)):
        one_pass(fraction / 4.0, threads)  # warm-up
        passes = [one_pass(fraction, threads) for _ in range(reps)]
        tr = statistics.median(p[0] for p in passes)
        rt = statistics.median(p[1] for p in passes)
        done = passes[0][2]
        out[label] = {"threads": threads, "rirs_per_s": round(done / tr, 5),
                      "rirs_per_s_with_rt60": round(done / (tr + rt), 5),
                      "seconds_per_pass": [round(p[0], 3) for p in passes], "trace_equivalents": round(done, 4)}
        _progress(f"cpu baseline {label}: {out[label]['rirs_per_s']} RIRs/s ({threads} threads)")
    out["picks"] = picks
    out["cores"] = cores

    # canonical traversal counts per query, per distinct scene
    counts = {}
    for r in requests:
        key = id(r.scene)
        if key in counts:
            continue
        res = O.trace(oscene(r), threads=cores, **kwargs(r, num_rays=counts_rays))["counters"]
        qb, qs = max(res["bounce_queries"], 1), max(res["shadow_queries"], 1)
        counts[key] = (res["nodes_bounce"] / qb, res["prims_bounce"] / qb, res["nodes_shadow"] / qs,
                       res["prims_shadow"] / qs)
    out["canonical_counts"] = counts

    # accuracy on full traces: the fewest rays in each triangle-count class
    tris = np.asarray([r.scene.info["num_triangles"] for r in requests])
    order = np.argsort(tris, kind="stable")
    acc = []
    for chunk in np.array_split(order, accuracy_traces):
        i = int(min(chunk, key=lambda k: requests[k].num_rays))
        r = requests[i]
        ref = O.trace(oscene(r), threads=cores, **kwargs(r))
        err = band_errors(results[i]["echogram"][0], ref["echogram"], r.bin_dur)
        err.update(trace=list(r.tag) if isinstance(r.tag, tuple) else r.tag, rays=r.num_rays,
                   triangles=int(tris[i]))
        acc.append(err)
    out["accuracy"] = acc
    return out


def algorithmic_bytes(requests, results, counts) -> int:
    total = 0.0
    for r, res in zip(requests, results):
        nb, tb, ns, ts = counts[id(r.scene)]
        c = res["counters"]
        total += (c["bounce_queries"] * (BYTES_PER_NODE * nb + BYTES_PER_TRI * tb)
                  + c["shadow_queries"] * (BYTES_PER_NODE * ns + BYTES_PER_TRI * ts)
                  + c["contributions"] * BYTES_PER_CONTRIB)
    return int(total)


# ---------------------------------------------------------------------------
# timing helpers
# ---------------------------------------------------------------------------


def time_batch(engine, requests, tuning, steps: int, warmup: int):
    """Untimed warm-up submissions, then ``steps`` timed ones: wall time of
    the K steps between synchronisations, and the HIP-event time of each
    submission on its stream.  Returns (wall s, event ms list, last results
    in request order)."""
    import torch

    stream = torch.cuda.current_stream(0)
    for _ in range(warmup):
        engine.submit(requests, tuning, stream=stream, order=False).wait()
    torch.cuda.synchronize()
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    pending = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in events:
        a.record(stream)
        pending.append(engine.submit(requests, tuning, stream=stream, order=False))
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = [a.elapsed_time(b) for a, b in events]
    return wall, ms, pending[-1].wait()


def sample_baselines(requests, sample: int, seed: int) -> dict:
    """The same seeded sample of single-source traces, one per launch (each
    synchronised and copied back, as one reference call per trace): the plain
    one-thread-per-ray kernel and the persistent single-trace kernel."""
    from acoustic_agent_amd.batch import trace_sequential

    units = [dataclasses.replace(r, sources=np.asarray(r.sources).reshape(-1, 3)[k:k + 1], tag=(r.tag, k))
             for r in requests for k in range(r.num_sources)]
    rng = np.random.default_rng(seed)
    pick = sorted(int(i) for i in rng.choice(len(units), size=min(sample, len(units)), replace=False))
    chosen = [units[i] for i in pick]
    trace_sequential(chosen[:1], plain=True)
    trace_sequential(chosen[:1])
    _, plain_s = trace_sequential(chosen, plain=True)
    _, single_s = trace_sequential(chosen)
    _progress(f"one trace per launch on {len(chosen)} traces: plain {plain_s:.3f} s, single {single_s:.3f} s")
    return {"sample_traces": len(chosen), "plain_rirs_per_s": round(len(chosen) / plain_s, 3),
            "single_trace_rirs_per_s": round(len(chosen) / single_s, 3)}


def c1_single_trace(steps: int, seed: int, plain: bool = False) -> dict:
    """The single-scene trace: furnished L plan (~20k triangles), 200k rays,
    200 bounces, 2.0 s at 1 ms, one persistent launch per trace (``plain``:
    one launch of the one-thread-per-ray kernel per trace instead)."""
    import torch

    w = scenes.l_plan_workload(seed=seed, target_triangles=20_000)
    a = w.scene.kernel_arrays()
    scene = hip_backend.DeviceScene(w.mesh, np.asarray(a["reflection"]), np.asarray(a["scattering"]), device=0)
    N, B, K, bin_dur = 200_000, 200, 2000, 1e-3
    bank = torch.from_numpy(np.ascontiguousarray(sampling.diffuse_sample_bank(128), dtype=np.float32)).to("cuda:0")
    delay = float(np.linalg.norm(w.source - w.listener)) / C
    src = torch.tensor([[*w.source, delay]], dtype=torch.float32, device="cuda:0")
    buffers = hip_backend.TraceBuffers.allocate(0, 1, K, scene.num_surfaces, False)
    params = hip_backend.make_params(
        num_rays=N, num_bounces=B, num_bins=K, num_diffuse_samples=128, num_sources=1, rng_mode=_lib.AA_RNG_COUNTER,
        use_bvh=True, render_ambisonics=False, seed=seed, listener=w.listener, source_forward=(1.0, 0.0, 0.0),
        bin_dur=bin_dur, speed_of_sound=C, max_path_len=2.0 * C, listener_radius=0.1, source_radius=0.1,
        irradiance_min_distance=1.0, specular_exponent=100.0, dipole_weight=0.0, dipole_power=1.0)
    stream = torch.cuda.current_stream(0)
    for _ in range(2):
        hip_backend.launch_trace(scene, params, src, bank, buffers, stream=stream, plain=plain)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a_, b_ in ev:
        a_.record(stream)
        hip_backend.launch_trace(scene, params, src, bank, buffers, stream=stream, plain=plain)
        b_.record(stream)
    torch.cuda.synchronize()
    ms = [a_.elapsed_time(b_) for a_, b_ in ev]
    counters = dict(zip(_lib.COUNTER_NAMES, (int(x) for x in buffers.counters.cpu().numpy())))
    # one more (untimed) launch with each ray's start and end stamped on the
    # device's 100 MHz clock: when the rays end, relative to the first start
    buffers.ray_times = torch.zeros((N, 2), dtype=torch.int64, device="cuda:0")
    hip_backend.launch_trace(scene, params, src, bank, buffers, stream=stream, plain=plain)
    torch.cuda.synchronize()
    t = buffers.ray_times.cpu().numpy().astype(np.float64)
    scene.close()
    ends = np.sort(t[:, 1] - t[:, 0].min()) * 1e-5  # ticks of 10 ns -> ms
    ray_ends = {f"p{q:g}": round(float(np.percentile(ends, q)), 3) for q in (50, 90, 99, 99.9)}
    ray_ends.update(last=round(float(ends[-1]), 3), last_start=round(float((t[:, 0].max() - t[:, 0].min()) * 1e-5), 3),
                    rays_stamped=int(np.count_nonzero(t[:, 1] >= t[:, 0])))
    return {"workload": f"furnished 3-room L plan, {w.mesh.num_triangles} triangles, 200k rays, 200 bounces, "
                        "2.0 s @ 1 ms" + (", plain one-thread-per-ray kernel" if plain else ""),
            "kernel_ms_p50": round(float(np.percentile(ms, 50)), 4),
            "rirs_per_s": round(1000.0 / float(np.mean(ms)), 2),
            "primary_rays_per_s": round(N / (float(np.mean(ms)) / 1000.0), 1), "counters": counters,
            "ray_end_ms": ray_ends}


def served_report(rates, seconds: float, seed: int, sequential_rates, plain_rates) -> dict:
    from acoustic_agent_amd.serving import SceneCache, ServingLoop, serve_sequential

    report = {}
    for rate in rates:
        stream = workloads.served_stream(rate, seconds, seed=seed + rate)
        loop = ServingLoop(0, slots=3, max_batch_traces=256)
        loop.warm(stream)
        _, served = loop.run(stream)
        loop.close()
        entry = {"requests": served["requests"], "rirs": served["rirs"], "served": {
            k: served[k] for k in ("rirs_per_s", "latency_all", "latency_by_class", "latency_by_kind")}}
        _progress(f"served {rate}/s: {served['latency_all']} {served['rirs_per_s']} RIRs/s")
        for label, rates_for, plain in (("sequential_single_trace", sequential_rates, False),
                                        ("sequential_plain", plain_rates, True)):
            if rate not in rates_for:
                continue
            cache = SceneCache(0)
            for r in stream:
                cache.release(list(cache.acquire(h for h in (t.scene for t in r.traces)
                                                 if isinstance(h, hip_backend.HostScene))))
            _, seq = serve_sequential(stream, plain=plain, scene_cache=cache)
            cache.close()
            entry[label] = {k: seq[k] for k in ("rirs_per_s", "latency_all", "latency_by_class", "latency_by_kind")}
            _progress(f"{label} {rate}/s: {seq['latency_all']} {seq['rirs_per_s']} RIRs/s")
        report[str(rate)] = entry
    return report


# ---------------------------------------------------------------------------
# the default line: batched mic arrays
# ---------------------------------------------------------------------------


def run_headline(args) -> int:
    import torch

    from acoustic_agent_amd.autotune import Autotuner
    from acoustic_agent_amd.batch import BatchEngine, BatchTuning

    t = time.perf_counter()
    reqs = workloads.mic_array_batch(num_scenes=args.scenes or 256, seed=args.seed, num_bounces=args.bounces)
    reqs.sort(key=lambda r: -r.work)  # the engine's launch order, kept fixed for all submissions
    build_s = time.perf_counter() - t
    _progress(f"{len(reqs)} traces over {len({id(r.scene) for r in reqs})} scenes built in {build_s:.1f} s")
    t = time.perf_counter()
    if args.tuning:
        tuning = BatchTuning(**json.loads(args.tuning))
    elif args.no_tune:
        tuning = BatchTuning()
    else:
        tuning = Autotuner(0).tune_batch(reqs, sample=64, reps=1, progress=_progress)
    tune_s = time.perf_counter() - t
    engine = BatchEngine(0, max_traces=len(reqs))
    wall, ms, results = time_batch(engine, reqs, tuning, args.steps, args.warmup)
    engine.close()
    ms_per_step = wall * 1000.0 / args.steps
    rirs = sum(r.num_sources for r in reqs)
    rays = sum(r.num_rays for r in reqs)
    totals = {k: sum(res["counters"][k] for res in results) for k in (
        "bounce_queries", "shadow_queries", "nodes_bounce", "nodes_shadow", "contributions")}
    kernel_s = statistics.mean(ms) / 1000.0
    _progress(f"batch: {ms_per_step:.2f} ms per step, {rirs / (ms_per_step / 1000.0):.1f} RIRs/s")
    tri = [r.scene.info["num_triangles"] for r in reqs]
    result = {
        "metric": METRIC,
        "value": round(rirs / (ms_per_step / 1000.0), 3),
        "unit": "RIRs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scenes, materials and counter-RNG ray sets)",
        "config": {
            "workload": f"batched mic arrays: {len({id(r.scene) for r in reqs})} scenes (triangles log-uniform "
                        f"12-50k, shoeboxes and furnished multi-room plans) x 4-mic ring (5 cm) = {len(reqs)} "
                        f"traces, 20k-200k rays, {args.bounces} bounces, 2.0 s @ 1 ms x 6 bands, one persistent "
                        "launch",
            "traces": len(reqs), "rirs": rirs, "rays": rays, "triangles_min_max": [min(tri), max(tri)],
            "bounces": args.bounces, "bins": reqs[0].num_bins, "rng": "counter (Philox4x32-10)",
            "tuning": dataclasses.asdict(tuning), "autotune_s": round(tune_s, 2), "scene_build_s": round(build_s, 2),
        },
        "primary_rays_per_s": round(rays / (ms_per_step / 1000.0), 1),
        "segment_queries_per_s": round((totals["bounce_queries"] + totals["shadow_queries"]) / kernel_s, 1),
        "bvh4_nodes_per_s": round((totals["nodes_bounce"] + totals["nodes_shadow"]) / kernel_s, 1),
        "latency_ms": {"p50": round(float(np.percentile(ms, 50)), 3), "p99": round(float(np.percentile(ms, 99)), 3)},
        "kernel_ms_avg": round(kernel_s * 1000.0, 3),
        "gpu_counters": totals,
    }

    roofline = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None}
    traffic = _traffic("batch_kernel_c2_traffic.json")
    if traffic:
        roofline["traffic"] = traffic.get("hbm_bytes_per_launch")
        roofline["measured"] = {k: traffic[k] for k in ("hbm_gbs", "hbm_frac", "valu_busy", "valu_issue_frac", "limiter")
                                if k in traffic}
        roofline["measured"]["source"] = "profiles/batch_kernel_c2_traffic.json (rocprofv3 PMC of this kernel and workload)"
        roofline["limiter"] = "valu-issue" if traffic.get("valu_busy", 0.0) > traffic.get("hbm_frac", 0.0) else "hbm"
    valid = True
    if not args.no_cpu_baseline:
        leg = cpu_baseline_leg(reqs, results, sample=args.cpu_sample, seed=args.seed + 1, counts_rays=256,
                               accuracy_traces=4)
        alg = algorithmic_bytes(reqs, results, leg["canonical_counts"])
        achieved = alg / kernel_s / 1e9
        roofline.update(achieved=round(achieved, 1), frac=round(achieved / HBM_PEAK_GBS, 4),
                        algorithmic_bytes_per_launch=alg,
                        algorithmic_note="the reference's binary-BVH traversal bytes (canonical counts of the "
                                         "binary64 oracle); the scenes' working set is served from L2/MALL, so "
                                         "they can exceed what DRAM delivers: see measured")
        allc, one = leg["all_cores"], leg["one_thread"]
        result["cpu_baseline"] = {
            "value": allc["rirs_per_s"], "unit": "RIRs/s", "cores": leg["cores"], "kind": "port",
            "sample": f"{len(leg['picks'])} seeded traces of this workload, each over 8 evenly spaced ray ranges "
                      f"covering 1/8 of its rays (binary64 oracle, OpenMP over rays), scaled to whole traces; "
                      f"median of 3 after a warm-up; CPU {cpu_model()}",
            "rirs_per_s_with_rt60": allc["rirs_per_s_with_rt60"],
            "one_thread": {"rirs_per_s": one["rirs_per_s"], "rirs_per_s_with_rt60": one["rirs_per_s_with_rt60"],
                           "sample": "the same traces over 1/128 of their rays"},
            "seconds_per_pass": allc["seconds_per_pass"],
        }
        result["speedup_vs_cpu"] = round(result["value"] / allc["rirs_per_s"], 1)
        result["speedup_vs_cpu_one_thread"] = round(result["value"] / one["rirs_per_s"], 1)
        worst = max(a["rt60_max_rel_err"] or 0.0 for a in leg["accuracy"])
        worst_edt = max(a["edt_max_rel_err"] or 0.0 for a in leg["accuracy"])
        result["accuracy_vs_fp64_oracle"] = {
            "traces": leg["accuracy"], "rt60_max_rel_err": worst, "edt_max_rel_err": worst_edt,
            "rt60_within_2pct": worst <= RT60_TOLERANCE, "edt_within_2pct": worst_edt <= RT60_TOLERANCE}
        valid = worst <= RT60_TOLERANCE and worst_edt <= RT60_TOLERANCE
    else:
        result["accuracy_vs_fp64_oracle"] = None
    result["roofline"] = roofline
    result["valid"] = valid

    if not args.no_extras:
        base = sample_baselines(reqs, args.plain_sample, args.seed + 2)
        base["speedup_vs_plain"] = round(result["value"] / base["plain_rirs_per_s"], 2)
        base["speedup_vs_single_trace"] = round(result["value"] / base["single_trace_rirs_per_s"], 2)
        base["single_scene_trace"] = c1_single_trace(10, args.seed)
        _progress(f"single-scene trace: {base['single_scene_trace']['rirs_per_s']} RIRs/s")
        result["baselines"] = base
        result["served"] = served_report((50, 100, 200, 400), args.serve_seconds, args.seed,
                                         sequential_rates=(50, 100, 200, 400), plain_rates=(50,))
        result["served_config"] = {"arrival_seconds": args.serve_seconds, "slots": 3, "hbm_budget_gb": 64,
                                   "mix": dict(zip(workloads.SERVED_KINDS, workloads.SERVED_MIX)),
                                   "latency": "arrival -> echogram, counters and band RT60 on the host"}
    torch.cuda.synchronize()
    print(json.dumps(result), flush=True)
    return 0 if valid else 3


# ---------------------------------------------------------------------------
# other modes
# ---------------------------------------------------------------------------


def run_keyframes(args) -> int:
    from acoustic_agent_amd.autotune import Autotuner
    from acoustic_agent_amd.batch import BatchEngine, BatchTuning

    reqs = workloads.keyframe_batch(num_scenes=args.scenes or 32, seed=args.seed, num_bounces=args.bounces)
    reqs.sort(key=lambda r: -r.work)
    if args.tuning:
        tuning = BatchTuning(**json.loads(args.tuning))
    elif args.no_tune:
        tuning = BatchTuning()
    else:
        tuning = Autotuner(0).tune_batch(reqs, sample=4, reps=1, progress=_progress)
    tuning.use_graph = True
    engine = BatchEngine(0, max_traces=len(reqs))
    wall, ms, results = time_batch(engine, reqs, tuning, args.steps, args.warmup)
    engine.close()
    rirs = sum(r.num_sources for r in reqs)
    out = {"workload": f"{len(reqs)} furnished L plans x 64-keyframe moving source (one 64-source trace per "
                       f"scene), 100k rays, {args.bounces} bounces, 1.0 s @ 1 ms, hipGraph replay",
           "rirs": rirs, "batch_s_per_step": round(wall / args.steps, 4),
           "rirs_per_s": round(rirs / (wall / args.steps), 2), "kernel_ms_p50": round(float(np.percentile(ms, 50)), 2),
           "tuning": dataclasses.asdict(tuning)}
    if not args.no_extras:
        base = sample_baselines(reqs, args.plain_sample, args.seed + 2)
        out.update(base, speedup_vs_plain=round(out["rirs_per_s"] / base["plain_rirs_per_s"], 2),
                   speedup_vs_single_trace=round(out["rirs_per_s"] / base["single_trace_rirs_per_s"], 2))
    print(json.dumps({"metric": "RIRs/s (batched 64-keyframe motion)", "unit": "RIRs/s", "value": out["rirs_per_s"],
                      "dtype": "f32", "data": "synthetic", "config": out}), flush=True)
    return 0


def run_single(args) -> int:
    out = c1_single_trace(args.steps, args.seed, plain=args.plain)
    print(json.dumps({"metric": "RIRs/s (single-scene trace)", "unit": "RIRs/s", "value": out["rirs_per_s"],
                      "dtype": "f32", "data": "synthetic", "config": out}), flush=True)
    return 0


def run_served(args) -> int:
    report = served_report((50, 100, 200, 400), args.serve_seconds, args.seed, sequential_rates=(50, 100, 200, 400),
                           plain_rates=(50, 100))
    print(json.dumps({"metric": "p50/p99 served latency (continuous batching vs sequential plain / single-trace)",
                      "unit": "ms", "dtype": "f32", "data": "synthetic Poisson arrivals",
                      "config": {"arrival_seconds": args.serve_seconds, "slots": 3, "hbm_budget_gb": 64},
                      "rates": report}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("c2", "c1", "c3", "c4"), default="c2",
                    help="c2: batched mic arrays (the default line); c1: single-scene trace; c3: batched "
                         "moving-source keyframes; c4: served Poisson stream")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scenes", type=int, default=0)
    ap.add_argument("--bounces", type=int, default=200)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--plain-sample", type=int, default=32, help="traces timed one per launch (plain / single)")
    ap.add_argument("--cpu-sample", type=int, default=8, help="traces in the CPU baseline sample")
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--plain", action="store_true", help="c1 only: the plain one-thread-per-ray kernel on the same trace")
    ap.add_argument("--tuning", type=str, default="", help="JSON BatchTuning fields (skips the autotuner)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the baselines and the served stream")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 256 MiB device-to-device three times (calibrates rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE against known bytes) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build (traces are independent "
              "and would shard by trace across GPUs)", file=sys.stderr)
        return 2

    import torch

    if args.copy_probe:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        src.fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe": {"bytes_read": src.numel() * 4, "bytes_written": src.numel() * 4,
                                         "copies": 3}}), flush=True)
        return 0
    if not hip_backend.hip_available(0):
        print("no HIP device", file=sys.stderr)
        return 1
    return {"c2": run_headline, "c1": run_single, "c3": run_keyframes, "c4": run_served}[args.workload](args)


if __name__ == "__main__":
    sys.exit(main())
