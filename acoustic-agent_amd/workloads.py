"""Seeded request sets of the batched, keyframe and served benchmark workloads.

* ``mic_array_batch``: scenes with triangle counts log-uniform in [12, 50k]
  (shoeboxes and furnished multi-room plans), one source and a 4-microphone
  circular array (5 cm) each, rays uniform in [20k, 200k], 2.0 s at 1 ms.
  Every microphone channel is its own trace with the scene's seed, as the
  reference traces each channel of an array (engine.py:64-73).
* ``keyframe_batch``: furnished L plans (~20k triangles) with a source moving
  along a seeded 64-keyframe spline at 0.5-1.5 m/s and a fixed listener; all
  keyframes of a scene share one ray set (the reference reuses config and seed
  for every keyframe, api.py:683-696), so they are one multi-source trace.
* ``served_stream``: Poisson arrivals of requests drawn from the four
  workload kinds, over a pool of host scenes (``HostScene``, keyed by content
  hash: the serving loop uploads each once and reuses it):
    - ``shoebox``: the 5 x 4 x 3 m shoebox (12 triangles), 10k rays, 50
      bounces, 1.0 s;
    - ``single``: one trace of a scene of the 12-50k triangle mix, 10k-200k
      rays, 200 bounces, 2.0 s;
    - ``mic_array``: the same with a 1-8 microphone ring (5 cm), one trace
      per channel;
    - ``keyframes``: a furnished L plan (~20k triangles) with its seeded
      64-keyframe source spline, one 64-source trace of 100k rays, 200
      bounces, 1.0 s.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import scenes
from .batch import TraceRequest
from .hip_backend import DeviceScene, HostScene


def device_scene(w: scenes.Workload, device: int = 0) -> DeviceScene:
    """Upload ``w``'s triangles; the scene keeps ``w`` as ``workload`` (its
    host description, e.g. for a CPU check of the same scene)."""
    a = w.scene.kernel_arrays()
    ds = DeviceScene(w.mesh, np.asarray(a["reflection"]), np.asarray(a["scattering"]), device=device)
    ds.workload = w
    return ds


def host_scene(w: scenes.Workload) -> HostScene:
    """``w``'s triangles on the host, keyed by content hash (keeps ``w``)."""
    a = w.scene.kernel_arrays()
    return HostScene(w.mesh, np.asarray(a["reflection"]), np.asarray(a["scattering"]), workload=w)


def log_uniform_triangles(rng: np.random.Generator, lo: int = 12, hi: int = 50_000) -> int:
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def mic_array_batch(num_scenes: int = 256, seed: int = 0, num_bounces: int = 200, rays=(20_000, 200_000),
                    device: int = 0, num_mics: int = 4) -> list[TraceRequest]:
    rng = np.random.default_rng(seed)
    out: list[TraceRequest] = []
    for i in range(num_scenes):
        tri = log_uniform_triangles(rng)
        s_seed = int(rng.integers(0, 1 << 31))
        w, mics = scenes.mic_array_workload(s_seed, tri, num_mics=num_mics)
        ds = device_scene(w, device)
        n = int(rng.integers(rays[0], rays[1] + 1))
        for m in range(num_mics):
            out.append(TraceRequest(scene=ds, listener=mics[m], sources=w.source[None, :], num_rays=n,
                                    num_bounces=num_bounces, duration_s=2.0, seed=s_seed,
                                    tag=("mic", i, m, w.mesh.num_triangles)))
    return out


def keyframe_batch(num_scenes: int = 32, seed: int = 0, keyframes: int = 64, num_rays: int = 100_000,
                   num_bounces: int = 200, device: int = 0) -> list[TraceRequest]:
    rng = np.random.default_rng(seed)
    out: list[TraceRequest] = []
    for i in range(num_scenes):
        s_seed = int(rng.integers(0, 1 << 31))
        w, path = scenes.keyframe_workload(s_seed, 20_000, keyframes)
        ds = device_scene(w, device)
        out.append(TraceRequest(scene=ds, listener=w.listener, sources=path, num_rays=num_rays,
                                num_bounces=num_bounces, duration_s=1.0, seed=s_seed,
                                tag=("keyframes", i, keyframes, w.mesh.num_triangles)))
    return out


@dataclass
class ServedRequest:
    arrival_s: float
    size_class: str
    traces: list[TraceRequest]
    kind: str = "single"


SERVED_KINDS = ("shoebox", "single", "mic_array", "keyframes")
SERVED_MIX = (0.3, 0.3, 0.25, 0.15)


def served_stream(rate_hz: float, duration_s: float, seed: int = 0, num_scenes: int = 12,
                  num_keyframe_scenes: int = 2, max_rays: int = 200_000) -> list[ServedRequest]:
    """Poisson arrivals at ``rate_hz`` over ``duration_s`` of mixed solves."""
    rng = np.random.default_rng(seed)
    shoebox = scenes.shoebox_workload(seed=int(rng.integers(0, 1 << 31)))
    shoebox_scene = host_scene(shoebox)
    pool = []
    for _ in range(num_scenes):
        w = scenes.random_scene(int(rng.integers(0, 1 << 31)), log_uniform_triangles(rng, 12, 50_000))
        pool.append((w, host_scene(w)))
    moving = []
    for _ in range(num_keyframe_scenes):
        w, path = scenes.keyframe_workload(int(rng.integers(0, 1 << 31)), 20_000, 64)
        moving.append((w, path, host_scene(w)))
    t = 0.0
    out: list[ServedRequest] = []
    while True:
        t += float(rng.exponential(1.0 / rate_hz))
        if t > duration_s:
            break
        kind = str(rng.choice(SERVED_KINDS, p=SERVED_MIX))
        req_seed = int(rng.integers(0, 1 << 31))
        if kind == "shoebox":
            w, hs = shoebox, shoebox_scene
            traces = [TraceRequest(scene=hs, listener=w.listener, sources=w.source[None, :], num_rays=10_000,
                                   num_bounces=50, duration_s=1.0, seed=req_seed)]
        elif kind == "keyframes":
            w, path, hs = moving[int(rng.integers(0, len(moving)))]
            traces = [TraceRequest(scene=hs, listener=w.listener, sources=path, num_rays=100_000, num_bounces=200,
                                   duration_s=1.0, seed=req_seed)]
        else:
            w, hs = pool[int(rng.integers(0, len(pool)))]
            rays = int(rng.integers(10_000, max_rays + 1))
            count = int(rng.integers(1, 9)) if kind == "mic_array" else 1
            listeners = scenes.circular_array(w.listener, count, 0.05) if kind == "mic_array" else [w.listener]
            traces = [TraceRequest(scene=hs, listener=m, sources=w.source[None, :], num_rays=rays, num_bounces=200,
                                   duration_s=2.0, seed=req_seed) for m in listeners]
        tri = w.mesh.num_triangles
        size = "small" if tri < 1000 else ("medium" if tri < 10_000 else "large")
        out.append(ServedRequest(arrival_s=t, size_class=size, traces=traces, kind=kind))
    return out
