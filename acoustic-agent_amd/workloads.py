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
* ``served_stream``: Poisson arrivals of requests drawn from the shoebox,
  single-scene, mic-array and keyframe mixes, over a pool of scenes whose
  BVHs stay resident and are reused across requests.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import scenes
from .batch import TraceRequest
from .hip_backend import DeviceScene


def device_scene(w: scenes.Workload, device: int = 0) -> DeviceScene:
    """Upload ``w``'s triangles; the scene keeps ``w`` as ``workload`` (its
    host description, e.g. for a CPU check of the same scene)."""
    a = w.scene.kernel_arrays()
    ds = DeviceScene(w.mesh, np.asarray(a["reflection"]), np.asarray(a["scattering"]), device=device)
    ds.workload = w
    return ds


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


def served_stream(rate_hz: float, duration_s: float, seed: int = 0, num_scenes: int = 12,
                  device: int = 0, max_rays: int = 200_000) -> list[ServedRequest]:
    """Poisson arrivals at ``rate_hz`` over ``duration_s`` of mixed solves."""
    rng = np.random.default_rng(seed)
    pool = []
    for i in range(num_scenes):
        tri = log_uniform_triangles(rng, 12, 50_000)
        w = scenes.random_scene(int(rng.integers(0, 1 << 31)), tri)
        pool.append((w, device_scene(w, device)))
    t = 0.0
    out: list[ServedRequest] = []
    while True:
        t += float(rng.exponential(1.0 / rate_hz))
        if t > duration_s:
            break
        w, ds = pool[int(rng.integers(0, len(pool)))]
        kind = rng.choice(["shoebox_small", "single", "mic_array", "keyframes"], p=[0.3, 0.3, 0.25, 0.15])
        rays = int(rng.integers(10_000, max_rays + 1))
        req_seed = int(rng.integers(0, 1 << 31))
        if kind == "shoebox_small":
            rays = 10_000
            traces = [TraceRequest(scene=ds, listener=w.listener, sources=w.source[None, :], num_rays=rays,
                                   num_bounces=50, duration_s=1.0, seed=req_seed)]
        elif kind == "single":
            traces = [TraceRequest(scene=ds, listener=w.listener, sources=w.source[None, :], num_rays=rays,
                                   num_bounces=200, duration_s=2.0, seed=req_seed)]
        elif kind == "mic_array":
            count = int(rng.integers(1, 9))
            mics = scenes.circular_array(w.listener, count, 0.05)
            traces = [TraceRequest(scene=ds, listener=m, sources=w.source[None, :], num_rays=rays, num_bounces=200,
                                   duration_s=2.0, seed=req_seed) for m in mics]
        else:
            path = w.source[None, :] + np.linspace(0.0, 0.5, 16)[:, None] * np.asarray([1.0, 0.3, 0.0])[None, :]
            traces = [TraceRequest(scene=ds, listener=w.listener, sources=path, num_rays=rays, num_bounces=200,
                                   duration_s=1.0, seed=req_seed)]
        size = "small" if w.mesh.num_triangles < 1000 else ("medium" if w.mesh.num_triangles < 10_000 else "large")
        out.append(ServedRequest(arrival_s=t, size_class=f"{size}", traces=traces))
    return out
