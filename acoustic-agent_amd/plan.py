"""Receivers and motion frames -> trace descriptors for the batch engine.

The reference solves an array receiver as one mono simulation per channel
position (engine.py:64-73) and a motion as one simulation per
frame with the same configuration and seed (api.py:669-713).  Here both
become ``TraceRequest`` descriptors for one persistent batch:

* ``channel_traces``: one trace per channel position (the same seed, so the
  same ray set), listener = the channel position.
* ``motion_traces``: frames are grouped by receiver position; within a group
  every frame shares listener, configuration and seed, so its ray
  trajectories do not depend on the source (steam_rt.py:4750-4849) and the
  group's sources (up to 64) are one multi-source trace whose per-source
  results equal the frame-by-frame traces.  ``MotionPlan.frames`` maps each
  frame and channel back to (trace, source).

Every planned trace carries the configuration ``trace_energy_field`` would
use for the same call: rays, bounces, bin count (``histogram_bins``), the
diffuse bank size, the intersection backend, the radii and exponents, the
source directivity (weight, power and forward axis of ``source_model``) and
the ambisonic echogram.  Batched traces draw their rays on the device, so the
configuration must use ``rt_rng="counter"``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Mapping, Sequence

import numpy as np

from .batch import TraceRequest
from .directivity import SourceModel
from .mic import channel_positions
from .models import SimConfig
from .steam_rt import _resolve_intersection_backend

MAX_SOURCES_PER_TRACE = 64  # AA_MAX_SOURCES


def _request(scene, listener, sources, config: SimConfig, tag: Any, source_model=None,
             render_ambisonics: bool = True) -> TraceRequest:
    if str(config.rt_rng).lower() != "counter":
        raise ValueError("batched traces draw their rays on the device: use rt_rng='counter'")
    emitter = SourceModel.parse(source_model)
    surfaces = getattr(scene, "num_surfaces", None)
    use_bvh = True if surfaces is None else _resolve_intersection_backend(
        config.intersection_backend, int(surfaces), config.bvh_min_surfaces) == "bvh"
    return TraceRequest(
        scene=scene, listener=np.asarray(listener, dtype=np.float64), sources=np.asarray(sources, dtype=np.float64),
        num_rays=int(config.rt_num_rays), num_bounces=int(config.rt_num_bounces),
        duration_s=float(max(config.rt_duration_s, config.duration_s)), bin_dur=float(config.rt_bin_duration_s),
        seed=int(config.seed), c=float(config.c), listener_radius=float(config.rt_listener_radius),
        source_radius=float(config.rt_source_radius), irradiance_min_distance=float(config.rt_irradiance_min_distance),
        specular_exponent=float(config.rt_specular_exponent),
        source_forward=tuple(float(v) for v in emitter.forward()), dipole_weight=emitter.weight,
        dipole_power=emitter.power, render_ambisonics=bool(render_ambisonics),
        num_diffuse_samples=int(config.rt_num_diffuse_samples), use_bvh=use_bvh, tag=tag)


def channel_traces(scene, source: Sequence[float], receiver: Sequence[float], receiver_model: Mapping[str, Any] | None,
                   config: SimConfig, source_model=None, render_ambisonics: bool = True) -> list[TraceRequest]:
    """One trace per channel of ``receiver_model`` at ``receiver``."""
    src = np.asarray(source, dtype=np.float64).reshape(1, 3)
    return [_request(scene, pos, src, config, ("channel", c), source_model, render_ambisonics) for c, pos in
            enumerate(channel_positions(receiver, receiver_model))]


@dataclass
class MotionPlan:
    traces: list[TraceRequest]
    # frames[f][c] = (trace index, source index) of frame f, channel c
    frames: list[list[tuple[int, int]]] = field(default_factory=list)


def motion_traces(scene, frames: Sequence[Mapping[str, Any]], config: SimConfig,
                  receiver_model: Mapping[str, Any] | None = None, source_model=None,
                  render_ambisonics: bool = True) -> MotionPlan:
    """Multi-source traces for a motion ``{"source", "receiver"}`` frame list."""
    if not frames:
        raise ValueError("motion must contain at least one frame")
    groups: dict[tuple, list[int]] = {}  # receiver position -> frame indices, in order
    for f, frame in enumerate(frames):
        if not isinstance(frame, Mapping) or "source" not in frame or "receiver" not in frame:
            raise ValueError("each motion frame must be an object with a source and a receiver")
        key = tuple(float(v) for v in np.asarray(frame["receiver"], dtype=float)[:3])
        groups.setdefault(key, []).append(f)
    plan = MotionPlan(traces=[], frames=[[] for _ in frames])
    for receiver, members in groups.items():
        positions = channel_positions(receiver, receiver_model)
        for start in range(0, len(members), MAX_SOURCES_PER_TRACE):
            chunk = members[start:start + MAX_SOURCES_PER_TRACE]
            sources = np.stack([np.asarray(frames[f]["source"], dtype=np.float64)[:3] for f in chunk])
            for c, pos in enumerate(positions):
                plan.traces.append(_request(scene, pos, sources, config, ("motion", chunk[0], c), source_model,
                                            render_ambisonics))
                t = len(plan.traces) - 1
                for k, f in enumerate(chunk):
                    plan.frames[f].append((t, k))
    return plan
