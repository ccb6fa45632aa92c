"""Receiver layouts: where each channel of a receiver listens.

The tracer is listener-centric, so a receiver enters it only as the list of
listener positions it needs: one per channel of a linear or circular array
(the reference traces each array channel on its own, engine.py:64-73), one
for a mono or HRTF receiver (both ears share the centre; binaural rendering
happens after tracing and is not part of this package).

Receiver models are dictionaries (``type``, ``orientation_deg``,
``channels`` = ``[{"id", "offset"}]``) so that models built by the
reference's ``microphone_array`` (acoustic_agent/mic.py:11-35) can be passed
in unchanged.  ``microphone_array`` builds them from ``LAYOUTS``: per kind,
the parameter that sets the channel count and the one that sets the size,
each with its default and clamping range (the reference's values), and the
offset generator.  Offsets lie in the horizontal plane, rotated by the
orientation, and are rounded to 1 um.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Any, Callable, Mapping, Sequence

import numpy as np


def _linear_offsets(count: int, spacing: float, yaw: float) -> np.ndarray:
    # channels centred on the receiver along the orientation axis
    along = (np.arange(count) - 0.5 * (count - 1)) * spacing
    return np.stack([along * math.cos(yaw), along * math.sin(yaw), np.zeros(count)], axis=1)


def _circular_offsets(count: int, radius: float, yaw: float) -> np.ndarray:
    # channel 0 on the orientation axis, then counter-clockwise
    angle = yaw + 2.0 * math.pi * np.arange(count) / count
    return np.stack([radius * np.cos(angle), radius * np.sin(angle), np.zeros(count)], axis=1)


@dataclass(frozen=True)
class Layout:
    kind: str
    count: tuple[str, int, int, int]           # keyword, default, min, max
    size: tuple[str, float, float, float]      # keyword, default, min, max
    offsets: Callable[[int, float, float], np.ndarray]


LAYOUTS: dict[str, Layout] = {
    "linear_array": Layout("linear_array", ("count", 4, 2, 64), ("spacing_m", 0.08, 0.005, 2.0), _linear_offsets),
    "circular_array": Layout("circular_array", ("count", 8, 3, 128), ("radius_m", 0.12, 0.005, 5.0),
                             _circular_offsets),
}
# receivers with one listening position: channel ids at the centre
CENTRED: dict[str, tuple[str, ...]] = {"mono": ("M",), "hrtf": ("L", "R")}
ALIASES = {"linear": "linear_array", "circular": "circular_array"}


def microphone_array(kind: str = "mono", **kwargs: Any) -> dict[str, Any]:
    """Receiver model dictionary of ``kind`` (mono, hrtf, linear[_array],
    circular[_array]); array parameters are clamped to their ranges."""
    name = ALIASES.get(str(kind).lower(), str(kind).lower())
    yaw_deg = float(kwargs.get("orientation_deg", 0.0))
    if name in CENTRED:
        channels = [{"id": cid, "offset": [0.0, 0.0, 0.0]} for cid in CENTRED[name]]
        return {"type": name, "orientation_deg": yaw_deg, "channels": channels}
    layout = LAYOUTS.get(name)
    if layout is None:
        raise ValueError(f"unsupported microphone model: {kind!r}")
    key, default, lo, hi = layout.count
    count = min(hi, max(lo, int(kwargs.get(key, default))))
    key, default, flo, fhi = layout.size
    size = min(fhi, max(flo, float(kwargs.get(key, default))))
    offsets = np.round(layout.offsets(count, size, math.radians(yaw_deg)), 6) + 0.0  # + 0.0: no negative zeros
    channels = [{"id": f"M{i + 1}", "offset": [float(v) for v in row]} for i, row in enumerate(offsets)]
    return {"type": name, "orientation_deg": yaw_deg, "channels": channels}


def channel_positions(receiver: Sequence[float], model: Mapping[str, Any] | None) -> list[tuple[float, float, float]]:
    """Listener position of every traced channel of ``model`` placed at
    ``receiver`` (the reference's mic.py:38-49): centre + offset for array
    channels, the centre alone for any other receiver."""
    centre = np.asarray(receiver, dtype=np.float64)[:3]
    if model is None or str(model.get("type")) not in LAYOUTS:
        return [tuple(float(v) for v in centre)]
    return [tuple(float(v) for v in centre + np.asarray(ch.get("offset", (0.0, 0.0, 0.0)), dtype=np.float64)[:3])
            for ch in model.get("channels", [])]


def circular_array(centre: Sequence[float], count: int = 4, radius: float = 0.05,
                   orientation_deg: float = 0.0) -> np.ndarray:
    """[count, 3] channel positions of a circular array at ``centre``."""
    model = microphone_array("circular", count=count, radius_m=radius, orientation_deg=orientation_deg)
    return np.asarray(channel_positions(centre, model))
