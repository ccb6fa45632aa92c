"""Receiver models and their channel positions.

Mirrors acoustic_agent/mic.py:11-80: ``microphone_array(kind, **kwargs)``
describes a mono, HRTF, linear or circular receiver (same kinds, defaults and
clamping: linear 2..64 channels at 0.005..2 m spacing, default 4 at 8 cm;
circular 3..128 channels of radius 0.005..5 m, default 8 at 12 cm; offsets
rounded to 1 um), and ``channel_positions(receiver, model)`` gives the
listener position of every channel (one position for mono and HRTF).  Each
position is one listener-origin trace, as the reference traces each array
channel on its own (engine.py:64-73).
"""
from __future__ import annotations

import math
from typing import Any, Mapping, Sequence

import numpy as np

SUPPORTED_ARRAYS = {"mono", "hrtf", "linear", "linear_array", "circular", "circular_array"}


def _linear_channels(count: int, spacing_m: float, orientation_deg: float) -> list[dict[str, Any]]:
    a = math.radians(orientation_deg)
    ux, uy = math.cos(a), math.sin(a)
    mid = 0.5 * (count - 1)
    return [{"id": f"M{i + 1}",
             "offset": [round(ux * (i - mid) * spacing_m, 6), round(uy * (i - mid) * spacing_m, 6), 0.0]}
            for i in range(count)]


def _circular_channels(count: int, radius_m: float, orientation_deg: float) -> list[dict[str, Any]]:
    a0 = math.radians(orientation_deg)
    out = []
    for i in range(count):
        a = a0 + 2.0 * math.pi * i / count
        out.append({"id": f"M{i + 1}", "offset": [round(math.cos(a) * radius_m, 6), round(math.sin(a) * radius_m, 6), 0.0]})
    return out


def microphone_array(kind: str = "mono", **kwargs: Any) -> dict[str, Any]:
    """Receiver model dictionary (mic.py:11-35)."""
    k = str(kind).lower()
    if k not in SUPPORTED_ARRAYS:
        raise ValueError(f"unsupported microphone model: {kind!r}")
    orientation = float(kwargs.get("orientation_deg", 0.0))
    if k == "hrtf":
        return {
            "type": "hrtf",
            "orientation_deg": orientation,
            "interpolation": str(kwargs.get("interpolation", "bilinear")),
            "spatial_blend": float(np.clip(float(kwargs.get("spatial_blend", 1.0)), 0.0, 1.0)),
            "loudness_normalization": str(kwargs.get("loudness_normalization", "energy")),
            "sofa_path": kwargs.get("sofa_path"),
            "channels": [{"id": "L", "offset": [0.0, 0.0, 0.0]}, {"id": "R", "offset": [0.0, 0.0, 0.0]}],
        }
    if k in ("linear", "linear_array"):
        count = int(np.clip(int(kwargs.get("count", 4)), 2, 64))
        spacing = float(np.clip(float(kwargs.get("spacing_m", 0.08)), 0.005, 2.0))
        return {"type": "linear_array", "orientation_deg": orientation,
                "channels": _linear_channels(count, spacing, orientation)}
    if k in ("circular", "circular_array"):
        count = int(np.clip(int(kwargs.get("count", 8)), 3, 128))
        radius = float(np.clip(float(kwargs.get("radius_m", 0.12)), 0.005, 5.0))
        return {"type": "circular_array", "orientation_deg": orientation,
                "channels": _circular_channels(count, radius, orientation)}
    return {"type": "mono", "orientation_deg": orientation, "channels": [{"id": "M", "offset": [0.0, 0.0, 0.0]}]}


def channel_positions(receiver: Sequence[float], model: Mapping[str, Any] | None) -> list[tuple[float, float, float]]:
    """Listener position of every channel of ``model`` at ``receiver`` (mic.py:38-49)."""
    m = model or microphone_array("mono")
    centre = np.asarray(receiver, dtype=float)[:3]
    if str(m.get("type")) not in ("linear_array", "circular_array"):
        return [tuple(float(v) for v in centre)]
    out = []
    for ch in m.get("channels", []):
        p = centre + np.asarray(ch.get("offset", (0.0, 0.0, 0.0)), dtype=float)[:3]
        out.append(tuple(float(v) for v in p))
    return out
