"""Source directivity of the next-event deposit: a weighted dipole.

A connection from a hit to the source carries the gain
``|(1 - w) + w cos|^p`` with ``cos = -(forward . towards_source)``
(steam_rt.py:4803-4804).  Callers describe the source with the dictionary
the reference's API accepts (acoustic_agent/directivity.py:9-67): a preset
name, or ``{"type": ..., "pattern": ..., "dipole_weight": ...,
"dipole_power": ..., "orientation_deg": ..., "elevation_deg": ...}``.

``SourceModel.parse`` resolves such a description in three table lookups:
  1. the type (``type``, else ``pattern``, else ``"omni"``; case and
     hyphens ignored) is either a preset, which fixes the pattern, or one of
     the generic types, which take the pattern from ``pattern`` (a pattern
     that is not a preset is ``"custom"``);
  2. the pattern's preset (custom: the cardioid's) supplies the defaults of
     weight and power;
  3. every numeric field is read, then checked against ``_CHECKS``.
``as_dict`` gives the normalised dictionary (the keys the field dictionary
reports as ``source_directivity``); ``forward`` the unit axis from yaw
(orientation) and pitch (elevation, clamped to +-90 degrees).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Any, Mapping

import numpy as np

# pattern -> (dipole weight w, dipole power p)
PRESETS: dict[str, tuple[float, float]] = {
    "omni": (0.0, 1.0),
    "cardioid": (0.5, 1.0),
    "dipole": (1.0, 1.0),
    "focused": (0.5, 4.0),
}
GENERIC_TYPES = frozenset({"weighted_dipole", "directivity"})
_CUSTOM_DEFAULTS = "cardioid"

# (predicate over the parsed fields, message of the ValueError)
_CHECKS = (
    (lambda f: all(math.isfinite(v) for v in f.values()), "source directivity parameters must be finite"),
    (lambda f: 0.0 <= f["weight"] <= 1.0, "dipole_weight must be between 0 and 1"),
    (lambda f: f["power"] >= 0.0, "dipole_power must be non-negative"),
)


def _name(value: Any) -> str:
    return str(value).strip().lower().replace("-", "_")


@dataclass(frozen=True)
class SourceModel:
    pattern: str
    weight: float
    power: float
    yaw_deg: float
    pitch_deg: float

    @classmethod
    def parse(cls, value: str | Mapping[str, Any] | None) -> "SourceModel":
        spec: Mapping[str, Any] = {} if value is None else ({"type": value} if isinstance(value, str) else value)
        kind = _name(spec.get("type", spec.get("pattern", "omni")))
        if kind in PRESETS:
            pattern = kind
        elif kind in GENERIC_TYPES:
            pattern = _name(spec.get("pattern", kind))
            pattern = pattern if pattern in PRESETS else "custom"
        else:
            raise ValueError(f"unknown source directivity {kind!r}; expected one of "
                             f"{', '.join(PRESETS)} or weighted_dipole")
        weight0, power0 = PRESETS.get(pattern, PRESETS[_CUSTOM_DEFAULTS])
        fields = {
            "weight": float(spec.get("dipole_weight", weight0)),
            "power": float(spec.get("dipole_power", power0)),
            "yaw": float(spec.get("orientation_deg", 0.0)),
            "pitch": float(spec.get("elevation_deg", 0.0)),
        }
        for ok, message in _CHECKS:
            if not ok(fields):
                raise ValueError(message)
        return cls(pattern, fields["weight"], fields["power"], fields["yaw"], min(90.0, max(-90.0, fields["pitch"])))

    @property
    def is_omni(self) -> bool:
        return self.weight == 0.0

    def forward(self) -> np.ndarray:
        yaw, pitch = math.radians(self.yaw_deg), math.radians(self.pitch_deg)
        return np.asarray([math.cos(pitch) * math.cos(yaw), math.cos(pitch) * math.sin(yaw), math.sin(pitch)])

    def as_dict(self) -> dict[str, Any]:
        return {"type": "weighted_dipole", "pattern": self.pattern, "orientation_deg": self.yaw_deg,
                "elevation_deg": self.pitch_deg, "dipole_weight": self.weight, "dipole_power": self.power}


def source_directivity(value: str | Mapping[str, Any] | None = None) -> dict[str, Any]:
    """The normalised source dictionary (the reference's
    ``source_directivity``, directivity.py:9-55)."""
    return SourceModel.parse(value).as_dict()


def source_forward(model: Mapping[str, Any]) -> np.ndarray:
    """Unit forward axis of a normalised source dictionary (directivity.py:58-66)."""
    yaw, pitch = float(model.get("orientation_deg", 0.0)), float(model.get("elevation_deg", 0.0))
    return SourceModel("custom", 0.0, 1.0, yaw, pitch).forward()
