"""Tracer configuration contract.

Mirrors the tracer- and late-tail-facing part of the reference's
``SimConfig`` (acoustic_agent/models.py:60-112) with the same field names and
defaults, plus
the fields of the HIP backend (``rt_hip_device``, ``rt_rng``,
``rt_tessellation_max_edge_m``).  The six bands are those of
acoustic_agent/models.py:9.
"""
from __future__ import annotations

from dataclasses import dataclass

FREQUENCY_BANDS: tuple[str, ...] = ("125", "250", "500", "1000", "2000", "4000")


@dataclass(frozen=True)
class SimConfig:
    fs: int = 16000
    c: float = 343.0
    duration_s: float = 2.0
    seed: int = 1729
    late_tail: bool = True
    hybrid_transition_s: float = 1.0
    hybrid_overlap_fraction: float = 0.25
    rt_accelerator: str = "hip"
    rt_precision: str = "float32"
    rt_cuda_device: int = 0
    rt_num_rays: int = 32768
    rt_num_bounces: int = 96
    rt_num_diffuse_samples: int = 128
    rt_duration_s: float = 2.0
    rt_bin_duration_s: float = 0.01
    rt_specular_exponent: float = 100.0
    rt_irradiance_min_distance: float = 1.0
    rt_source_radius: float = 0.1
    rt_listener_radius: float = 0.1
    intersection_backend: str = "auto"
    bvh_min_surfaces: int = 16
    collect_visual_paths: bool = False
    rt_visual_num_rays: int | None = None
    rt_visual_num_bounces: int | None = None
    render_ambisonics: bool = True
    # HIP backend
    rt_hip_device: int | None = None   # None: use rt_cuda_device
    rt_rng: str = "counter"            # "counter" (Philox on device) or "reference" (NumPy streams)
    rt_tessellation_max_edge_m: float | None = None  # None: coarsest tessellation

    @property
    def device(self) -> int:
        return int(self.rt_cuda_device if self.rt_hip_device is None else self.rt_hip_device)
