"""Shared inputs for the parity tests: one workload, one ray set, run through
both the HIP tracer (C ABI via hip_backend) and the oracle."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from acoustic_agent_amd import sampling, scenes
from oracle import oracle as O

C = 343.0


@dataclass
class Case:
    workload: scenes.Workload
    num_rays: int
    num_bounces: int
    num_bins: int
    bin_dur: float = 1e-3
    seed: int = 0
    rng: str = "reference"
    ambisonics: bool = False
    log_bounces: int = 0
    sources: np.ndarray | None = None
    extra: dict = field(default_factory=dict)

    @property
    def arrays(self):
        return self.workload.scene.kernel_arrays()

    @property
    def duration(self) -> float:
        return self.num_bins * self.bin_dur

    def ray_set(self):
        bank = sampling.diffuse_sample_bank(128)
        if self.rng == "reference":
            dirs = sampling.sphere_samples(self.num_rays, self.seed)
            U, J = sampling.diffuse_random_sequence(self.num_rays, self.num_bounces, 128, self.seed)
            return dirs, bank, U, J
        return None, bank, None, None

    def direct_delay(self, source=None) -> float:
        s = self.workload.source if source is None else source
        return float(np.linalg.norm(np.asarray(s) - self.workload.listener)) / C


def run_gpu(case: Case, *, use_bvh: bool = True, tuning=None, log_rays: int = 0, scene=None, source=None,
            plain: bool = False, visual_limit: int = 0, visual_stride: int = 1):
    from acoustic_agent_amd import hip_backend

    dirs, bank, U, J = case.ray_set()
    w = case.workload
    src = w.source if source is None else np.asarray(source)
    kw = dict(source=src, direct_delay=case.direct_delay(src))
    if case.sources is not None and source is None:
        kw = dict(source=None, direct_delay=None, sources=case.sources)
    return hip_backend.trace_energy_field_hip(
        listener=w.listener, directions=dirs, diffuse_bank=bank, diffuse_random=U, diffuse_indices=J,
        arrays=case.arrays, use_bvh=use_bvh, num_bounces=case.num_bounces, num_bins=case.num_bins,
        bin_dur=case.bin_dur, speed_of_sound=C, max_path_len=case.duration * C, listener_radius=0.1,
        source_radius=0.1, irradiance_min_distance=1.0, specular_exponent=100.0,
        source_forward_vector=np.asarray(case.extra.get("source_forward", (1.0, 0.0, 0.0))),
        dipole_weight=float(case.extra.get("dipole_weight", 0.0)),
        dipole_power=float(case.extra.get("dipole_power", 1.0)),
        render_ambisonics=case.ambisonics, device_id=0, rng_mode=case.rng, seed=case.seed,
        num_rays=case.num_rays, scene=scene, tessellation_max_edge=w.max_edge, log_rays=log_rays,
        log_bounces=case.log_bounces if log_rays else 0, tuning=tuning, plain=plain,
        visual_candidate_limit=visual_limit, visual_stride=visual_stride, **kw)


def run_oracle(case: Case, precision: str = "float32", *, triangles: bool = True, log_rays: int = 0,
               source=None, threads: int = 0, use_bvh: bool = True, ray_range=None, visual_limit: int = 0,
               visual_stride: int = 1):
    dirs, bank, U, J = case.ray_set()
    w = case.workload
    src = w.source if source is None else np.asarray(source)
    scene = O.OracleScene(case.arrays, w.mesh if triangles else None, precision, use_bvh=use_bvh)
    return O.trace(scene, source=src, listener=w.listener, num_rays=case.num_rays, num_bounces=case.num_bounces,
                   num_bins=case.num_bins, bin_dur=case.bin_dur, c=C, max_path_len=case.duration * C,
                   direct_delay=case.direct_delay(src), diffuse_bank=bank, rng=case.rng, directions=dirs,
                   source_forward=case.extra.get("source_forward", (1.0, 0.0, 0.0)),
                   dipole_weight=float(case.extra.get("dipole_weight", 0.0)),
                   dipole_power=float(case.extra.get("dipole_power", 1.0)),
                   diffuse_random=U, diffuse_indices=J, seed=case.seed, render_ambisonics=case.ambisonics,
                   log_rays=log_rays, log_bounces=case.log_bounces if log_rays else 0, threads=threads,
                   ray_range=ray_range, visual_limit=visual_limit, visual_stride=visual_stride)


def rel_l1(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def band_rel(a, b) -> np.ndarray:
    """Per-band relative difference of the total energy."""
    a = np.asarray(a, dtype=np.float64).sum(axis=-1)
    b = np.asarray(b, dtype=np.float64).sum(axis=-1)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
