"""Host side of the tracer: backend dispatch, the field dictionary, and the
echogram -> band RT60 / EDT metric.

``trace_energy_field`` mirrors acoustic_agent/steam_rt.py:1757-1804 and the
host preparation of ``_trace_energy_field_cuda`` (steam_rt.py:1822-1923):
``num_bins = ceil(max(rt_duration_s, duration_s) / bin)``, ``max_path_len =
duration * c``, ``direct_delay = |S - L| / c``, the ray set, and the field
dictionary every downstream consumer reads.  The accelerator ``"hip"`` is the
MI355X backend (``hip_backend``); it accepts float32 only, as the reference's
GPU backend does (steam_rt.py:1777-1778).  This package ships no CPU tracer:
``"numba"`` (and ``"auto"`` without a HIP device or with float64) raise.

``estimate_reverb_times`` and ``_schroeder_fit_rt60`` restate
steam_rt.py:2727-2742 and 3052-3067 (Schroeder integral, least squares on
-5..-25 dB, RT60 = -60/slope) on air-absorption weighted echograms
(steam_rt.py:2196-2202, acoustics.py:12-22).  The reference computes no EDT;
``estimate_early_decay_times`` uses the ISO 3382-1 definition (fit on
0..-10 dB, times 6) on the same weighted echogram.
"""
from __future__ import annotations

import math
from typing import Any, Mapping

import numpy as np

from .directivity import source_directivity, source_forward
from .models import FREQUENCY_BANDS, SimConfig
from .sampling import diffuse_random_sequence, diffuse_sample_bank, sphere_samples

AIR_ABSORPTION_NP_PER_M = {
    "125": 0.0,
    "250": 0.00011513,
    "500": 0.00034539,
    "1000": 0.00057565,
    "2000": 0.0011513,
    "4000": 0.0034539,
}
_SH_Y00 = 1.0 / math.sqrt(4.0 * math.pi)
_RT_VISUAL_RETAIN_LIMIT = 512  # steam_rt.py:41


def visual_candidate_plan(config: SimConfig, num_rays: int) -> tuple[int, int]:
    """(limit, stride) of the visual path candidates (steam_rt.py:1853-1860):
    up to 4 x 512 (or rt_visual_num_rays) rays spread evenly over the ray
    set, none unless collect_visual_paths."""
    default = int(_RT_VISUAL_RETAIN_LIMIT * 4)
    requested = default if config.rt_visual_num_rays is None else max(1, min(int(config.rt_visual_num_rays), default))
    limit = min(int(num_rays), requested) if config.collect_visual_paths else 0
    return limit, max(1, int(num_rays) // max(limit, 1))


def histogram_bins(duration_s: float, bin_dur: float) -> int:
    """Bins of the time histogram: ceil(duration / bin), at least one (steam_rt.py:1947-1949)."""
    return max(1, int(math.ceil(float(duration_s) / float(bin_dur))))


def _normalize_rt_accelerator(value: str) -> str:
    accelerator = str(value).strip().lower()
    if accelerator not in {"auto", "numba", "cuda", "hip"}:
        raise ValueError("rt_accelerator must be auto, numba, cuda, or hip")
    return accelerator


def _normalize_rt_precision(value: str) -> str:
    aliases = {"fp32": "float32", "single": "float32", "fp64": "float64", "double": "float64"}
    precision = aliases.get(str(value).strip().lower(), str(value).strip().lower())
    if precision not in {"float32", "float64"}:
        raise ValueError("rt_precision must be float32 or float64")
    return precision


def _resolve_intersection_backend(value: str, surface_count: int, min_surfaces: int) -> str:
    backend = str(value).strip().lower()
    if backend not in {"auto", "linear", "bvh"}:
        raise ValueError("intersection_backend must be auto, linear, or bvh")
    threshold = max(1, int(min_surfaces))
    return "bvh" if backend == "bvh" or (backend == "auto" and int(surface_count) >= threshold) else "linear"


def trace_energy_field(
    scene: Any,
    source: np.ndarray,
    listener: np.ndarray,
    config: SimConfig,
    source_model: str | Mapping[str, Any] | None = None,
    *,
    render_ambisonics: bool = True,
) -> dict[str, Any]:
    """Trace the listener-centric energy field of ``scene`` (a ``RoomScene``)."""
    emitter = source_directivity(source_model)
    accelerator = _normalize_rt_accelerator(config.rt_accelerator)
    precision = _normalize_rt_precision(config.rt_precision)
    if accelerator in {"numba", "cuda"}:
        raise RuntimeError(
            f"rt_accelerator={accelerator!r} is the reference engine's own backend; this package provides 'hip'"
        )
    from .hip_backend import hip_available

    available = hip_available(config.device)
    if accelerator == "hip" and not available:
        raise RuntimeError(f"HIP accelerator requested for device {config.device}, but no HIP device is available")
    if precision != "float32":
        raise ValueError("HIP tracing supports rt_precision='float32' only")
    if not available:
        raise RuntimeError("rt_accelerator='auto' found no HIP device and this package has no CPU tracer")
    return _trace_energy_field_hip(scene, source, listener, config, emitter, render_ambisonics=render_ambisonics)


def _trace_energy_field_hip(
    scene: Any,
    source: np.ndarray,
    listener: np.ndarray,
    config: SimConfig,
    emitter: Mapping[str, Any],
    *,
    render_ambisonics: bool = True,
) -> dict[str, Any]:
    from .hip_backend import trace_energy_field_hip

    arrays = scene.kernel_arrays()
    intersection_backend = _resolve_intersection_backend(config.intersection_backend, scene.num_surfaces,
                                                         config.bvh_min_surfaces)
    num_rays = int(config.rt_num_rays)
    num_bounces = int(config.rt_num_bounces)
    bin_dur = float(config.rt_bin_duration_s)
    duration = max(float(config.rt_duration_s), float(config.duration_s))
    num_bins = histogram_bins(duration, bin_dur)
    bank = diffuse_sample_bank(config.rt_num_diffuse_samples)
    rng = str(config.rt_rng).lower()
    if rng == "reference":
        directions = sphere_samples(num_rays, int(config.seed))
        diffuse_random, diffuse_indices = diffuse_random_sequence(num_rays, num_bounces, bank.shape[0],
                                                                  int(config.seed))
    elif rng == "counter":
        directions = diffuse_random = diffuse_indices = None
    else:
        raise ValueError("rt_rng must be 'counter' or 'reference'")
    direct_delay = float(np.linalg.norm(np.asarray(source, float) - np.asarray(listener, float))) / float(config.c)
    visual_limit, visual_stride = visual_candidate_plan(config, num_rays)
    raw = trace_energy_field_hip(
        source=np.asarray(source, dtype=np.float32),
        listener=np.asarray(listener, dtype=np.float32),
        directions=directions,
        diffuse_bank=bank,
        diffuse_random=diffuse_random,
        diffuse_indices=diffuse_indices,
        arrays=arrays,
        use_bvh=intersection_backend == "bvh",
        num_bounces=num_bounces,
        num_bins=num_bins,
        bin_dur=bin_dur,
        speed_of_sound=float(config.c),
        max_path_len=duration * float(config.c),
        direct_delay=direct_delay,
        listener_radius=float(config.rt_listener_radius),
        source_radius=float(config.rt_source_radius),
        irradiance_min_distance=float(config.rt_irradiance_min_distance),
        specular_exponent=float(config.rt_specular_exponent),
        source_forward_vector=np.asarray(source_forward(emitter), dtype=np.float32),
        dipole_weight=float(emitter["dipole_weight"]),
        dipole_power=float(emitter["dipole_power"]),
        render_ambisonics=render_ambisonics,
        device_id=config.device,
        rng_mode=rng,
        seed=int(config.seed),
        num_rays=num_rays,
        tessellation_max_edge=config.rt_tessellation_max_edge_m,
        visual_candidate_limit=visual_limit,
        visual_stride=visual_stride,
    )
    return field_from_raw(raw, arrays["names"], num_bins, bin_dur, direct_delay, render_ambisonics,
                          intersection_backend, emitter,
                          visual=(visual_limit, visual_stride) if config.collect_visual_paths else None)


def field_from_raw(raw: Mapping[str, Any], names, num_bins: int, bin_dur: float, direct_delay: float,
                   render_ambisonics: bool, intersection_backend: str, emitter: Mapping[str, Any],
                   visual: tuple[int, int] | None = None) -> dict[str, Any]:
    """The field dictionary of steam_rt.py:1889-1923 from one raw trace;
    ``visual`` = (candidate limit, stride) when visual paths were collected."""
    echogram = np.asarray(raw["echogram"])
    total_energy = np.sum(echogram, axis=0)
    nonzero = np.flatnonzero(total_energy > 0.0)
    hits, contribs, energy = raw["hit_counts"], raw["contrib_counts"], raw["surface_energy"]
    return {
        "echogram": echogram,
        "ambisonic_echogram": raw["ambisonic"] if render_ambisonics else None,
        "num_bins": int(num_bins),
        "bin_duration_s": float(bin_dur),
        "direct_delay_s": float(direct_delay),
        "actual_bounces": int(raw["actual_bounces"]),
        "active_ray_count": int(raw["active_count"]),
        "last_energy_time_s": float(nonzero[-1] * bin_dur) if nonzero.size else 0.0,
        "surface_hit_count": {names[i]: int(hits[i]) for i in range(len(names)) if int(hits[i]) > 0},
        "surface_contribution_count": {names[i]: int(contribs[i]) for i in range(len(names)) if int(contribs[i]) > 0},
        "surface_energy": {names[i]: float(energy[i]) for i in range(len(names)) if float(energy[i]) > 0.0},
        "visual_candidates": {
            "hit_points": raw["visual_hit_points"],
            "surface_indices": raw["visual_surface_indices"],
            "ray_indices": raw["visual_ray_indices"],
            "orders": raw["visual_orders"],
            "distances": raw["visual_distances"],
            "gains": raw["visual_gains"],
            "surface_names": tuple(names),
            "candidate_limit": int(visual[0]),
            "stride": int(visual[1]),
        } if visual is not None else None,
        "accelerator": "hip",
        "precision": "float32",
        "hip": raw.get("device"),
        "device_input_cache": raw.get("device_input_cache"),
        "kernel_time_s": float(raw.get("kernel_time_s", 0.0)),
        "transfer_time_s": float(raw.get("transfer_time_s", 0.0)),
        "intersection_backend": intersection_backend,
        "source_directivity": dict(emitter),
        "counters": dict(raw.get("counters", {})),
    }


# ---------------------------------------------------------------------------
# echogram -> RT60 / EDT
# ---------------------------------------------------------------------------


def _air_absorption_energy_weights(coefficient_per_m: float, travel_times_s: np.ndarray, speed_m_s: float) -> np.ndarray:
    distances_m = np.maximum(np.asarray(travel_times_s, dtype=np.float64), 0.0) * max(0.0, float(speed_m_s))
    return np.exp(-2.0 * max(0.0, float(coefficient_per_m)) * distances_m)


def _fit_decay(energy: np.ndarray, bin_dur: float, hi_db: float, lo_db: float, scale: float,
               min_total_energy: float) -> float:
    total = float(np.sum(energy))
    if total < float(min_total_energy):
        return 0.0
    edc = np.cumsum(np.asarray(energy, dtype=np.float64)[::-1])[::-1]
    db = 10.0 * np.log10(np.clip(edc / total, 1e-12, None))
    x = float(bin_dur) * np.arange(edc.size, dtype=np.float64)
    fit = (db <= hi_db) & (db >= lo_db)
    if int(np.count_nonzero(fit)) < 2:
        return 0.0
    xs, ys = x[fit], db[fit]
    n = xs.size
    denom = n * np.sum(xs * xs) - np.sum(xs) ** 2
    numer = n * np.sum(xs * ys) - np.sum(xs) * np.sum(ys)
    slope = numer / denom if abs(denom) > 1e-12 else 0.0
    return float(max(0.0, scale / slope)) if slope < -1e-9 else 0.0


def _schroeder_fit_rt60(energy: np.ndarray, bin_dur: float, *, min_total_energy: float = 1e-12) -> float:
    return _fit_decay(energy, bin_dur, -5.0, -25.0, -60.0, min_total_energy)


def _schroeder_fit_edt(energy: np.ndarray, bin_dur: float, *, min_total_energy: float = 1e-12) -> float:
    # ISO 3382-1 early decay time: regression over 0 .. -10 dB, extrapolated to 60 dB.
    return _fit_decay(energy, bin_dur, 0.0, -10.0, -60.0, min_total_energy)


def _weighted_bands(field: Mapping[str, Any], config: SimConfig) -> list[np.ndarray]:
    echogram = np.asarray(field["echogram"], dtype=np.float64)
    bin_dur = float(field["bin_duration_s"])
    times = bin_dur * np.arange(int(field["num_bins"]), dtype=np.float64)
    return [(_SH_Y00 * echogram[bi]) * _air_absorption_energy_weights(AIR_ABSORPTION_NP_PER_M[band], times,
                                                                        float(config.c))
            for bi, band in enumerate(FREQUENCY_BANDS)]


def estimate_reverb_times(field: Mapping[str, Any], config: SimConfig) -> dict[str, float]:
    bin_dur = float(field["bin_duration_s"])
    return {band: _schroeder_fit_rt60(w, bin_dur, min_total_energy=1e-4)
            for band, w in zip(FREQUENCY_BANDS, _weighted_bands(field, config))}


def estimate_early_decay_times(field: Mapping[str, Any], config: SimConfig) -> dict[str, float]:
    bin_dur = float(field["bin_duration_s"])
    return {band: _schroeder_fit_edt(w, bin_dur, min_total_energy=1e-4)
            for band, w in zip(FREQUENCY_BANDS, _weighted_bands(field, config))}


# echogram -> band impulse responses (host), under the reference's names
from .rir import (  # noqa: E402,F401
    bandlimit_band_signals,
    estimate_reconstructed_reverb_times,
    estimate_signal_reverb_time,
    hybridize_ambisonic_tail as _hybridize_ambisonic_tail,
    reconstruct_ambisonic_band_irs,
    reconstruct_band_irs,
)
# echogram -> decay profiles and late-tail targets, under the reference's names
from .decay import (  # noqa: E402,F401
    apply_coupled_late_reverb_prior as _apply_coupled_late_reverb_prior,
    energy_decay_profile as _energy_decay_profile,
    estimate_late_reverb_times,
)
