"""The reflection part of the reference's ``simulate_rir``
(acoustic_agent/engine.py:30-91 and the reflection branch of
``simulate_steam_room``, steam_rt.py:944-1125): for each traced receiver
channel, trace the energy field on the GPU, estimate its band RT60 and EDT,
render the band impulse responses with the FDN late tail, and place them
from the direct-path arrival in an impulse response of ``duration_s``.

Array receivers trace one channel per position (engine.py:64-73) and return
[channels, n]; a mono receiver returns [1, n] and, with
``config.render_ambisonics``, the first-order ambisonic impulse response.
The direct path, edge diffraction, material-based priors and binaural
(HRTF) rendering belong to the reference's engine around this path and are
not part of this package: the result holds the reflections only.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Mapping, Sequence

import numpy as np

from . import rir, steam_rt
from .directivity import source_directivity
from .mic import LAYOUTS, channel_positions, microphone_array
from .models import FREQUENCY_BANDS, SimConfig


@dataclass(frozen=True)
class ReflectionResult:
    """Shaped like the reference's ``SimulationResult`` (engine.py:20-27),
    without ``paths``; ``fields`` are the traced energy fields, one per
    channel."""

    rir: np.ndarray                      # [channels, n] float32
    rt60: Mapping[str, Any]
    receiver_model: Mapping[str, Any]
    metadata: Mapping[str, Any] = field(default_factory=dict)
    ambisonic_rir: np.ndarray | None = None   # [4, n] float32 (W, X, Y, Z)
    source_model: Mapping[str, Any] = field(default_factory=dict)
    fields: tuple = ()


def _mean_positive(values: Mapping[str, float]) -> float:
    live = [float(v) for v in values.values() if float(v) > 0.0]
    return round(float(np.mean(live)), 4) if live else 0.0


def _rounded(values: Mapping[str, float]) -> dict[str, float]:
    return {band: round(float(values[band]), 4) for band in FREQUENCY_BANDS}


def _place(signal: np.ndarray, start: int, total: int) -> np.ndarray:
    """``signal`` [..., m] starting at sample ``start`` of a [..., total] buffer."""
    out = np.zeros(signal.shape[:-1] + (total,), dtype=np.float32)
    if 0 <= start < total:
        seg = min(signal.shape[-1], total - start)
        out[..., start:start + seg] = signal[..., :seg]
    return out


def _render_channel(fld: Mapping[str, Any], config: SimConfig, multi_room: bool,
                    late_reverb_prior: Mapping[str, float] | None, ambisonics: bool) -> dict[str, Any]:
    fs = max(1, int(config.fs))
    total = max(1, int(round(float(config.duration_s) * fs)))
    start = int(round(float(fld.get("direct_delay_s", 0.0)) * fs))
    traced = steam_rt.estimate_reverb_times(fld, config)
    edt = steam_rt.estimate_early_decay_times(fld, config)
    bands, tail, meta = rir.render_with_tail(fld, config, traced, multi_room, late_reverb_prior)
    placed = _place(np.asarray(bands, dtype=np.float32), start, total)
    mono = np.sum(placed, axis=0, dtype=np.float32)
    out = {"rir": mono, "bands": placed, "meta": meta, "traced": traced, "edt": edt,
           "hybrid": rir.estimate_reconstructed_reverb_times(bands, config),
           "rir_bands": rir.estimate_reconstructed_reverb_times(placed, config),
           "rir_s": rir.estimate_signal_reverb_time(mono, config)}
    if ambisonics:
        amb = rir.hybridize_ambisonic_tail(rir.reconstruct_ambisonic_band_irs(fld, config), tail, meta, config)
        out["ambisonic"] = _place(np.sum(amb, axis=0, dtype=np.float64).astype(np.float32), start, total)
    return out


def _result(fields: Sequence[Mapping[str, Any]], cfg: SimConfig, model: Mapping[str, Any],
            emitter: Mapping[str, Any], multi_room: bool, late_reverb_prior: Mapping[str, float] | None,
            ambisonics: bool, array: bool, batched: bool) -> ReflectionResult:
    """Render each channel's field and assemble the result (the RT60s are
    the first channel's, as in engine.py:72-76)."""
    channels = [_render_channel(f, cfg, multi_room, late_reverb_prior, ambisonics) for f in fields]
    first = channels[0]
    rt60 = {
        "rt60_bands": _rounded(first["rir_bands"]), "rt60_s": round(float(first["rir_s"]), 4),
        "model": "schroeder_fit_from_final_rir",
        "traced_rt60_bands": _rounded(first["traced"]), "traced_rt60_s": _mean_positive(first["traced"]),
        "traced_model": "schroeder_fit_from_path_traced_echogram",
        "hybrid_rt60_bands": _rounded(first["hybrid"]), "hybrid_rt60_s": _mean_positive(first["hybrid"]),
        "edt_bands": _rounded(first["edt"]), "edt_model": "iso3382_0_to_minus10_db_from_path_traced_echogram",
    }
    meta = {"model": "pathtraced_reflections_with_fdn_late_tail", "late_tail": first["meta"],
            "direct_delay_s": float(fields[0].get("direct_delay_s", 0.0)), "multi_room": bool(multi_room),
            "batched": bool(batched)}
    if multi_room:
        profiles = first["meta"].get("late_decay_profiles", {})
        rt60["late_tail_target_rt60_bands"] = {
            band: float(profiles.get(band, {}).get("selected_target_rt60_s", 0.0)) for band in FREQUENCY_BANDS}
    if array:
        meta["array_channel_count"] = len(channels)
    out = np.stack([c["rir"] for c in channels], axis=0).astype(np.float32)
    return ReflectionResult(rir=out, rt60=rt60, receiver_model=dict(model), metadata=meta,
                            ambisonic_rir=first.get("ambisonic"), source_model=dict(emitter), fields=tuple(fields))


def _batched_fields(scene: Any, source: np.ndarray, positions: Sequence[Sequence[float]], config: SimConfig,
                    emitter: Mapping[str, Any], ambisonics: bool) -> tuple[dict[str, Any], ...]:
    """The fields of all channels from one persistent batch launch (the same
    arithmetic and ray set as one ``trace_energy_field`` per channel)."""
    from .batch import BatchEngine
    from .hip_backend import cached_scene
    from .plan import _request

    arrays = scene.kernel_arrays()
    device_scene = cached_scene(arrays, device=config.device, max_edge=config.rt_tessellation_max_edge_m)
    reqs = [_request(device_scene, p, source.reshape(1, 3), config, ("channel", i), emitter, ambisonics)
            for i, p in enumerate(positions)]
    engine = BatchEngine(config.device, max_traces=len(reqs))
    try:
        results = {res["tag"]: res for res in engine.submit(reqs).wait()}
    finally:
        engine.close()
    backend = "bvh" if reqs[0].use_bvh else "linear"
    fields = []
    for r in reqs:
        res = results[r.tag]
        raw = {k: res[k][0] for k in ("echogram", "hit_counts", "contrib_counts", "surface_energy",
                                        "actual_bounces", "active_count")}
        raw.update(ambisonic=res["ambisonic"][0] if ambisonics else None, counters=res["counters"],
                   device=config.device)
        direct = float(np.linalg.norm(source - np.asarray(r.listener))) / float(config.c)
        fields.append(steam_rt.field_from_raw(raw, arrays["names"], r.num_bins, r.bin_dur, direct, ambisonics,
                                              backend, emitter))
    return tuple(fields)


def _receiver(receiver_model: Mapping[str, Any] | None, cfg: SimConfig) -> tuple[dict[str, Any], bool, bool]:
    model = dict(receiver_model or microphone_array("mono"))
    if str(model.get("type", "mono")) == "hrtf":
        raise ValueError("binaural (HRTF) rendering is not part of this package; trace a 'mono' receiver with "
                         "config.render_ambisonics to get the first-order ambisonic response it renders from")
    array = str(model.get("type", "mono")) in LAYOUTS
    return model, array, bool(cfg.render_ambisonics) and not array


def simulate_reflections(scene: Any, source: Sequence[float], receiver: Sequence[float], *,
                         config: SimConfig | None = None, receiver_model: Mapping[str, Any] | None = None,
                         source_model: str | Mapping[str, Any] | None = None, multi_room: bool = False,
                         late_reverb_prior: Mapping[str, float] | None = None,
                         batched: bool | None = None) -> ReflectionResult:
    """Reflection impulse response(s) of ``scene`` (a ``RoomScene``) for a
    source and a receiver (engine.py:30-91).  ``multi_room`` applies the
    late-tail targets of coupled spaces (steam_rt.py:1015-1035);
    ``late_reverb_prior`` is the room-system RT60 per band the reference
    derives from its material model.  ``batched`` traces all channels in one
    persistent launch (default: for several channels with the counter RNG,
    which the batch kernel draws on the device) instead of one
    ``trace_energy_field`` per channel."""
    cfg = config or SimConfig()
    model, array, ambisonics = _receiver(receiver_model, cfg)
    emitter = source_directivity(source_model)
    positions = channel_positions(receiver, model)
    src = np.asarray(source, dtype=np.float64)
    if batched is None:
        batched = len(positions) > 1 and str(cfg.rt_rng).lower() == "counter"
    if batched:
        fields = _batched_fields(scene, src, positions, cfg, emitter, ambisonics)
    else:
        fields = tuple(steam_rt.trace_energy_field(scene, src, np.asarray(p), cfg, emitter,
                                                   render_ambisonics=ambisonics) for p in positions)
    return _result(fields, cfg, model, emitter, multi_room, late_reverb_prior, ambisonics, array, batched)


def simulate_motion(scene: Any, frames: Sequence[Mapping[str, Any]], *, config: SimConfig | None = None,
                    receiver_model: Mapping[str, Any] | None = None,
                    source_model: str | Mapping[str, Any] | None = None, multi_room: bool = False,
                    late_reverb_prior: Mapping[str, float] | None = None) -> tuple[ReflectionResult, ...]:
    """One ``ReflectionResult`` per motion frame ``{"source", "receiver"}``
    (the reference's frame loop, api.py:669-713), traced in one persistent
    batch: the frames that share a receiver position are one multi-source
    trace per channel (ray trajectories do not depend on the source), so a
    frame's result equals its own ``simulate_reflections``."""
    from .batch import BatchEngine
    from .hip_backend import cached_scene
    from .plan import motion_traces

    cfg = config or SimConfig()
    model, array, ambisonics = _receiver(receiver_model, cfg)
    emitter = source_directivity(source_model)
    arrays = scene.kernel_arrays()
    device_scene = cached_scene(arrays, device=cfg.device, max_edge=cfg.rt_tessellation_max_edge_m)
    mplan = motion_traces(device_scene, frames, cfg, model, emitter, ambisonics)
    engine = BatchEngine(cfg.device, max_traces=len(mplan.traces))
    try:
        by_tag = {res["tag"]: res for res in engine.submit(mplan.traces).wait()}
    finally:
        engine.close()
    backend = "bvh" if mplan.traces[0].use_bvh else "linear"
    out = []
    for f, pairs in enumerate(mplan.frames):
        src = np.asarray(frames[f]["source"], dtype=np.float64)[:3]
        fields = []
        for t, k in pairs:
            req = mplan.traces[t]
            res = by_tag[req.tag]
            raw = {key: res[key][k] for key in ("echogram", "hit_counts", "contrib_counts", "surface_energy",
                                                "actual_bounces", "active_count")}
            raw.update(ambisonic=res["ambisonic"][k] if ambisonics else None, counters=res["counters"],
                       device=cfg.device)
            direct = float(np.linalg.norm(src - np.asarray(req.listener))) / float(cfg.c)
            fields.append(steam_rt.field_from_raw(raw, arrays["names"], req.num_bins, req.bin_dur, direct,
                                                  ambisonics, backend, emitter))
        out.append(_result(fields, cfg, model, emitter, multi_room, late_reverb_prior, ambisonics, array, True))
    return tuple(out)
