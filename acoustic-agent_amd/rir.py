"""Echogram -> band impulse responses, on the host.

The traced field is turned into sound the way the reference does after its
tracer (acoustic_agent/steam_rt.py:2205-2232, 3070-3083, 4123-4212,
2824-2851):

* ``reconstruct_band_irs``: one white-noise sequence (``default_rng(seed +
  7)``, uniform in [-1, 1)), shared by the six bands, is scaled per band by
  the amplitude envelope ``sqrt(E / 4 pi)`` of the echogram (bins below
  1e-9 are silent), interpolated linearly over the samples of each bin from
  the previous bin's amplitude, times the air-absorption amplitude
  ``exp(-m_b c t)`` at the bin's centre time (plus the direct delay); the
  band signals are then band-limited.
* ``bandlimit_band_signals``: a complementary crossover at 176.78, 353.55,
  707.11, 1414.2 and 2828.4 Hz.  Band 0 is the low-pass at the first
  crossover, band i the difference of the low-passes at crossovers i and
  i-1, the top band the input minus the low-pass at the last crossover, so
  that the six outputs of six equal inputs sum to that input.  Each low-pass
  is an 8th-order Butterworth: four biquad sections (RBJ low-pass form, Q of
  section k = 1 / (2 cos((2k+1) pi / 16))), run in float64 with each
  section's output rounded to float32, as the reference's biquad does.
* ``estimate_reconstructed_reverb_times``: RT60 of rendered band signals by
  the same Schroeder fit as the echogram RT60, on the energy in bins of
  ``rt_bin_duration_s``.
* ``render_band_rirs`` / ``render_rir``: the reconstruction with the FDN late
  tail crossfaded in (``fdn.py``) at the field's band RT60, summed to mono.

This is host post-processing of the tracer's output (a few milliseconds per
RIR); it runs on NumPy/SciPy and needs no GPU.
"""
from __future__ import annotations

import math
from typing import Any, Mapping

import numpy as np
from scipy.signal import lfilter

from .models import FREQUENCY_BANDS, SimConfig

_NUM_BANDS = len(FREQUENCY_BANDS)
_ENERGY_THRESHOLD = 1e-9  # steam_rt.py:39
_CROSSOVERS_HZ = (176.78, 353.55, 707.11, 1414.2, 2828.4)  # steam_rt.py:4125


def _band_edges(fs: int) -> list[tuple[float, float]]:
    """(lo, hi) of each band: the crossovers between 0 and Nyquist (steam_rt.py:4123-4127)."""
    bounds = [0.0, *_CROSSOVERS_HZ, fs * 0.5]
    return [(bounds[i], bounds[i + 1]) for i in range(_NUM_BANDS)]


def _apply_biquad(values: np.ndarray, b0: float, b1: float, b2: float, a1: float, a2: float) -> np.ndarray:
    """y[n] = b0 x[n] + b1 x[n-1] + b2 x[n-2] - a1 y[n-1] - a2 y[n-2] from rest,
    in float64, returned as float32 (steam_rt.py:4133-4145, 4201-4212)."""
    x = np.asarray(values, dtype=np.float64).reshape(-1)
    if x.size == 0:
        return x.astype(np.float32)
    return lfilter([b0, b1, b2], [1.0, a1, a2], x).astype(np.float32)


def _butterworth_lowpass(signal: np.ndarray, cutoff_hz: float, fs: int) -> np.ndarray:
    """8th-order Butterworth low-pass at ``cutoff_hz`` (the crossover filter of
    steam_rt.py:4148-4172): four RBJ low-pass biquads with section Qs
    1 / (2 cos((2k + 1) pi / 16)), the cutoff clamped below Nyquist."""
    values = np.asarray(signal, dtype=np.float64).reshape(-1)
    if values.size == 0:
        return values.astype(np.float32)
    rate = max(1, int(fs))
    cutoff = float(np.clip(float(cutoff_hz), 1e-3, 0.5 * rate * 0.999))
    w0 = 2.0 * math.pi * cutoff / rate
    cw0, sw0 = math.cos(w0), math.sin(w0)
    out = values
    for section in range(4):
        q = 1.0 / (2.0 * math.cos((2.0 * section + 1.0) * math.pi / 16.0))
        alpha = sw0 / (2.0 * q)
        a0 = 1.0 + alpha
        b0 = ((1.0 - cw0) / 2.0) / a0
        out = _apply_biquad(out, b0, (1.0 - cw0) / a0, b0, (-2.0 * cw0) / a0, (1.0 - alpha) / a0)
    return out


def bandlimit_band_signals(signals: np.ndarray, fs: int) -> np.ndarray:
    """Complementary six-band crossover of [6, n] band signals -> float32 [6, n]
    (steam_rt.py:3070-3083)."""
    values = np.asarray(signals, dtype=np.float64)
    if values.ndim != 2 or values.shape[0] != _NUM_BANDS:
        raise ValueError(f"expected {_NUM_BANDS} band signals, got shape {values.shape}")
    crossovers = [edge[1] for edge in _band_edges(fs)[:-1]]
    out = np.zeros_like(values, dtype=np.float32)
    out[0] = _butterworth_lowpass(values[0], crossovers[0], fs)
    for b in range(1, _NUM_BANDS - 1):
        out[b] = _butterworth_lowpass(values[b], crossovers[b], fs) - _butterworth_lowpass(values[b], crossovers[b - 1],
                                                                                           fs)
    out[-1] = values[-1] - _butterworth_lowpass(values[-1], crossovers[-1], fs)
    return out


def _air_absorption_amplitude(coefficient_per_m: float, travel_time_s: float, speed_m_s: float) -> float:
    """exp(-m d) for d = c t (steam_rt.py:2191-2193)."""
    distance_m = max(0.0, float(travel_time_s)) * max(0.0, float(speed_m_s))
    return math.exp(-max(0.0, float(coefficient_per_m)) * distance_m)


def estimate_signal_reverb_time(signal: np.ndarray, config: SimConfig) -> float:
    """RT60 of one rendered signal: energy per bin of ``rt_bin_duration_s``,
    then the Schroeder fit (steam_rt.py:2854-2869)."""
    from . import steam_rt as _rt

    fs = max(1, int(config.fs))
    values = np.asarray(signal, dtype=np.float64).reshape(-1)
    if values.size == 0:
        return 0.0
    bin_samples = max(1, int(round(float(config.rt_bin_duration_s) * fs)))
    num_bins = int(math.ceil(values.size / bin_samples))
    squared = np.zeros(num_bins * bin_samples, dtype=np.float64)
    squared[:values.size] = values * values
    return _rt._schroeder_fit_rt60(squared.reshape(num_bins, bin_samples).sum(axis=1), bin_samples / fs)


def _bin_ramps(values: np.ndarray, spb: int) -> np.ndarray:
    """Per-sample values [..., K * spb] rising linearly within each bin from
    the previous bin's value (the first bin from its own) to the bin's value
    (steam_rt.py:2222-2228)."""
    w = np.arange(spb, dtype=np.float64) / spb
    prev = np.concatenate([values[..., :1], values[..., :-1]], axis=-1)
    ramp = (1.0 - w) * prev[..., None] + w * values[..., None]
    return ramp.reshape(*values.shape[:-1], -1)


def _band_amplitudes(field: Mapping[str, Any], config: SimConfig) -> tuple[np.ndarray, np.ndarray, int]:
    """(per-sample amplitude envelopes [6, n] with air absorption, bin
    energies [6, K] clipped at 0, samples per bin) of an echogram field."""
    from . import steam_rt as _rt  # the absorption table lives there (steam_rt re-exports this module)

    echogram = np.asarray(field["echogram"], dtype=np.float64)
    num_bins = int(field["num_bins"])
    fs = int(config.fs)
    spb = max(1, int(math.ceil(float(field["bin_duration_s"]) * fs)))
    direct_delay_s = max(0.0, float(field.get("direct_delay_s", 0.0)))
    # distance travelled at each bin's centre time (>= 0), for the absorption
    centre_m = (direct_delay_s + (np.arange(num_bins, dtype=np.float64) + 0.5) * spb / fs) * max(0.0, float(config.c))
    energy = np.clip(echogram[:, :num_bins], 0.0, None)
    amps = np.sqrt(energy / (4.0 * math.pi))
    amps[echogram[:, :num_bins] < _ENERGY_THRESHOLD] = 0.0
    coeffs = np.asarray([max(0.0, float(_rt.AIR_ABSORPTION_NP_PER_M[b])) for b in FREQUENCY_BANDS])
    air = np.exp(-coeffs[:, None] * centre_m[None, :])  # _air_absorption_amplitude per band and bin
    return _bin_ramps(amps, spb) * np.repeat(air, spb, axis=1), energy, spb


def reconstruct_band_irs(field: Mapping[str, Any], config: SimConfig) -> np.ndarray:
    """Band impulse responses [6, K * ceil(bin * fs)] (float32) of an echogram
    field (steam_rt.py:2205-2232)."""
    env, _, _ = _band_amplitudes(field, config)
    white = np.random.default_rng(config.seed + 7).uniform(-1.0, 1.0, size=env.shape[1]).astype(np.float64)
    return bandlimit_band_signals(env * white, int(config.fs))


def reconstruct_ambisonic_band_irs(field: Mapping[str, Any], config: SimConfig) -> np.ndarray:
    """First-order ambisonic band impulse responses [6, 4, n] (float32;
    channels W, X, Y, Z) of a field with an ``ambisonic_echogram`` [6, 4, K]
    (steam_rt.py:2680-2724).  Each channel is the band's amplitude envelope
    times its direction ratio (channel energy over band energy, clipped to
    [-1, 1]; W = 1; 0 in silent bins), ramped within bins like the
    amplitudes, on one white-noise sequence shared by all bands and
    channels.  Without an ambisonic echogram, W carries the band impulse
    responses and X, Y, Z are silent."""
    amb = field.get("ambisonic_echogram")
    if amb is None:
        mono = reconstruct_band_irs(field, config)
        out = np.zeros((_NUM_BANDS, 4, mono.shape[1]), dtype=np.float32)
        out[:, 0] = mono
        return out
    env, energy, spb = _band_amplitudes(field, config)
    amb = np.asarray(amb, dtype=np.float64)[:, :, :energy.shape[1]]
    live = energy > _ENERGY_THRESHOLD
    ratios = np.where(live[:, None, :], np.clip(amb / np.maximum(energy, 1e-12)[:, None, :], -1.0, 1.0), 0.0)
    ratios[:, 0] = live
    white = np.random.default_rng(config.seed + 11).uniform(-1.0, 1.0, size=env.shape[1]).astype(np.float64)
    raw = env[:, None, :] * _bin_ramps(ratios, spb) * white
    out = np.empty(raw.shape, dtype=np.float32)
    for ch in range(4):
        out[:, ch] = bandlimit_band_signals(raw[:, ch], int(config.fs))
    return out


def hybridize_ambisonic_tail(traced: np.ndarray, fdn_component: np.ndarray, meta: Mapping[str, Any],
                             config: SimConfig) -> np.ndarray:
    """The traced ambisonic bands [6, 4, n] faded out over each band's
    late-tail transition (gain sqrt of a linear ramp from 1 towards 0, zero
    after the transition) with the FDN tail [6, n] added to W
    (steam_rt.py:2654-2677); unchanged when no tail was applied."""
    out = np.asarray(traced, dtype=np.float64).copy()
    if not bool(meta.get("applied", False)):
        return out.astype(np.float32)
    fs, n = max(1, int(config.fs)), int(out.shape[-1])
    by_band = meta.get("transition_by_band") if isinstance(meta.get("transition_by_band"), Mapping) else {}
    for bi, band in enumerate(FREQUENCY_BANDS):
        tr = by_band.get(band) if isinstance(by_band.get(band), Mapping) else {}
        start = min(n, max(0, int(math.floor(float(tr.get("start_s", meta["transition_start_s"])) * fs))))
        end = min(n, max(start + 1, int(math.ceil(float(tr.get("end_s", meta["transition_end_s"])) * fs))))
        gain = np.ones(n)
        gain[start:end] = np.sqrt(1.0 - np.arange(end - start) / max(1, end - start))
        gain[end:] = 0.0
        out[bi] *= gain
    out[:, 0] += np.asarray(fdn_component, dtype=np.float64)
    return out.astype(np.float32)


def estimate_reconstructed_reverb_times(band_rirs: np.ndarray, config: SimConfig) -> dict[str, float]:
    """Band RT60 of rendered band signals [6, n]: energy per bin of
    ``rt_bin_duration_s``, then the Schroeder fit (steam_rt.py:2824-2851)."""
    from . import steam_rt as _rt

    fs = max(1, int(config.fs))
    signals = np.asarray(band_rirs, dtype=np.float64)
    if signals.ndim != 2 or signals.shape[0] != _NUM_BANDS:
        return {band: 0.0 for band in FREQUENCY_BANDS}
    bin_samples = max(1, int(round(float(config.rt_bin_duration_s) * fs)))
    n = int(signals.shape[1])
    num_bins = max(1, int(math.ceil(n / bin_samples)))
    squared = np.zeros((_NUM_BANDS, num_bins * bin_samples), dtype=np.float64)
    squared[:, :n] = np.square(signals)
    energy = squared.reshape(_NUM_BANDS, num_bins, bin_samples).sum(axis=2)
    return {band: _rt._schroeder_fit_rt60(energy[bi], bin_samples / fs) for bi, band in enumerate(FREQUENCY_BANDS)}


def render_with_tail(field: Mapping[str, Any], config: SimConfig, rt60_bands: Mapping[str, float] | None = None,
                     multi_room: bool = False, late_reverb_prior: Mapping[str, float] | None = None
                     ) -> tuple[np.ndarray, np.ndarray, dict[str, Any]]:
    """(band impulse responses [6, n] with the late tail, the tail alone
    [6, n], metadata): ``render_band_rirs`` with the FDN component kept for
    the ambisonic hybridisation."""
    from . import steam_rt as _rt
    from .decay import apply_coupled_late_reverb_prior, estimate_late_reverb_times
    from .fdn import _render_parametric_fdn_late_reverb

    traced = reconstruct_band_irs(field, config)
    rt60 = dict(rt60_bands) if rt60_bands is not None else _rt.estimate_reverb_times(field, config)
    profiles = None
    if multi_room:
        rt60, profiles = estimate_late_reverb_times(field, config, fallback=rt60)
        rt60, profiles = apply_coupled_late_reverb_prior(rt60, profiles, late_reverb_prior)
    bands, fdn_part, meta = _render_parametric_fdn_late_reverb(traced, field, rt60, config, decay_profiles=profiles)
    if profiles is not None:
        meta = {**meta, "late_decay_profiles": profiles}
    return bands, fdn_part, meta


def render_band_rirs(field: Mapping[str, Any], config: SimConfig,
                     rt60_bands: Mapping[str, float] | None = None, *, multi_room: bool = False,
                     late_reverb_prior: Mapping[str, float] | None = None) -> tuple[np.ndarray, dict[str, Any]]:
    """Band impulse responses [6, n] of a traced field with the FDN late tail
    (``config.late_tail``) at the field's band RT60 (or ``rt60_bands``), and
    the late-tail metadata.  For a multi-room scene (steam_rt.py:1015-1035)
    the tail follows each band's late decay slope where the echogram shows a
    coupled-space break, and starts its crossfade at that break; a
    room-system ``late_reverb_prior`` replaces the late targets.  The
    metadata then carries the profiles as ``late_decay_profiles``."""
    bands, _, meta = render_with_tail(field, config, rt60_bands, multi_room, late_reverb_prior)
    return bands, meta


def render_ambisonic_rir(field: Mapping[str, Any], config: SimConfig,
                         rt60_bands: Mapping[str, float] | None = None, *, multi_room: bool = False,
                         late_reverb_prior: Mapping[str, float] | None = None) -> np.ndarray:
    """First-order ambisonic impulse response [4, n] (W, X, Y, Z): the
    ambisonic band reconstruction with the same late tail as
    ``render_band_rirs`` hybridised into W, summed over bands
    (steam_rt.py:1046-1053)."""
    _, fdn_part, meta = render_with_tail(field, config, rt60_bands, multi_room, late_reverb_prior)
    amb = hybridize_ambisonic_tail(reconstruct_ambisonic_band_irs(field, config), fdn_part, meta, config)
    return np.sum(amb, axis=0, dtype=np.float64).astype(np.float32)


def render_rir(field: Mapping[str, Any], config: SimConfig) -> np.ndarray:
    """Mono impulse response: the sum of the band impulse responses."""
    bands, _ = render_band_rirs(field, config)
    return np.sum(bands, axis=0, dtype=np.float64).astype(np.float32)
