"""Parametric late reverberation: a 16-line Hadamard FDN crossfaded into the
traced band impulse responses (host).

Follows acoustic_agent/steam_rt.py:2235-2563:

* delay lines: the 16 primes 2..53, line p gets p^round(log t / log p) with
  t = 0.15 * 10 * fs / 16 + U{0..100} (``default_rng(seed + 1701)``);
  all-pass stages of 225, 341, 441 and 556 samples;
* per line, a cascade of six EQ sections (low shelf at the first crossover,
  peaking sections over the inner bands, high shelf at the last crossover;
  RBJ forms, Q 0.707, shelf slope a = sqrt(gain)) whose gains are solved in
  the log domain so that the cascade's gain at the band centres equals the
  per-pass decay exp(-6.91 d / (RT60_b fs)) of a line of d samples (the
  response matrix is the log-gain of each section at a probe gain of 0.5);
* a tone cascade with gains sqrt(1 / RT60_b), normalised to a maximum of 1;
* the FDN impulse is pre-rolled by max(delay) + sum(all-pass delays) samples,
  band-limited like the traced signals, and each band is scaled so that its
  energy after the anchor sample equals that of an exponential decay at the
  band's RT60 starting from the traced power at the anchor;
* band b crossfades at equal power from the traced signal into the FDN over
  [start, end): by default [(1 - overlap) T, T] with T = hybrid_transition_s;
  a band whose decay profile selected a fitted late slope with a breakpoint
  inside the response uses [(1 - overlap) breakpoint, breakpoint] and anchors
  at the breakpoint.

The per-sample recursion runs natively (``aa_fdn_impulse``).
"""
from __future__ import annotations

import ctypes
import math
from typing import Any, Mapping

import numpy as np

from . import _lib
from .models import FREQUENCY_BANDS, SimConfig
from .rir import _band_edges, bandlimit_band_signals

_NUM_BANDS = len(FREQUENCY_BANDS)
_PRIMES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53)
ALLPASS_DELAYS = (225, 341, 441, 556)
MODEL = "steam_style_16_line_hadamard_fdn"


def _steam_fdn_delays(fs: int, seed: int) -> np.ndarray:
    """Line lengths: a power of each prime near the jittered target (steam_rt.py:2333-2342)."""
    base = max(1, int(0.15 * 10.0 * max(int(fs), 1) / len(_PRIMES)))
    jitter = np.random.default_rng(int(seed) + 1701).integers(0, 101, size=len(_PRIMES))
    out = np.empty(len(_PRIMES), dtype=np.int64)
    for i, p in enumerate(_PRIMES):
        target = max(1, base + int(jitter[i]))
        out[i] = max(1, p ** max(1, int(round(math.log(target) / math.log(p)))))
    return out


def _shelf(cutoff: float, gain: float, fs: int, high: bool) -> np.ndarray:
    """RBJ shelf, Q 0.707, A = sqrt(gain): DC (low) or Nyquist (high) gain = gain
    (steam_rt.py:2345-2378)."""
    w0 = 2.0 * math.pi * float(cutoff) / max(float(fs), 1.0)
    cw, sw = math.cos(w0), math.sin(w0)
    alpha = sw / (2.0 * 0.707)
    a = math.sqrt(max(float(gain), 1e-6))
    ra = math.sqrt(a)
    s = -1.0 if high else 1.0  # the high shelf is the low shelf with cos(w0) negated in the sums
    a0 = (a + 1.0) + s * (a - 1.0) * cw + 2.0 * ra * alpha
    return np.asarray((
        a * ((a + 1.0) - s * (a - 1.0) * cw + 2.0 * ra * alpha) / a0,
        s * 2.0 * a * ((a - 1.0) - s * (a + 1.0) * cw) / a0,
        a * ((a + 1.0) - s * (a - 1.0) * cw - 2.0 * ra * alpha) / a0,
        -s * 2.0 * ((a - 1.0) + s * (a + 1.0) * cw) / a0,
        ((a + 1.0) + s * (a - 1.0) * cw - 2.0 * ra * alpha) / a0,
    ), dtype=np.float64)


def _low_shelf_coefficients(cutoff: float, gain: float, fs: int) -> np.ndarray:
    return _shelf(cutoff, gain, fs, high=False)


def _high_shelf_coefficients(cutoff: float, gain: float, fs: int) -> np.ndarray:
    return _shelf(cutoff, gain, fs, high=True)


def _peaking_coefficients(low: float, high: float, gain: float, fs: int) -> np.ndarray:
    """RBJ peaking section at sqrt(low high), bandwidth high - low: gain at the
    centre = gain (steam_rt.py:2381-2395)."""
    centre = math.sqrt(max(float(low) * float(high), 1e-12))
    w0 = 2.0 * math.pi * centre / max(float(fs), 1.0)
    cw, sw = math.cos(w0), math.sin(w0)
    alpha = sw * ((float(high) - float(low)) / centre) / 2.0
    a = math.sqrt(max(float(gain), 1e-6))
    a0 = 1.0 + alpha / a
    return np.asarray(((1.0 + alpha * a) / a0, -2.0 * cw / a0, (1.0 - alpha * a) / a0, -2.0 * cw / a0,
                       (1.0 - alpha / a) / a0), dtype=np.float64)


def _steam_multiband_filter_coefficients(gains: np.ndarray, fs: int) -> np.ndarray:
    """[6,5] sections: low shelf, four peaking, high shelf (steam_rt.py:2398-2414)."""
    g = np.asarray(gains, dtype=np.float64).reshape(_NUM_BANDS)
    nyquist = max(1.0, 0.5 * float(fs))
    out = np.empty((_NUM_BANDS, 5), dtype=np.float64)
    for b, (lo, hi) in enumerate(_band_edges(fs)):
        if b == 0:
            out[b] = _low_shelf_coefficients(float(np.clip(hi, 1e-3, nyquist * 0.999)), g[b], fs)
        elif b == _NUM_BANDS - 1:
            out[b] = _high_shelf_coefficients(float(np.clip(lo, 1e-3, nyquist * 0.999)), g[b], fs)
        else:
            lower = float(np.clip(lo, 1e-3, nyquist * 0.998))
            upper = float(np.clip(hi, lower + 1e-3, nyquist * 0.999))
            out[b] = _peaking_coefficients(lower, upper, g[b], fs)
    return out


def _cascade_response(coeffs: np.ndarray, z: np.ndarray) -> np.ndarray:
    """Complex response of a biquad cascade at z = exp(-j w)."""
    h = np.ones(z.shape, dtype=np.complex128)
    for b0, b1, b2, a1, a2 in coeffs:
        h *= (b0 + b1 * z + b2 * z * z) / (1.0 + a1 * z + a2 * z * z)
    return h


def _band_centres(fs: int) -> np.ndarray:
    return np.minimum(np.asarray([float(b) for b in FREQUENCY_BANDS]), 0.9 * 0.5 * max(float(fs), 1.0))


def _steam_multiband_response_matrix(fs: int) -> np.ndarray:
    """R[c, s] = log|H(centre c)| / log 0.5 of the cascade with section s at gain
    0.5 and the others at 1 (steam_rt.py:2417-2437)."""
    z = np.exp(-2.0j * math.pi * _band_centres(fs) / max(float(fs), 1.0))
    out = np.empty((_NUM_BANDS, _NUM_BANDS), dtype=np.float64)
    for s in range(_NUM_BANDS):
        gains = np.ones(_NUM_BANDS)
        gains[s] = 0.5
        h = _cascade_response(_steam_multiband_filter_coefficients(gains, fs), z)
        out[:, s] = np.log(np.maximum(np.abs(h), 1e-12)) / math.log(0.5)
    return out


def _steam_compensated_multiband_coefficients(gains: np.ndarray, fs: int,
                                              response: np.ndarray | None = None) -> np.ndarray:
    """Section gains solving R log g = log(desired), clipped to [1e-3, 4]
    (steam_rt.py:2440-2455)."""
    desired = np.clip(np.asarray(gains, dtype=np.float64).reshape(_NUM_BANDS), 1e-3, 0.999999)
    r = response if response is not None else _steam_multiband_response_matrix(fs)
    try:
        log_g = np.linalg.solve(r, np.log(desired))
    except np.linalg.LinAlgError:
        log_g = np.linalg.pinv(r) @ np.log(desired)
    return _steam_multiband_filter_coefficients(np.clip(np.exp(log_g), 1e-3, 4.0), fs)


def _band_rt60(rt60_bands: Mapping[str, float]) -> np.ndarray:
    return np.asarray([max(0.1, float(rt60_bands.get(b, 0.0) or 0.0)) for b in FREQUENCY_BANDS])


def _steam_fdn_filter_coefficients(delays: np.ndarray, rt60_bands: Mapping[str, float],
                                   fs: int) -> tuple[np.ndarray, np.ndarray]:
    """Absorptive [L,6,5] and tone [6,5] cascades (steam_rt.py:2458-2470)."""
    rt60 = _band_rt60(rt60_bands)
    r = _steam_multiband_response_matrix(fs)
    absorptive = np.empty((len(delays), _NUM_BANDS, 5))
    for i, d in enumerate(np.asarray(delays, dtype=np.float64)):
        per_pass = np.maximum(np.exp(-(6.91 * d) / (rt60 * max(float(fs), 1.0))), 1e-3)
        absorptive[i] = _steam_compensated_multiband_coefficients(per_pass, fs, r)
    tone = np.sqrt(1.0 / rt60)
    tone /= max(float(np.max(tone)), 1e-12)
    return absorptive, _steam_multiband_filter_coefficients(tone, fs)


def fdn_impulse(delays: np.ndarray, absorptive: np.ndarray, num_samples: int, allpass_delays: np.ndarray,
                tone: np.ndarray) -> np.ndarray:
    """The FDN's impulse response (``aa_fdn_impulse``), float64 [num_samples]."""
    d = np.ascontiguousarray(delays, dtype=np.int64)
    ab = np.ascontiguousarray(absorptive, dtype=np.float64)
    ap = np.ascontiguousarray(allpass_delays, dtype=np.int64)
    tn = np.ascontiguousarray(tone, dtype=np.float64)
    if ab.shape[:1] != d.shape or ab.ndim != 3 or ab.shape[2] != 5 or tn.shape != (ab.shape[1], 5):
        raise ValueError("absorptive must be [lines, bands, 5] and tone [bands, 5]")
    out = np.empty(int(num_samples), dtype=np.float64)
    p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    _lib.check(_lib.load().aa_fdn_impulse(p(d), int(d.size), p(ab), int(ab.shape[1]), int(num_samples), p(ap),
                                          int(ap.size), p(tn), p(out)))
    return out


def _late_tail_start_power(signal: np.ndarray, start_sample: int, rt60: float, fs: int, bin_samples: int) -> float:
    """Power at ``start_sample``: the larger of the mean power over the next
    bin and the median, over the last five non-empty bins before it, of the
    bin power decayed at ``rt60`` to ``start_sample`` (steam_rt.py:2473-2497)."""
    sq = np.square(np.asarray(signal, dtype=np.float64))
    start = min(max(0, int(start_sample)), sq.size)
    ahead = sq[start:min(sq.size, start + bin_samples)]
    power = float(np.mean(ahead)) if ahead.size else 0.0
    if start <= 0:
        return power
    nb = int(math.ceil(start / bin_samples))
    padded = np.zeros(nb * bin_samples)
    padded[:start] = sq[:start]
    energy = padded.reshape(nb, bin_samples).sum(axis=1)
    nonzero = np.flatnonzero(energy > 1e-18)
    if nonzero.size == 0:
        return power
    rate = 6.0 * math.log(10.0) / max(float(rt60), 0.1)
    est = [(energy[i] / bin_samples) * math.exp(-rate * max(0.0, (start - (i + 0.5) * bin_samples) / max(float(fs), 1.0)))
           for i in nonzero[-min(5, nonzero.size):]]
    return max(power, float(np.median(est)))


def _render_parametric_fdn_late_reverb(
    traced_band_irs: np.ndarray,
    field: Mapping[str, Any],
    rt60_bands: Mapping[str, float],
    config: SimConfig,
    *,
    decay_profiles: Mapping[str, Any] | None = None,
) -> tuple[np.ndarray, np.ndarray, dict[str, Any]]:
    """(rendered [6,n], FDN component [6,n], metadata): the traced band
    signals crossfaded into the calibrated FDN tail (steam_rt.py:2500-2634)."""
    traced = np.asarray(traced_band_irs, dtype=np.float64)
    if traced.ndim != 2 or traced.shape[0] != _NUM_BANDS:
        raise ValueError(f"expected {_NUM_BANDS} traced band IRs, got {traced.shape}")
    fs = max(1, int(config.fs))
    n = int(traced.shape[1])
    duration_s = n / fs
    T = min(max(1.0 / fs, float(config.hybrid_transition_s)), max(1.0 / fs, (n - 1) / fs))
    overlap = float(np.clip(config.hybrid_overlap_fraction, 0.0, 0.95))
    start_s = (1.0 - overlap) * T
    start0 = min(n, max(0, int(math.floor(start_s * fs))))
    starts = np.full(_NUM_BANDS, start0, dtype=np.int64)
    ends = np.full(_NUM_BANDS, min(n, max(start0 + 1, int(math.ceil(T * fs)))), dtype=np.int64)
    by_band: dict[str, dict[str, Any]] = {}
    profiles = decay_profiles if isinstance(decay_profiles, Mapping) else {}
    for b, band in enumerate(FREQUENCY_BANDS):
        prof = profiles.get(band)
        chosen = str(prof.get("selected_target_source", "")) if isinstance(prof, Mapping) else ""
        brk = float(prof.get("transition_time_s", 0.0) or 0.0) if isinstance(prof, Mapping) else 0.0
        if chosen == "fitted_late_slope" and 0.0 < brk < duration_s:
            starts[b] = min(n - 1, max(0, int(math.floor((1.0 - overlap) * brk * fs))))
            ends[b] = min(n, max(int(starts[b]) + 1, int(math.ceil(brk * fs))))
            model = "coupled_space_breakpoint"
        else:
            model = "configured_hybrid_transition"
        by_band[band] = {"start_s": round(float(starts[b]) / fs, 6), "end_s": round(float(ends[b]) / fs, 6),
                         "model": model}
    first, last = int(np.min(starts)), int(np.max(ends))
    if not config.late_tail or first >= n:
        return traced.astype(np.float32), np.zeros_like(traced, dtype=np.float32), {
            "applied": False, "added": False,
            "reason": "disabled" if not config.late_tail else "tail_start_outside_rir",
            "model": MODEL, "transition_start_s": round(start_s, 6), "transition_end_s": round(T, 6),
            "overlap_fraction": overlap,
        }

    delays = _steam_fdn_delays(fs, config.seed)
    allpass = np.asarray(ALLPASS_DELAYS, dtype=np.int64)
    rt60_used = {band: round(float(rt60_bands.get(band, 0.0) or 0.0), 4) for band in FREQUENCY_BANDS
                 if float(rt60_bands.get(band, 0.0) or 0.0) > 0.0}
    absorptive, tone = _steam_fdn_filter_coefficients(delays, rt60_bands, fs)
    # run the network through its longest delay and the all-passes first, so
    # the tail starts from populated states (a breakpoint may come early)
    preroll = int(np.max(delays) + np.sum(allpass))
    raw = fdn_impulse(delays, absorptive, n + preroll, allpass, tone)[preroll:]
    fdn = np.asarray(bandlimit_band_signals(np.repeat(raw[None, :], _NUM_BANDS, axis=0), fs), dtype=np.float64)

    bin_samples = max(1, int(round(float(field.get("bin_duration_s", config.rt_bin_duration_s)) * fs)))
    scales, anchor_power, fdn_energy_by_band = {}, {}, {}
    for b, band in enumerate(FREQUENCY_BANDS):
        rt = float(rt60_bands.get(band, 0.0) or 0.0)
        if rt <= 0.0:
            continue
        s0 = int(starts[b])
        anchor = int(ends[b]) if by_band[band]["model"] == "coupled_space_breakpoint" else s0
        power = _late_tail_start_power(traced[b], anchor, rt, fs, bin_samples)
        rate = 6.0 * math.log(10.0) / max(rt, 0.1)
        target = float(np.sum(power * np.exp(-rate * np.arange(n - anchor, dtype=np.float64) / fs)))
        unit = float(np.sum(np.square(fdn[b, anchor:])))
        scale = math.sqrt(target / max(unit, 1e-24)) if target > 0.0 else 0.0
        fdn[b, s0:] *= scale
        scales[band] = round(scale, 8)
        anchor_power[band] = round(power, 14)
        by_band[band]["anchor_s"] = round(float(anchor) / fs, 6)
        fdn_energy_by_band[band] = round(float(np.sum(np.square(fdn[b, s0:]))), 12)

    out = traced.copy()
    component = np.zeros_like(traced)
    traced_tail = 0.0
    for b in range(_NUM_BANDS):
        s0, e0 = int(starts[b]), int(ends[b])
        ramp = max(1, e0 - s0)
        fade = np.linspace(1.0, 0.0, ramp, endpoint=False, dtype=np.float64)  # traced weight^2 over the ramp
        out[b, s0:e0] *= np.sqrt(fade)
        if e0 < n:
            out[b, e0:] = 0.0
        w = np.ones(n - s0)
        w[:ramp] = np.sqrt(1.0 - fade)
        component[b, s0:] = fdn[b, s0:] * w
        traced_tail += float(np.sum(np.square(traced[b, s0:])))
    out += component
    rendered_tail = sum(float(np.sum(np.square(out[b, int(starts[b]):]))) for b in range(_NUM_BANDS))
    fdn_energy = float(np.sum(np.square(component)))
    return out.astype(np.float32), component.astype(np.float32), {
        "applied": bool(rt60_used and fdn_energy > 1e-18),
        "changed": bool(not np.allclose(out[:, first:], traced[:, first:], rtol=1e-7, atol=1e-12)),
        "added": bool(fdn_energy > 1e-18),
        "model": MODEL,
        "rt60_source": "steam_audio_reverb_estimator_from_traced_energy_field",
        "calibration": "steady-state pre-rolled feedback-loop multiband absorption, traced cutoff power, "
                       "duration-integrated energy normalization, and equal-power hybrid crossfade",
        "transition_start_s": round(float(first) / fs, 6),
        "transition_end_s": round(float(last) / fs, 6),
        "configured_transition_end_s": round(T, 6),
        "transition_by_band": by_band,
        "overlap_fraction": overlap,
        "delay_line_count": int(delays.size),
        "delay_samples": [int(v) for v in delays],
        "allpass_delay_samples": [int(v) for v in allpass],
        "preroll_samples": preroll,
        "traced_tail_energy": round(traced_tail, 12),
        "fdn_tail_energy": round(fdn_energy, 12),
        "rendered_tail_energy": round(rendered_tail, 12),
        "anchor_power_by_band": anchor_power,
        "scale_by_band": scales,
        "fdn_energy_by_band": fdn_energy_by_band,
        "rt60_bands": rt60_used,
    }
