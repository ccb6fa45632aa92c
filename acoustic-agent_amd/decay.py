"""Decay profiles of a traced echogram band, and the late-tail RT60 targets of
multi-room scenes that follow from them.

A band's energy decay curve (backward-integrated energy, in dB relative to
the total) is fitted over -5 .. -35 dB either by one straight line or by two
lines that meet at a break, and the Bayesian information criterion decides
between them (the reference's `_energy_decay_profile`,
acoustic_agent/steam_rt.py:2896-2984): a break is kept only if it improves
the criterion by at least 10, the two slopes differ by a factor of at least
1.35 and both lines fit with r^2 >= 0.95.  Each side of a candidate break
needs at least 5 points spanning at least 7 dB.  For a multi-room scene the
late tail then follows the late line (`estimate_late_reverb_times`,
steam_rt.py:2745-2792), which is also where the parametric FDN tail starts
its coupled-space crossfade.

The profile dictionaries keep the reference's keys, because the FDN tail
and callers of the reference read them.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Any, Mapping

import numpy as np

from .models import FREQUENCY_BANDS, SimConfig

FIT_START_DB = -5.0
FIT_END_DB = -35.0
MIN_POINTS = 10           # points in the fit range for any profile
MIN_RANGE_DB = 14.0       # dB the fit range must span
SIDE_POINTS = 5           # points on each side of a break
SIDE_RANGE_DB = 7.0       # dB each side must span
MIN_BIC_GAIN = 10.0
MIN_SLOPE_RATIO = 1.35
MIN_R2 = 0.95


@dataclass(frozen=True)
class _Line:
    """Least-squares line y = slope x + intercept of a decaying segment."""

    slope: float
    intercept: float
    sse: float
    r2: float

    @property
    def rt60(self) -> float:
        return -60.0 / self.slope

    @staticmethod
    def fit(x: np.ndarray, y: np.ndarray) -> "_Line | None":
        """None for fewer than two points, a degenerate abscissa or a slope
        that does not decay."""
        n = x.size
        if n < 2:
            return None
        dx = x - x.mean()
        dy = y - y.mean()
        sxx = float(dx @ dx)
        if n * sxx <= 1e-12:
            return None
        slope = float(dx @ dy) / sxx
        if slope >= -1e-9:
            return None
        intercept = float(y.mean()) - slope * float(x.mean())
        resid = y - (slope * x + intercept)
        sse = float(resid @ resid)
        return _Line(slope, intercept, sse, 1.0 - sse / max(float(dy @ dy), 1e-15))

    def segment(self, label: str, x: np.ndarray, y: np.ndarray) -> dict[str, Any]:
        return {"label": label, "equivalent_rt60_s": round(self.rt60, 4), "slope_db_per_s": round(self.slope, 4),
                "r2": round(self.r2, 4), "start_time_s": round(float(x[0]), 4), "end_time_s": round(float(x[-1]), 4),
                "start_level_db": round(float(y[0]), 4), "end_level_db": round(float(y[-1]), 4)}


def _bic(sse: float, n: int, params: int) -> float:
    return n * math.log(max(sse / n, 1e-15)) + params * math.log(n)


def _decay_curve_db(energy: np.ndarray) -> np.ndarray:
    edc = np.cumsum(energy[::-1])[::-1]
    return 10.0 * np.log10(np.clip(edc / float(np.sum(energy)), 1e-12, None))


def energy_decay_profile(energy: np.ndarray, bin_dur: float, *, fit_start_db: float = FIT_START_DB,
                         fit_end_db: float = FIT_END_DB) -> dict[str, Any]:
    """One- or two-slope profile of the band energies `energy` (bins of
    `bin_dur` seconds); "unavailable" when the decay is too short, silent or
    does not decay."""
    e = np.asarray(energy, dtype=np.float64).reshape(-1)
    profile: dict[str, Any] = {"model": "unavailable", "segments": [],
                               "fit_range_db": [float(fit_start_db), float(fit_end_db)]}
    if e.size < 2 or float(np.sum(e)) < 1e-12 or bin_dur <= 0.0:
        return profile
    db = _decay_curve_db(e)
    keep = (db <= float(fit_start_db)) & (db >= float(fit_end_db))
    x = float(bin_dur) * np.flatnonzero(keep).astype(np.float64)
    y = db[keep]
    n = int(x.size)
    if n < MIN_POINTS or float(np.ptp(y)) < MIN_RANGE_DB:
        return profile
    whole = _Line.fit(x, y)
    if whole is None:
        return profile
    full = whole.segment("full", x, y)
    profile.update(model="single_slope", segments=[full], single_rt60_s=full["equivalent_rt60_s"],
                   single_r2=full["r2"], dynamic_range_db=round(float(np.ptp(y)), 4))

    # breaks k (first point of the late side) that leave enough points and
    # range on both sides: running extremes from the front and the back
    ks = np.arange(SIDE_POINTS, n - SIDE_POINTS + 1)
    if ks.size == 0:
        return profile
    head = np.maximum.accumulate(y) - np.minimum.accumulate(y)                 # ptp of y[:k + 1]
    tail = (np.maximum.accumulate(y[::-1]) - np.minimum.accumulate(y[::-1]))[::-1]  # ptp of y[k:]
    ks = ks[(head[ks - 1] >= SIDE_RANGE_DB) & (tail[ks] >= SIDE_RANGE_DB)]
    best: tuple[float, int, _Line, _Line] | None = None
    for k in ks.tolist():
        early, late = _Line.fit(x[:k], y[:k]), _Line.fit(x[k:], y[k:])
        if early is None or late is None:
            continue
        score = _bic(early.sse + late.sse, n, 5)
        if best is None or score < best[0]:
            best = (score, k, early, late)
    if best is None:
        return profile
    score, k, early, late = best
    gain = _bic(whole.sse, n, 2) - score
    ratio = max(early.rt60, late.rt60) / max(min(early.rt60, late.rt60), 1e-12)
    profile.update(candidate_bic_improvement=round(gain, 4), candidate_slope_ratio=round(ratio, 4))
    if not (gain >= MIN_BIC_GAIN and ratio >= MIN_SLOPE_RATIO and early.r2 >= MIN_R2 and late.r2 >= MIN_R2):
        return profile
    profile.update(model="double_slope",
                   segments=[early.segment("early", x[:k], y[:k]), late.segment("late", x[k:], y[k:])],
                   transition_time_s=round(float(x[k]), 4), transition_level_db=round(float(y[k]), 4),
                   slope_ratio=round(ratio, 4), bic_improvement=round(gain, 4),
                   slope_order="early_slower" if early.rt60 > late.rt60 else "late_slower")
    return profile


def estimate_late_reverb_times(field: Mapping[str, Any], config: SimConfig, *,
                               fallback: Mapping[str, float] | None = None
                               ) -> tuple[dict[str, float], dict[str, Any]]:
    """Late-tail RT60 target and decay profile per band of a traced field
    (band energies weighted as for RT60: Y00 and air absorption).  The target
    is the late line of a two-slope decay, else the band RT60 (`fallback`, or
    the Schroeder RT60 of the field), else, where that is 0, the single line
    fitted over the wider -5 .. -35 dB range."""
    from .steam_rt import _weighted_bands, estimate_reverb_times

    standard = dict(fallback) if fallback is not None else estimate_reverb_times(field, config)
    bin_dur = float(field["bin_duration_s"])
    targets: dict[str, float] = {}
    profiles: dict[str, Any] = {}
    for band, weighted in zip(FREQUENCY_BANDS, _weighted_bands(field, config)):
        profile = energy_decay_profile(weighted, bin_dur)
        target, source = float(standard.get(band, 0.0) or 0.0), "steam_single_slope"
        late = [s for s in profile["segments"] if s.get("label") == "late"]
        if profile["model"] == "double_slope" and late and late[0]["equivalent_rt60_s"] > 0.0:
            target, source = float(late[0]["equivalent_rt60_s"]), "fitted_late_slope"
        elif target <= 0.0 and profile["model"] == "single_slope" and profile.get("single_rt60_s", 0.0) > 0.0:
            target, source = float(profile["single_rt60_s"]), "fitted_wide_range_single_slope"
        targets[band] = target
        profiles[band] = {**profile, "selected_target_rt60_s": round(target, 4), "selected_target_source": source}
    return targets, profiles


def apply_coupled_late_reverb_prior(traced_targets: Mapping[str, float], profiles: Mapping[str, Any],
                                    prior: Mapping[str, float] | None
                                    ) -> tuple[dict[str, float], dict[str, Any]]:
    """Replace each band's late-tail target by a positive, finite room-system
    prior when one is given (steam_rt.py:2795-2821), keeping the traced
    target in the profile as `traced_target_rt60_s`."""
    targets = {band: float(traced_targets.get(band, 0.0) or 0.0) for band in FREQUENCY_BANDS}
    updated = {band: dict(profiles.get(band, {})) for band in FREQUENCY_BANDS}
    for band in FREQUENCY_BANDS if isinstance(prior, Mapping) else ():
        value = float(prior.get(band, 0.0) or 0.0)
        if math.isfinite(value) and value > 0.0:
            updated[band].update(traced_target_rt60_s=round(targets[band], 4), selected_target_rt60_s=round(value, 4),
                                 selected_target_source="coupled_room_energy_matrix")
            targets[band] = value
    return targets, updated
