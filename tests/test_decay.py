"""Decay profiles and late-tail targets (acoustic-agent_amd/decay.py).

Known answers of the reference's own tests
(tests/test_steam_audio_parity.py:70-92, 100-151, 154-167) and closed forms:
an energy decay curve built piecewise linear in dB, so each fitted line has
a known slope, break and level.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from acoustic_agent_amd import rir, steam_rt
from acoustic_agent_amd.decay import (
    apply_coupled_late_reverb_prior,
    energy_decay_profile,
    estimate_late_reverb_times,
)
from acoustic_agent_amd.models import FREQUENCY_BANDS, SimConfig


def _energy_from_edc_db(edc_db: np.ndarray) -> np.ndarray:
    """Bin energies whose backward integral is exactly 10^(edc_db/10)."""
    edc = np.power(10.0, np.asarray(edc_db, dtype=np.float64) / 10.0)
    return np.maximum(0.0, np.concatenate((edc[:-1] - edc[1:], edc[-1:])))


def _two_slopes(times, t_break, rate1, rate2):
    return np.where(times <= t_break, -rate1 * times, -rate1 * t_break - rate2 * (times - t_break))


@pytest.fixture
def no_air(monkeypatch):
    for band in FREQUENCY_BANDS:
        monkeypatch.setitem(steam_rt.AIR_ABSORPTION_NP_PER_M, band, 0.0)


def test_single_exponential_keeps_one_slope():
    # test_steam_audio_parity.py:100-108
    times = np.arange(200) * 0.01
    profile = energy_decay_profile(_energy_from_edc_db(-60.0 * times), 0.01)
    assert profile["model"] == "single_slope" and len(profile["segments"]) == 1
    assert profile["segments"][0]["equivalent_rt60_s"] == pytest.approx(1.0, rel=0.01)


def test_significant_coupled_space_break_is_detected():
    # test_steam_audio_parity.py:111-125
    times = np.arange(200) * 0.01
    profile = energy_decay_profile(_energy_from_edc_db(_two_slopes(times, 0.25, 60.0, 120.0)), 0.01)
    assert profile["model"] == "double_slope"
    assert profile["transition_time_s"] == pytest.approx(0.25, abs=0.02)
    assert profile["segments"][0]["equivalent_rt60_s"] == pytest.approx(1.0, rel=0.03)
    assert profile["segments"][1]["equivalent_rt60_s"] == pytest.approx(0.5, rel=0.03)


@pytest.mark.parametrize("t_break, rate1, rate2", [(0.3, 60.0, 150.0), (0.15, 150.0, 50.0), (0.4, 45.0, 90.0)])
def test_two_lines_recover_the_constructed_slopes_break_and_levels(t_break, rate1, rate2):
    """Exact piecewise-linear decay curves with the break 17.5-22.5 dB down
    (both sides span well over 7 dB of the -5 .. -35 dB range): each side is
    a straight line in dB, so the fitted lines are exact (r^2 = 1), their
    RT60s are 60 / rate and the late side starts at the break."""
    bin_dur = 0.005
    times = np.arange(400) * bin_dur
    db = _two_slopes(times, t_break, rate1, rate2)
    profile = energy_decay_profile(_energy_from_edc_db(db), bin_dur)
    assert profile["model"] == "double_slope"
    early, late = profile["segments"]
    assert (early["label"], late["label"]) == ("early", "late")
    assert early["equivalent_rt60_s"] == pytest.approx(60.0 / rate1, rel=1e-3)
    assert late["equivalent_rt60_s"] == pytest.approx(60.0 / rate2, rel=1e-3)
    assert early["r2"] == pytest.approx(1.0, abs=1e-4) and late["r2"] == pytest.approx(1.0, abs=1e-4)
    fit = (db <= -5.0) & (db >= -35.0)
    xs = times[fit]
    assert early["start_time_s"] == pytest.approx(xs[0], abs=1e-9)
    assert late["end_time_s"] == pytest.approx(xs[-1], abs=1e-9)
    assert t_break - bin_dur <= profile["transition_time_s"] <= t_break + 2 * bin_dur
    assert profile["slope_order"] == ("early_slower" if rate1 < rate2 else "late_slower")
    assert profile["slope_ratio"] == pytest.approx(max(rate1, rate2) / min(rate1, rate2), rel=1e-3)


def test_breaks_below_the_thresholds_stay_single():
    bin_dur, times = 0.01, np.arange(300) * 0.01
    # slopes within a factor 1.3 (< 1.35) of each other
    near = energy_decay_profile(_energy_from_edc_db(_two_slopes(times, 0.3, 60.0, 78.0)), bin_dur)
    assert near["model"] == "single_slope" and near["candidate_slope_ratio"] == pytest.approx(1.3, rel=1e-3)
    # a break 3 dB into the fit range leaves the early side under 7 dB
    short = energy_decay_profile(_energy_from_edc_db(_two_slopes(times, 0.08, 100.0, 40.0)), bin_dur)
    assert short["model"] == "single_slope"


def test_unavailable_profiles():
    assert energy_decay_profile(np.zeros(100), 0.01)["model"] == "unavailable"
    assert energy_decay_profile(np.ones(1), 0.01)["model"] == "unavailable"
    # a decay spanning only 10 dB of the -5 .. -35 dB range
    times = np.arange(100) * 0.01
    shallow = energy_decay_profile(_energy_from_edc_db(-15.0 * times), 0.01)
    assert shallow["model"] == "unavailable" and shallow["segments"] == []


def test_low_energy_cross_room_decay_uses_the_wide_range_fit(no_air):
    # test_steam_audio_parity.py:70-92
    times = np.arange(200) * 0.01
    energy = 1e-8 * np.power(10.0, -6.0 * times / 0.8)
    field = {"echogram": np.repeat(energy[None, :], 6, axis=0), "bin_duration_s": 0.01, "num_bins": energy.size}
    targets, profiles = estimate_late_reverb_times(field, SimConfig(), fallback={b: 0.0 for b in FREQUENCY_BANDS})
    for band in FREQUENCY_BANDS:
        assert targets[band] == pytest.approx(0.8, rel=0.03)
        assert profiles[band]["selected_target_source"] == "fitted_wide_range_single_slope"


def test_coupled_space_tail_target_uses_the_fitted_late_slope(no_air):
    # test_steam_audio_parity.py:128-151
    times = np.arange(200) * 0.01
    energy = _energy_from_edc_db(_two_slopes(times, 0.2, 100.0, 50.0))
    field = {"echogram": np.tile(energy, (6, 1)), "num_bins": energy.size, "bin_duration_s": 0.01}
    targets, profiles = estimate_late_reverb_times(field, SimConfig(), fallback={b: 0.6 for b in FREQUENCY_BANDS})
    for band in FREQUENCY_BANDS:
        assert profiles[band]["model"] == "double_slope"
        assert profiles[band]["selected_target_source"] == "fitted_late_slope"
        assert targets[band] == pytest.approx(1.2, rel=0.03)


def test_single_slope_band_keeps_its_schroeder_rt60(no_air):
    times = np.arange(300) * 0.01
    energy = _energy_from_edc_db(-60.0 * times / 0.7) * math.sqrt(4.0 * math.pi)
    field = {"echogram": np.tile(energy, (6, 1)), "num_bins": energy.size, "bin_duration_s": 0.01}
    standard = steam_rt.estimate_reverb_times(field, SimConfig())
    targets, profiles = estimate_late_reverb_times(field, SimConfig())
    for band in FREQUENCY_BANDS:
        assert targets[band] == standard[band] and standard[band] == pytest.approx(0.7, rel=1e-6)
        assert profiles[band]["selected_target_source"] == "steam_single_slope"


def test_coupled_room_prior_replaces_only_the_late_tail_target():
    # test_steam_audio_parity.py:154-167
    traced = {band: 0.7 + 0.1 * i for i, band in enumerate(FREQUENCY_BANDS)}
    prior = {band: 1.4 + 0.05 * i for i, band in enumerate(FREQUENCY_BANDS)}
    profiles = {band: {"selected_target_source": "fitted_late_slope", "selected_target_rt60_s": traced[band]}
                for band in FREQUENCY_BANDS}
    targets, updated = apply_coupled_late_reverb_prior(traced, profiles, prior)
    assert targets == prior
    for band in FREQUENCY_BANDS:
        assert updated[band]["traced_target_rt60_s"] == pytest.approx(traced[band])
        assert updated[band]["selected_target_source"] == "coupled_room_energy_matrix"
    # no prior, or a non-positive one: the traced targets stay
    assert apply_coupled_late_reverb_prior(traced, profiles, None)[0] == traced
    bad = dict(prior, **{FREQUENCY_BANDS[0]: 0.0})
    assert apply_coupled_late_reverb_prior(traced, profiles, bad)[0][FREQUENCY_BANDS[0]] == traced[FREQUENCY_BANDS[0]]


def test_multi_room_render_starts_the_crossfade_at_the_break(no_air):
    """A coupled-space echogram rendered as a multi-room scene: the FDN tail
    targets the late slope and its crossfade ends at the fitted break; as a
    single room it keeps the configured transition."""
    bin_dur = 0.005
    times = np.arange(400) * bin_dur
    energy = _energy_from_edc_db(_two_slopes(times, 0.15, 120.0, 50.0)) * 1e3  # break 18 dB down
    field = {"echogram": np.tile(energy, (6, 1)), "num_bins": energy.size, "bin_duration_s": bin_dur,
             "direct_delay_s": 0.0}
    config = SimConfig(fs=8000, seed=2, hybrid_transition_s=0.3)
    bands, meta = rir.render_band_rirs(field, config, multi_room=True)
    assert bands.shape[0] == 6 and meta["applied"]
    for band in FREQUENCY_BANDS:
        profile = meta["late_decay_profiles"][band]
        assert profile["selected_target_source"] == "fitted_late_slope"
        assert profile["selected_target_rt60_s"] == pytest.approx(1.2, rel=0.03)
        by_band = meta["transition_by_band"][band]
        assert by_band["model"] == "coupled_space_breakpoint"
        assert by_band["end_s"] == pytest.approx(profile["transition_time_s"], abs=1.0 / 8000)
    _, single = rir.render_band_rirs(field, config)
    assert "late_decay_profiles" not in single
    assert {v["model"] for v in single["transition_by_band"].values()} == {"configured_hybrid_transition"}


def test_names_match_the_reference_module():
    assert steam_rt._energy_decay_profile is energy_decay_profile
    assert steam_rt.estimate_late_reverb_times is estimate_late_reverb_times
    assert steam_rt._apply_coupled_late_reverb_prior is apply_coupled_late_reverb_prior
