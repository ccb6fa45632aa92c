"""Pins of the echogram -> band RT60 / EDT metric (steam_rt.py:2727-2742,
3052-3067, 2196-2202; EDT is the ISO 3382-1 definition, fit on 0..-10 dB).

The expected values come from closed forms: exact exponential decays, an
energy decay curve built piecewise linear in dB (so each fit range sees a
known straight line), and NumPy's own least squares on the known curve.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from acoustic_agent_amd import steam_rt
from acoustic_agent_amd.models import FREQUENCY_BANDS, SimConfig
from acoustic_agent_amd.steam_rt import (
    _schroeder_fit_edt,
    _schroeder_fit_rt60,
    estimate_early_decay_times,
    estimate_reverb_times,
)


def _energy_from_edc_db(edc_db: np.ndarray) -> np.ndarray:
    """Bin energies whose backward integral is exactly 10^(edc_db/10)."""
    edc = np.power(10.0, np.asarray(edc_db, dtype=np.float64) / 10.0)
    return np.concatenate((edc[:-1] - edc[1:], edc[-1:]))


def test_rt60_of_an_exponential_matches_the_reference_known_answer():
    # tests/test_steam_audio_parity.py:43-51
    times = np.arange(400, dtype=np.float64) * 0.01
    energy = np.power(10.0, -6.0 * times / 1.0)
    assert _schroeder_fit_rt60(energy, 0.01) == pytest.approx(1.0, rel=0.01)


@pytest.mark.parametrize("T", [0.35, 1.0, 2.4])
def test_edt_equals_rt60_on_a_single_slope(T):
    bin_dur = 1e-3
    times = np.arange(int(3 * T / bin_dur)) * bin_dur
    energy = _energy_from_edc_db(-60.0 * times / T)
    assert _schroeder_fit_edt(energy, bin_dur) == pytest.approx(T, rel=1e-9)
    assert _schroeder_fit_rt60(energy, bin_dur) == pytest.approx(T, rel=1e-9)


def test_edt_and_rt60_differ_on_a_two_slope_decay():
    """EDC: -60 t / T1 down to -15 dB, then -60 t / T2.  The EDT range
    (0..-10 dB) lies on the first line, so EDT = T1 exactly; the RT60 range
    (-5..-25 dB) spans the break, so RT60 is the least-squares slope of the
    known curve over that range."""
    bin_dur, T1, T2 = 1e-3, 0.4, 1.6
    times = np.arange(2000) * bin_dur
    t_break = 15.0 * T1 / 60.0
    db = np.where(times <= t_break, -60.0 * times / T1, -15.0 - 60.0 * (times - t_break) / T2)
    energy = _energy_from_edc_db(db)
    edt = _schroeder_fit_edt(energy, bin_dur)
    rt60 = _schroeder_fit_rt60(energy, bin_dur)
    assert edt == pytest.approx(T1, rel=1e-9)
    fit = (db <= -5.0) & (db >= -25.0)
    slope = np.polyfit(times[fit], db[fit], 1)[0]
    assert rt60 == pytest.approx(-60.0 / slope, rel=1e-9)
    assert T1 < rt60 < T2 and abs(rt60 - edt) / edt > 0.3


def test_silent_and_too_short_decays_give_zero():
    assert _schroeder_fit_rt60(np.zeros(100), 1e-3) == 0.0
    assert _schroeder_fit_edt(np.zeros(100), 1e-3) == 0.0
    # one bin: fewer than two points in the fit range
    assert _schroeder_fit_rt60(np.asarray([1.0]), 1e-3) == 0.0


def test_band_estimators_apply_air_absorption_and_the_energy_floor():
    """An exponential echogram exp(-a t) in every band: after the band's air
    absorption weight exp(-2 m c t) the decay rate is a + 2 m c, so
    RT60 = EDT = 60 / (10 log10(e) (a + 2 m c)).  The echogram is scaled by
    1 / Y00 so that the weighted total stays above the 1e-4 floor; scaled
    below it every band gives 0."""
    config = SimConfig()
    bin_dur, K = 1e-3, 3000
    times = np.arange(K) * bin_dur
    a = 60.0 / (0.9 * 10.0 * math.log10(math.e))  # RT60 0.9 s without air absorption
    energy = np.exp(-a * times) * math.sqrt(4.0 * math.pi)
    field = {"echogram": np.tile(energy, (6, 1)), "bin_duration_s": bin_dur, "num_bins": K}
    rt60 = estimate_reverb_times(field, config)
    edt = estimate_early_decay_times(field, config)
    for band in FREQUENCY_BANDS:
        rate = a + 2.0 * steam_rt.AIR_ABSORPTION_NP_PER_M[band] * config.c
        want = 60.0 / (10.0 * math.log10(math.e) * rate)
        # bins sample the continuous decay: the fits agree to the sampling error
        assert rt60[band] == pytest.approx(want, rel=2e-3), band
        assert edt[band] == pytest.approx(want, rel=2e-3), band
    quiet = dict(field, echogram=field["echogram"] * 1e-9)
    assert set(estimate_reverb_times(quiet, config).values()) == {0.0}
    assert set(estimate_early_decay_times(quiet, config).values()) == {0.0}
