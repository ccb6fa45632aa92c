"""MI355X-native reflection tracer (six-band Monte-Carlo energy field).

A drop-in backend for the reference engine's tracer selection
(``trace_energy_field``, acoustic_agent/steam_rt.py:1757-1804): the host code
here prepares scenes and ray sets and calls the hand-written gfx950 kernels of
``libaa_tracer.so`` through its C ABI (``include/aa_tracer.h``).

This directory is imported as the package ``acoustic_agent_amd``; the
``load_package`` helper of the repository's entry points registers it.
"""
from __future__ import annotations

from .models import FREQUENCY_BANDS, SimConfig

__all__ = ["FREQUENCY_BANDS", "SimConfig"]
__version__ = "0.1.0"
