"""Host ray sets of the reference ("reference" RNG mode).

Restates the reference samplers so that a trace can be fed exactly the ray
set the reference would draw:
  sphere_samples           acoustic_agent/steam_rt.py:3958-3968
  diffuse_sample_bank      acoustic_agent/steam_rt.py:3990-4006
  diffuse_random_sequence  acoustic_agent/steam_rt.py:4009-4044 (call order:
                           per bounce rng.random(N) then rng.integers(0, M, N))
The "counter" mode generates the same distributions on the device from
Philox4x32-10 instead and needs none of these arrays except the bank.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np


@lru_cache(maxsize=8)
def sphere_samples(n: int, seed: int) -> np.ndarray:
    n = max(1, int(n))
    rng = np.random.default_rng(seed)
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(np.clip(1.0 - z * z, 0.0, 1.0))
    theta = i * (math.pi * (3.0 - math.sqrt(5.0))) + rng.uniform(-0.02, 0.02, n)
    samples = np.stack([np.cos(theta) * r, np.sin(theta) * r, z], axis=1)
    samples.setflags(write=False)
    return samples


@lru_cache(maxsize=16)
def diffuse_sample_bank(count: int) -> np.ndarray:
    count = max(1, int(count))
    bits = np.arange(count, dtype=np.uint64)
    inverse = np.zeros(count, dtype=np.float64)
    factor = 0.5
    while np.any(bits):
        inverse += factor * (bits & np.uint64(1))
        bits >>= np.uint64(1)
        factor *= 0.5
    u = (np.arange(count, dtype=np.float64) + 0.5) / count
    r = np.sqrt(u)
    theta = 2.0 * math.pi * inverse
    samples = np.stack([r * np.cos(theta), r * np.sin(theta), np.sqrt(np.clip(1.0 - u, 0.0, 1.0))], axis=1)
    samples.setflags(write=False)
    return samples


def diffuse_random_sequence(num_rays: int, num_bounces: int, sample_count: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    rng = np.random.default_rng(int(seed) + 1)
    random_values = np.empty((num_bounces, num_rays), dtype=np.float64)
    sample_indices = np.empty((num_bounces, num_rays), dtype=np.int32)
    for bounce in range(num_bounces):
        random_values[bounce] = rng.random(num_rays)
        sample_indices[bounce] = rng.integers(0, sample_count, size=num_rays, dtype=np.int32)
    random_values.setflags(write=False)
    sample_indices.setflags(write=False)
    return random_values, sample_indices
