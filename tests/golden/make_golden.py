"""Regenerate the committed fixtures of tests/golden/.

philox_rocrand.npz: Philox4x32-10 outputs of rocRAND's engine
    (philox_rocrand.cpp, built with hipcc on the host).
ray_sets.npz: the reference ray set for N = 8, B = 3, seed 1729, drawn with
    NumPy exactly as acoustic_agent/steam_rt.py:3958-3968 and 4009-4044 do
    (directions with the default_rng(seed) jitter; per bounce random(N) then
    integers(0, 128, N) from default_rng(seed + 1)).
"""
from __future__ import annotations

import math
import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def philox() -> None:
    with tempfile.TemporaryDirectory() as tmp:
        exe = Path(tmp) / "philox"
        subprocess.run(["hipcc", "-O1", str(HERE / "philox_rocrand.cpp"), "-o", str(exe)], check=True)
        rows = [list(map(int, line.split())) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                                         text=True).stdout.splitlines()]
    a = np.asarray(rows, dtype=np.uint64).astype(np.uint32)
    np.savez(HERE / "philox_rocrand.npz", counters=a[:, 0:4], keys=a[:, 4:6], outputs=a[:, 6:10])


def ray_sets() -> None:
    n, b, m, seed = 8, 3, 128, 1729
    rng = np.random.default_rng(seed)
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(np.clip(1.0 - z * z, 0.0, 1.0))
    theta = i * (math.pi * (3.0 - math.sqrt(5.0))) + rng.uniform(-0.02, 0.02, n)
    directions = np.stack([np.cos(theta) * r, np.sin(theta) * r, z], axis=1)
    rng = np.random.default_rng(seed + 1)
    U = np.empty((b, n))
    J = np.empty((b, n), dtype=np.int32)
    for k in range(b):
        U[k] = rng.random(n)
        J[k] = rng.integers(0, m, size=n, dtype=np.int32)
    np.savez(HERE / "ray_sets.npz", directions_n8_s1729=directions, random_n8_b3_s1729=U, indices_n8_b3_s1729=J)


if __name__ == "__main__":
    philox()
    ray_sets()
