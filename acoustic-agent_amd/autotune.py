"""Autotuner of launch shapes, cached per device architecture and size class.

Tunables (none changes results, only speed):
  batch:  workgroup size, rays per work item (packet), persistent
          workgroups per CU;
  single: workgroup size, LDS-privatised histogram, persistent workgroups
          per CU.
Each candidate is timed with HIP events on a sample of the real work; the
winner is stored under (arch, size class) in memory and, if ``cache_path`` is
given, in a JSON file.  A size class is (log2 of triangles, log2 of rays).
"""
from __future__ import annotations

import itertools
import json
import math
import threading
from pathlib import Path
from typing import Any, Callable, Sequence

import numpy as np

from .batch import BatchEngine, BatchTuning, TraceRequest

BATCH_SPACE = {
    "block_size": (128, 256),
    "packet_rays": (0, 1024, 4096),
    "workgroups_per_cu": (0,),
}
TRACE_SPACE = {
    "block_size": (64, 128, 256),
    "lds_histogram": (0, 1),
    "workgroups_per_cu": (0,),
}


def size_class(triangles: int, rays: int) -> str:
    return f"t{int(math.log2(max(triangles, 1)))}_r{int(math.log2(max(rays, 1)))}"


def _grid(space: dict[str, Sequence[int]]) -> list[dict[str, int]]:
    keys = list(space)
    return [dict(zip(keys, vals)) for vals in itertools.product(*(space[k] for k in keys))]


class Autotuner:
    def __init__(self, device: int = 0, cache_path: str | Path | None = None) -> None:
        from .hip_backend import device_info

        self.device = int(device)
        self.arch = device_info(device)["arch"]
        self.cache_path = Path(cache_path) if cache_path else None
        self._lock = threading.Lock()
        self.cache: dict[str, dict[str, Any]] = {}
        if self.cache_path and self.cache_path.exists():
            try:
                self.cache = json.loads(self.cache_path.read_text())
            except ValueError:
                self.cache = {}
        self.log: list[tuple[str, dict[str, int], float]] = []

    def _store(self, kind: str, cls: str, value: dict[str, Any]) -> None:
        with self._lock:
            self.cache.setdefault(self.arch, {})[f"{kind}:{cls}"] = value
            if self.cache_path:
                self.cache_path.write_text(json.dumps(self.cache, indent=1, sort_keys=True))

    def lookup(self, kind: str, cls: str) -> dict[str, Any] | None:
        return self.cache.get(self.arch, {}).get(f"{kind}:{cls}")

    def tune_batch(self, requests: Sequence[TraceRequest], *, space: dict[str, Sequence[int]] | None = None,
                   sample: int = 48, reps: int = 2, engine: BatchEngine | None = None,
                   force: bool = False, progress: Callable[[str], None] | None = None) -> BatchTuning:
        """Best BatchTuning for batches like ``requests`` (timed on a sample).

        ``progress`` receives one line per timed candidate.
        """
        import torch

        tri = int(np.median([r.scene.info["num_triangles"] for r in requests]))
        rays = int(np.median([r.num_rays for r in requests]))
        cls = size_class(tri, rays)
        cached = None if force else self.lookup("batch", cls)
        if cached is not None:
            return BatchTuning(**cached)
        rng = np.random.default_rng(0)
        idx = rng.choice(len(requests), size=min(sample, len(requests)), replace=False)
        subset = [requests[int(i)] for i in sorted(idx)]
        own = engine is None
        engine = engine or BatchEngine(self.device, max_traces=max(len(subset), 1))
        best, best_t = None, float("inf")
        stream = torch.cuda.current_stream(self.device)
        for cand in _grid(space or BATCH_SPACE):
            tuning = BatchTuning(**cand)
            engine.submit(subset, tuning).wait()  # warm-up (and LDS attribute setup)
            times = []
            for _ in range(reps):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                res = engine.submit(subset, tuning, stream=stream)
                b.record(stream)
                res.wait()
                times.append(a.elapsed_time(b))
            t = float(np.median(times))
            self.log.append(("batch:" + cls, cand, t))
            if progress is not None:
                progress(f"autotune batch {cls} {cand}: {t:.2f} ms")
            if t < best_t:
                best, best_t = cand, t
        if own:
            engine.close()
        self._store("batch", cls, dict(best))
        return BatchTuning(**best)

    def tune_trace(self, run_one, triangles: int, rays: int, *, space: dict[str, Sequence[int]] | None = None,
                   reps: int = 3, force: bool = False) -> dict[str, int]:
        """Best single-trace tuning; ``run_one(tuning) -> milliseconds``."""
        cls = size_class(triangles, rays)
        cached = None if force else self.lookup("trace", cls)
        if cached is not None:
            return dict(cached)
        best, best_t = None, float("inf")
        for cand in _grid(space or TRACE_SPACE):
            run_one(cand)
            t = float(np.median([run_one(cand) for _ in range(reps)]))
            self.log.append(("trace:" + cls, cand, t))
            if t < best_t:
                best, best_t = cand, t
        self._store("trace", cls, dict(best))
        return dict(best)
