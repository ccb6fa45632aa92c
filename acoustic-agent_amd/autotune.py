"""Autotuner of launch shapes, cached per device architecture and size class.

Tunables (none changes results, only speed):
  batch:  rays per work item (packet), workgroup size, persistent workgroups
          per CU, and for batches with multi-source traces the sources that
          share one lane's trajectory;
  single: workgroup size, LDS-privatised histogram, persistent workgroups
          per CU.
Each candidate is timed with HIP events on a sample of the real work.  The
batch search is a coordinate descent: from the first value of every axis it
tries the other values of one axis at a time, moves to any faster point, and
sweeps again until a sweep finds nothing faster (``exhaustive=True`` times
the whole grid instead).  The winner is stored under (arch, size class) in
memory and, if ``cache_path`` is given, in a JSON file.  A size class is
(log2 of triangles, log2 of rays); batches with multi-source traces are a
class of their own.
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

# first value of each axis = the starting point of the batch search.  The
# packet size is not searched: the automatic (guided) sizes follow the work
# left in each XCD queue of the full batch, which a sample of it cannot show
# (measured on 256 scenes x 4-mic arrays: guided sizes 887 ms per batch, the
# fixed 1024 / 4096-ray packets a sample search picked from 942 / 929 ms).
BATCH_SPACE = {
    "packet_rays": (0,),
    "block_size": (256, 128),
    "workgroups_per_cu": (0, 2),
    "sources_per_lane": (0, 4, 16),
}
# axes that only matter for batches holding multi-source traces
MULTI_SOURCE_AXES = ("sources_per_lane",)
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


def coordinate_descent(space: dict[str, Sequence[int]], cost: Callable[[dict[str, int]], float],
                       max_sweeps: int = 3) -> tuple[dict[str, int], float]:
    """Minimise ``cost`` over the grid ``space`` one axis at a time, starting
    from the first value of every axis; each point is costed once.  The
    result is a point that no single-axis move improves: a local minimum
    when the axes interact, which on the launch-shape grid is checked
    against the exhaustive sweep by the GPU tests."""
    seen: dict[tuple, float] = {}

    def at(point: dict[str, int]) -> float:
        key = tuple(point[k] for k in space)
        if key not in seen:
            seen[key] = cost(point)
        return seen[key]

    point = {k: space[k][0] for k in space}
    best = at(point)
    for _ in range(max_sweeps):
        moved = False
        for k in space:
            for v in space[k]:
                if v == point[k]:
                    continue
                cand = dict(point, **{k: v})
                t = at(cand)
                if t < best:
                    point, best, moved = cand, t, True
        if not moved:
            break
    return point, best


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
                   force: bool = False, exhaustive: bool = False,
                   progress: Callable[[str], None] | None = None) -> BatchTuning:
        """Best BatchTuning for batches like ``requests`` (timed on a sample).

        ``progress`` receives one line per timed candidate; ``self.log`` keeps
        (class, candidate, milliseconds) for each.
        """
        import torch

        multi = any(r.num_sources > 1 for r in requests)
        space = dict(space or BATCH_SPACE)
        if not multi:
            for k in MULTI_SOURCE_AXES:
                space.pop(k, None)
        tri = int(np.median([r.scene.info["num_triangles"] for r in requests]))
        rays = int(np.median([r.num_rays for r in requests]))
        cls = size_class(tri, rays) + ("_multi" if multi else "")
        cached = None if force else self.lookup("batch", cls)
        if cached is not None:
            return BatchTuning(**cached)
        rng = np.random.default_rng(0)
        idx = rng.choice(len(requests), size=min(sample, len(requests)), replace=False)
        subset = [requests[int(i)] for i in sorted(idx)]
        own = engine is None
        engine = engine or BatchEngine(self.device, max_traces=max(len(subset), 1))
        stream = torch.cuda.current_stream(self.device)

        def cost(cand: dict[str, int]) -> float:
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
            self.log.append(("batch:" + cls, dict(cand), t))
            if progress is not None:
                progress(f"autotune batch {cls} {cand}: {t:.2f} ms")
            return t

        try:
            if exhaustive:
                best = min(_grid(space), key=cost)
            else:
                best, _ = coordinate_descent(space, cost)
        finally:
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
