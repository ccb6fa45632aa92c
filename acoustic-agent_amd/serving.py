"""Request loop: continuous batching of RIR solves on one GPU.

The reference serves solves one at a time: its HTTP server takes a global
lock around each simulation (web_server.py:34, 271) and its batch helpers fan
whole simulations out over threads (batch.py:139-184).  Here requests are
admitted as they arrive, grouped into persistent-kernel batches, and several
batches run at once on separate streams (``slots``), so requests that arrive
while others are on the GPU start as soon as a slot and HBM budget free up;
their work overlaps on the device.  Scenes (BVHs) stay resident and are
shared by every request that uses them; each batch's outputs live in one
arena whose size is charged against ``hbm_budget_bytes`` until the batch is
read back.  Latency is arrival -> results on the host (echogram, counters and
band RT60), reported as p50/p99 per size class beside throughput.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Any, Sequence

import numpy as np

from .batch import BatchEngine, BatchTuning, _plan, trace_sequential
from .models import SimConfig
from .steam_rt import estimate_reverb_times
from .workloads import ServedRequest


def request_bytes(req: ServedRequest) -> int:
    lay = _plan(req.traces)
    return 4 * lay.f32 + 4 * lay.i32 + 8 * lay.i64 + 16 * lay.src


def _rt60(echogram: np.ndarray, bin_dur: float) -> dict[str, float]:
    e = np.asarray(echogram, dtype=np.float64).reshape(-1, 6, echogram.shape[-1])[0]
    return estimate_reverb_times({"echogram": e, "bin_duration_s": bin_dur, "num_bins": e.shape[-1]}, SimConfig())


@dataclass
class Completion:
    request: ServedRequest
    latency_s: float
    results: list[dict[str, Any]] = field(default_factory=list)
    rt60: dict[str, float] | None = None


def _percentiles(values: Sequence[float]) -> dict[str, float]:
    if not values:
        return {"p50_ms": None, "p99_ms": None, "count": 0}
    a = np.asarray(values) * 1000.0
    return {"p50_ms": round(float(np.percentile(a, 50)), 3), "p99_ms": round(float(np.percentile(a, 99)), 3),
            "count": int(a.size)}


def summarize(completions: list[Completion], makespan_s: float) -> dict[str, Any]:
    classes: dict[str, list[float]] = {}
    for c in completions:
        classes.setdefault(c.request.size_class, []).append(c.latency_s)
    traces = sum(len(c.request.traces) for c in completions)
    rirs = sum(t.num_sources for c in completions for t in c.request.traces)
    rays = sum(t.num_rays for c in completions for t in c.request.traces)
    return {
        "requests": len(completions),
        "traces": traces,
        "rirs": rirs,
        "makespan_s": round(makespan_s, 4),
        "rirs_per_s": round(rirs / makespan_s, 3) if makespan_s > 0 else None,
        "primary_rays_per_s": round(rays / makespan_s, 1) if makespan_s > 0 else None,
        "latency_all": _percentiles([c.latency_s for c in completions]),
        "latency_by_class": {k: _percentiles(v) for k, v in sorted(classes.items())},
    }


class ServingLoop:
    def __init__(self, device: int = 0, slots: int = 3, max_batch_traces: int = 256,
                 hbm_budget_bytes: int = 64 << 30, tuning: BatchTuning | None = None,
                 compute_rt60: bool = True) -> None:
        import torch

        self.torch = torch
        self.device = int(device)
        self.max_batch_traces = int(max_batch_traces)
        self.hbm_budget = int(hbm_budget_bytes)
        self.tuning = tuning or BatchTuning()
        self.compute_rt60 = compute_rt60
        self.engines = [BatchEngine(device, max_batch_traces) for _ in range(slots)]
        self.streams = [torch.cuda.Stream(device) for _ in range(slots)]

    def close(self) -> None:
        for e in self.engines:
            e.close()

    def run(self, stream: Sequence[ServedRequest], realtime: bool = True) -> tuple[list[Completion], dict[str, Any]]:
        """Serve ``stream`` (sorted by arrival); returns completions and a summary."""
        reqs = sorted(stream, key=lambda r: r.arrival_s)
        for r in reqs:
            if len(r.traces) > self.max_batch_traces:
                raise ValueError("a request has more traces than a batch can hold")
            if request_bytes(r) > self.hbm_budget:
                raise ValueError("a request alone exceeds the HBM budget")
        pending: deque[ServedRequest] = deque()
        free = list(range(len(self.engines)))
        inflight: list[tuple[int, Any, list[ServedRequest], int]] = []
        done: list[Completion] = []
        in_use = 0
        i = 0
        t0 = time.perf_counter()
        offset = 0.0 if realtime else None
        while i < len(reqs) or pending or inflight:
            now = time.perf_counter() - t0
            while i < len(reqs) and (offset is None or reqs[i].arrival_s <= now):
                pending.append(reqs[i])
                i += 1
            # batches whose results (copied back on their own stream) are on the host
            finished = [x for x in inflight if x[1].done()]
            if finished:
                inflight = [x for x in inflight if all(x is not f for f in finished)]
                for slot, _, _, nbytes in finished:
                    free.append(slot)
                    in_use -= nbytes
            # refill free slots before any host-side work
            launched = False
            while free and pending:
                members, traces, nbytes = [], [], 0
                while pending:
                    r = pending[0]
                    b = request_bytes(r)
                    if len(traces) + len(r.traces) > self.max_batch_traces or in_use + nbytes + b > self.hbm_budget:
                        break
                    pending.popleft()
                    for j, t in enumerate(r.traces):
                        t.tag = (len(members), j)
                        traces.append(t)
                    members.append(r)
                    nbytes += b
                if not members:
                    break
                slot = free.pop()
                res = self.engines[slot].submit(traces, self.tuning, stream=self.streams[slot], copy_back=True)
                inflight.append((slot, res, members, nbytes))
                in_use += nbytes
                launched = True
            # host-side work of the finished batches: split per request, RT60
            for slot, res, members, nbytes in finished:
                results = res.wait()
                by_req: dict[int, list] = {}
                for r in results:
                    by_req.setdefault(r["tag"][0], []).append(r)
                for k, m in enumerate(members):
                    rs = sorted(by_req.get(k, []), key=lambda r: r["tag"][1])
                    rt60 = _rt60(rs[0]["echogram"], m.traces[0].bin_dur) if self.compute_rt60 and rs else None
                    t_done = time.perf_counter() - t0  # histogram and RT60 ready
                    done.append(Completion(m, t_done - (m.arrival_s if offset is not None else 0.0), rs, rt60))
            if not finished and not launched:
                time.sleep(0.0002)
        makespan = time.perf_counter() - t0
        return done, summarize(done, makespan)


def serve_sequential(stream: Sequence[ServedRequest], realtime: bool = True,
                     plain: bool = False) -> tuple[list[Completion], dict[str, Any]]:
    """Baseline: one request at a time, one launch per trace.

    ``plain`` runs each trace on the plain one-thread-per-ray kernel (the
    reference GPU backend's execution model); otherwise on the persistent
    single-trace kernel.
    """
    reqs = sorted(stream, key=lambda r: r.arrival_s)
    done = []
    t0 = time.perf_counter()
    for r in reqs:
        if realtime:
            wait = r.arrival_s - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        results, _ = trace_sequential(r.traces, plain=plain)
        rt60 = _rt60(results[0]["echogram"], r.traces[0].bin_dur) if results else None
        t_done = time.perf_counter() - t0  # histogram and RT60 ready, as for the served path
        done.append(Completion(r, t_done - (r.arrival_s if realtime else 0.0), results, rt60))
    return done, summarize(done, time.perf_counter() - t0)
