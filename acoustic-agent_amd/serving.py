"""Request loop: continuous batching of RIR solves on one GPU.

The reference serves solves one at a time: its HTTP server takes a global
lock around each simulation (web_server.py:34, 271) and its batch helpers fan
whole simulations out over threads (batch.py:139-184).  Here requests are
admitted as they arrive, grouped into persistent-kernel batches, and several
batches run at once on separate streams (``slots``), so requests that arrive
while others are on the GPU start as soon as a slot and HBM budget free up;
their work overlaps on the device.  Latency is arrival -> results on the host
(echogram, counters and band RT60), reported as p50/p99 per size class and
per request kind beside throughput.

HBM budget (``hbm_budget_bytes``, 64 GB by default).  Two things are charged:
  * resident scenes: requests name host scenes (``HostScene``); the loop's
    ``SceneCache`` uploads a scene (BVH build + copy) the first time its
    content hash is seen, keeps it resident for every later request with the
    same content, and evicts the least recently used scenes that no in-flight
    batch uses when the budget needs room.  A scene not yet resident is
    charged at its upper-bound size until it is uploaded;
  * output arenas: each batch's outputs, from submission until read back.
A request is admitted when both fit next to what is already charged.
"""
from __future__ import annotations

import dataclasses
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field
from typing import Any, Callable, Iterable, Sequence

import numpy as np

from .batch import BatchEngine, BatchTuning, TraceRequest, _plan, trace_sequential
from .hip_backend import HostScene
from .models import SimConfig
from .steam_rt import estimate_reverb_times
from .workloads import ServedRequest


def request_bytes(req: ServedRequest) -> int:
    """Device bytes of a request's output arena."""
    lay = _plan(req.traces)
    return 4 * lay.f32 + 4 * lay.i32 + 8 * lay.i64 + 16 * lay.src


class SceneCache:
    """Device-resident scenes keyed by content hash, least recently used first.

    ``acquire`` uploads the host scenes that are not resident and pins all of
    them (a pinned scene is in use by an in-flight batch and is never
    evicted); ``release`` unpins.  ``evict`` frees unpinned scenes, oldest
    first.  ``bytes`` is the device memory the resident scenes hold.  Scenes
    passed in already uploaded (``DeviceScene``) belong to the caller: they
    are used as they are and not charged."""

    def __init__(self, device: int = 0, upload: Callable[[Any, int], Any] | None = None) -> None:
        self.device = int(device)
        self._upload = upload or (lambda host, dev: host.upload(dev))
        self._scenes: "OrderedDict[int, Any]" = OrderedDict()
        self._pins: dict[int, int] = {}
        self.bytes = 0
        self.stats = {"hits": 0, "misses": 0, "evictions": 0}

    def __len__(self) -> int:
        return len(self._scenes)

    def resident(self, host: HostScene) -> bool:
        return host.content_hash in self._scenes

    def pending_bytes(self, hosts: Iterable[HostScene]) -> int:
        """Upper bound of the bytes uploading ``hosts`` adds."""
        seen, total = set(), 0
        for h in hosts:
            if h.content_hash not in self._scenes and h.content_hash not in seen:
                seen.add(h.content_hash)
                total += h.bound_device_bytes
        return total

    def evictable_bytes(self, keep: set[int] = frozenset()) -> int:
        return sum(s.device_bytes for k, s in self._scenes.items() if not self._pins.get(k) and k not in keep)

    def evict(self, need: int, keep: set[int] = frozenset()) -> int:
        """Free unpinned scenes (not in ``keep``), oldest first, until at least
        ``need`` bytes are freed or none is left; returns the bytes freed."""
        freed = 0
        for k in list(self._scenes):
            if freed >= need:
                break
            if self._pins.get(k) or k in keep:
                continue
            scene = self._scenes.pop(k)
            freed += scene.device_bytes
            self.bytes -= scene.device_bytes
            self.stats["evictions"] += 1
            scene.close()
        return freed

    def acquire(self, hosts: Iterable[HostScene]) -> dict[int, Any]:
        out = {}
        for h in hosts:
            k = h.content_hash
            if k in out:
                continue
            scene = self._scenes.get(k)
            if scene is None:
                scene = self._upload(h, self.device)
                self._scenes[k] = scene
                self.bytes += scene.device_bytes
                self.stats["misses"] += 1
            else:
                self._scenes.move_to_end(k)
                self.stats["hits"] += 1
            self._pins[k] = self._pins.get(k, 0) + 1
            out[k] = scene
        return out

    def release(self, keys: Iterable[int]) -> None:
        for k in keys:
            self._pins[k] -= 1
            if self._pins[k] <= 0:
                del self._pins[k]

    def close(self) -> None:
        for scene in self._scenes.values():
            scene.close()
        self._scenes.clear()
        self._pins.clear()
        self.bytes = 0


def _host_scenes(req: ServedRequest) -> list[HostScene]:
    return [t.scene for t in req.traces if isinstance(t.scene, HostScene)]


def _resolve(traces: Sequence[TraceRequest], resident: dict[int, Any]) -> list[TraceRequest]:
    """The traces with host scenes replaced by their resident device copies."""
    return [dataclasses.replace(t, scene=resident[t.scene.content_hash]) if isinstance(t.scene, HostScene) else t
            for t in traces]


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
    kinds: dict[str, list[float]] = {}
    for c in completions:
        kinds.setdefault(c.request.kind, []).append(c.latency_s)
    return {
        "requests": len(completions),
        "traces": traces,
        "rirs": rirs,
        "makespan_s": round(makespan_s, 4),
        "rirs_per_s": round(rirs / makespan_s, 3) if makespan_s > 0 else None,
        "primary_rays_per_s": round(rays / makespan_s, 1) if makespan_s > 0 else None,
        "latency_all": _percentiles([c.latency_s for c in completions]),
        "latency_by_class": {k: _percentiles(v) for k, v in sorted(classes.items())},
        "latency_by_kind": {k: _percentiles(v) for k, v in sorted(kinds.items())},
    }


class ServingLoop:
    def __init__(self, device: int = 0, slots: int = 3, max_batch_traces: int = 256,
                 hbm_budget_bytes: int = 64 << 30, tuning: BatchTuning | None = None,
                 compute_rt60: bool = True, scene_cache: SceneCache | None = None) -> None:
        import torch

        self.torch = torch
        self.device = int(device)
        self.max_batch_traces = int(max_batch_traces)
        self.hbm_budget = int(hbm_budget_bytes)
        self.tuning = tuning or BatchTuning()
        self.compute_rt60 = compute_rt60
        self.scenes = scene_cache or SceneCache(device)
        self.engines = [BatchEngine(device, max_batch_traces) for _ in range(slots)]
        self.streams = [torch.cuda.Stream(device) for _ in range(slots)]
        self.peak_bytes = 0  # largest charged total (resident scenes + in-flight outputs) seen

    def close(self) -> None:
        for e in self.engines:
            e.close()
        self.scenes.close()

    def warm(self, stream: Sequence[ServedRequest]) -> int:
        """Make the scenes of ``stream`` resident ahead of time (as a running
        server has its working set resident), within the budget; returns
        the number uploaded."""
        before = self.scenes.stats["misses"]
        for r in stream:
            hosts = _host_scenes(r)
            if self.scenes.bytes + self.scenes.pending_bytes(hosts) <= self.hbm_budget:
                self.scenes.release(list(self.scenes.acquire(hosts)))
        return self.scenes.stats["misses"] - before

    def _fits(self, outputs: int, new_scene_bytes: int, keep: set[int]) -> bool:
        """Whether ``outputs`` more output bytes and ``new_scene_bytes`` of
        uploads fit the budget, evicting unused scenes (not in ``keep``) if
        that makes them fit."""
        over = self.scenes.bytes + outputs + new_scene_bytes - self.hbm_budget
        if over <= 0:
            return True
        if self.scenes.evictable_bytes(keep) < over:
            return False
        self.scenes.evict(over, keep)
        return True

    def run(self, stream: Sequence[ServedRequest], realtime: bool = True) -> tuple[list[Completion], dict[str, Any]]:
        """Serve ``stream`` (sorted by arrival); returns completions and a summary."""
        reqs = sorted(stream, key=lambda r: r.arrival_s)
        for r in reqs:
            if len(r.traces) > self.max_batch_traces:
                raise ValueError("a request has more traces than a batch can hold")
            if request_bytes(r) + SceneCache().pending_bytes(_host_scenes(r)) > self.hbm_budget:
                raise ValueError("a request alone exceeds the HBM budget")
        pending: deque[ServedRequest] = deque()
        free = list(range(len(self.engines)))
        inflight: list[tuple[int, Any, list[ServedRequest], int, list[int]]] = []
        done: list[Completion] = []
        out_in_use = 0
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
                for slot, _, _, nbytes, keys in finished:
                    free.append(slot)
                    out_in_use -= nbytes
                    self.scenes.release(keys)
            # refill free slots before any host-side work
            launched = False
            while free and pending:
                members, traces, nbytes, hosts = [], [], 0, []
                keep: set[int] = set()
                while pending:
                    r = pending[0]
                    b = request_bytes(r)
                    rh = _host_scenes(r)
                    new = self.scenes.pending_bytes(hosts + rh)
                    if len(traces) + len(r.traces) > self.max_batch_traces or \
                            not self._fits(out_in_use + nbytes + b, new, keep | {h.content_hash for h in rh}):
                        break
                    pending.popleft()
                    for j, t in enumerate(r.traces):
                        t.tag = (len(members), j)
                        traces.append(t)
                    members.append(r)
                    nbytes += b
                    hosts += rh
                    keep |= {h.content_hash for h in rh}
                if not members:
                    break
                resident = self.scenes.acquire(hosts)
                slot = free.pop()
                res = self.engines[slot].submit(_resolve(traces, resident), self.tuning, stream=self.streams[slot],
                                                copy_back=True)
                inflight.append((slot, res, members, nbytes, list(resident)))
                out_in_use += nbytes
                self.peak_bytes = max(self.peak_bytes, out_in_use + self.scenes.bytes)
                launched = True
            # host-side work of the finished batches: split per request, RT60
            for slot, res, members, nbytes, _ in finished:
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
        summary = summarize(done, makespan)
        summary["scene_cache"] = dict(self.scenes.stats, resident=len(self.scenes), resident_bytes=self.scenes.bytes)
        summary["peak_charged_bytes"] = self.peak_bytes
        return done, summary


def serve_sequential(stream: Sequence[ServedRequest], realtime: bool = True, plain: bool = False,
                     scene_cache: SceneCache | None = None) -> tuple[list[Completion], dict[str, Any]]:
    """Baseline: one request at a time, one launch per trace.

    ``plain`` runs each trace on the plain one-thread-per-ray kernel (the
    reference GPU backend's execution model); otherwise on the persistent
    single-trace kernel.  Scenes stay resident across requests here too
    (``scene_cache``, unbounded), so only the execution model differs.
    """
    cache = scene_cache or SceneCache()
    reqs = sorted(stream, key=lambda r: r.arrival_s)
    done = []
    t0 = time.perf_counter()
    for r in reqs:
        if realtime:
            wait = r.arrival_s - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        resident = cache.acquire(_host_scenes(r))
        results, _ = trace_sequential(_resolve(r.traces, resident), plain=plain)
        cache.release(list(resident))
        rt60 = _rt60(results[0]["echogram"], r.traces[0].bin_dur) if results else None
        t_done = time.perf_counter() - t0  # histogram and RT60 ready, as for the served path
        done.append(Completion(r, t_done - (r.arrival_s if realtime else 0.0), results, rt60))
    summary = summarize(done, time.perf_counter() - t0)
    if scene_cache is None:
        cache.close()
    return done, summary
