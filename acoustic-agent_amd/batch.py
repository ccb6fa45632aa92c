"""Batched tracing: many traces in one persistent launch.

The reference traces one (scene, source, listener channel, keyframe) per call
and fans calls out over a ``ThreadPoolExecutor`` (acoustic_agent/batch.py:
139-184, api.py:477-639, engine.py:64-73).  ``BatchEngine.submit`` instead
packs any mix of traces into one ``aa_batch_submit``: outputs live in three
flat device arenas (no per-trace allocation), sources go up in one copy, and
the persistent kernel pulls ray packets from a device work queue.  Traces are
ordered largest first so the queue drains evenly.

``trace_sequential`` is the one-scene-per-launch path (one ``aa_trace`` or
``aa_trace_plain`` per trace, synchronised and copied back per call, as the
reference's backend does per ``trace_energy_field`` call); it is the baseline
the batched path is measured against.
"""
from __future__ import annotations

import ctypes
import dataclasses
import time
from dataclasses import dataclass, field
from typing import Any, Sequence

import numpy as np

from . import _lib
from .hip_backend import DeviceScene, TraceBuffers, launch_trace, make_params
from .sampling import diffuse_sample_bank
from .steam_rt import histogram_bins

C = 343.0


@dataclass
class TraceRequest:
    """One trace: a resident scene, a listener, one or more sources."""

    scene: DeviceScene
    listener: np.ndarray
    sources: np.ndarray  # [Ks,3]
    num_rays: int
    num_bounces: int
    duration_s: float = 2.0
    bin_dur: float = 1e-3
    seed: int = 1729
    c: float = C
    listener_radius: float = 0.1
    source_radius: float = 0.1
    irradiance_min_distance: float = 1.0
    specular_exponent: float = 100.0
    source_forward: tuple[float, float, float] = (1.0, 0.0, 0.0)
    dipole_weight: float = 0.0
    dipole_power: float = 1.0
    render_ambisonics: bool = False
    num_diffuse_samples: int = 128  # rows of the diffuse sample bank (rt_num_diffuse_samples)
    use_bvh: bool = True
    log_rays: int = 0     # hit log (triangle, t, point) of rays [0, log_rays), first log_bounces bounces
    log_bounces: int = 0
    tag: Any = None

    @property
    def num_bins(self) -> int:
        return histogram_bins(self.duration_s, self.bin_dur)

    @property
    def num_sources(self) -> int:
        return int(np.asarray(self.sources).reshape(-1, 3).shape[0])

    def params(self) -> _lib.TraceParams:
        return make_params(
            num_rays=self.num_rays, num_bounces=self.num_bounces, num_bins=self.num_bins,
            num_diffuse_samples=self.num_diffuse_samples, num_sources=self.num_sources,
            rng_mode=_lib.AA_RNG_COUNTER, use_bvh=self.use_bvh,
            render_ambisonics=self.render_ambisonics, seed=self.seed, listener=self.listener,
            source_forward=self.source_forward, bin_dur=self.bin_dur, speed_of_sound=self.c,
            max_path_len=self.duration_s * self.c, listener_radius=self.listener_radius,
            source_radius=self.source_radius, irradiance_min_distance=self.irradiance_min_distance,
            specular_exponent=self.specular_exponent, dipole_weight=self.dipole_weight,
            dipole_power=self.dipole_power, log_rays=min(self.log_rays, self.num_rays),
            log_bounces=self.log_bounces if self.log_rays > 0 else 0)

    def sources4(self) -> np.ndarray:
        src = np.asarray(self.sources, dtype=np.float64).reshape(-1, 3)
        delays = np.linalg.norm(src - np.asarray(self.listener, dtype=np.float64)[None, :], axis=1) / self.c
        return np.concatenate([src, delays[:, None]], axis=1)

    @property
    def work(self) -> float:
        """Relative cost used for ordering (rays x bounces x log triangles)."""
        return float(self.num_rays) * self.num_bounces * np.log2(2.0 + self.scene.info["num_triangles"])


@dataclass
class BatchTuning:
    block_size: int = 256
    packet_rays: int = 0  # 0: guided sizes per trace (large early in each XCD queue, 1024 at its end)
    workgroups_per_cu: int = 0
    use_graph: bool = False
    sources_per_lane: int = 0  # multi-source traces: sources per lane task (0: each trace's own setting)

    def as_struct(self) -> _lib.BatchTuning:
        return _lib.BatchTuning(int(self.block_size), int(self.packet_rays), int(self.workgroups_per_cu),
                                1 if self.use_graph else 0)

    def key(self) -> tuple:
        return (self.block_size, self.packet_rays, self.workgroups_per_cu, self.use_graph, self.sources_per_lane)


@dataclass
class _Layout:
    f32: int = 0
    i32: int = 0
    i64: int = 0
    src: int = 0
    slots: list = field(default_factory=list)


def _plan(requests: Sequence[TraceRequest]) -> _Layout:
    lay = _Layout()
    for r in requests:
        Ks, K, S = r.num_sources, r.num_bins, r.scene.num_surfaces
        slot = {}
        slot["echogram"] = (lay.f32, Ks * 6 * K)
        lay.f32 += Ks * 6 * K
        if r.render_ambisonics:
            slot["ambisonic"] = (lay.f32, Ks * 24 * K)
            lay.f32 += Ks * 24 * K
        slot["surface_energy"] = (lay.f32, Ks * S)
        lay.f32 += Ks * S
        for name, n in (("hit_counts", Ks * S), ("contrib_counts", Ks * S), ("actual_bounces", Ks),
                        ("active_count", Ks)):
            slot[name] = (lay.i32, n)
            lay.i32 += n
        slot["counters"] = (lay.i64, _lib.AA_CNT_COUNT)
        lay.i64 += _lib.AA_CNT_COUNT
        slot["sources"] = (lay.src, Ks)
        lay.src += Ks
        lay.slots.append(slot)
    return lay


class BatchResult:
    """Device outputs of one submission; ``wait`` copies them back."""

    def __init__(self, requests, layout, arenas, event, device, host=None, logs=None) -> None:
        self.requests = requests
        self.layout = layout
        self.arenas = arenas
        self.logs = logs or {}  # request index -> (triangle, t, point) device tensors
        self.event = event
        self.device = device
        self.host_arenas = host  # pinned copies queued on the submission stream (or None)
        self._host = None

    def done(self) -> bool:
        return bool(self.event.query())

    def wait(self) -> list[dict[str, Any]]:
        if self._host is None:
            self.event.synchronize()
            if self.host_arenas is not None:
                f32, i32, i64 = (a.numpy() for a in self.host_arenas)
            else:
                f32, i32, i64 = (a.cpu().numpy() for a in self.arenas[:3])
            out = []
            for r, slot in zip(self.requests, self.layout.slots):
                Ks, K, S = r.num_sources, r.num_bins, r.scene.num_surfaces

                def take(arr, name, shape):
                    off, n = slot[name]
                    return arr[off:off + n].reshape(shape).copy()

                res = {
                    "echogram": take(f32, "echogram", (Ks, 6, K)),
                    "surface_energy": take(f32, "surface_energy", (Ks, S)),
                    "hit_counts": take(i32, "hit_counts", (Ks, S)),
                    "contrib_counts": take(i32, "contrib_counts", (Ks, S)),
                    "actual_bounces": take(i32, "actual_bounces", (Ks,)),
                    "active_count": take(i32, "active_count", (Ks,)),
                    "counters": dict(zip(_lib.COUNTER_NAMES, (int(x) for x in take(i64, "counters",
                                                                                   (_lib.AA_CNT_COUNT,))))),
                    "tag": r.tag,
                }
                if "ambisonic" in slot:
                    res["ambisonic"] = take(f32, "ambisonic", (Ks, 6, 4, K))
                if len(out) in self.logs:
                    res["log_triangle"], res["log_t"], res["log_point"] = (x.cpu().numpy()
                                                                           for x in self.logs[len(out)])
                out.append(res)
            self._host = out
        return self._host


class BatchEngine:
    """Owner of one ``aa_batch_t`` (pinned descriptor table) on one device."""

    def __init__(self, device: int = 0, max_traces: int = 4096) -> None:
        import torch

        self.torch = torch
        self.device = int(device)
        self.max_traces = int(max_traces)
        lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(lib.aa_batch_create(self.device, self.max_traces, ctypes.byref(handle)))
        self._handle = handle
        self._banks: dict[int, Any] = {}

    def bank(self, rows: int):
        """The diffuse sample bank of ``rows`` rows on this device (uploaded once)."""
        if rows not in self._banks:
            self._banks[rows] = self.torch.from_numpy(
                np.ascontiguousarray(diffuse_sample_bank(rows), dtype=np.float32)).to(f"cuda:{self.device}")
        return self._banks[rows]

    def close(self) -> None:
        if self._handle is not None:
            _lib.load().aa_batch_release(self._handle)
            self._handle = None

    def __del__(self) -> None:
        try:
            from . import hip_backend

            if not hip_backend._SHUTDOWN:
                self.close()
        except Exception:
            pass

    def pack(self, requests: Sequence[TraceRequest], tuning: BatchTuning | None = None, order: bool = True):
        """The C descriptor table of ``requests`` with freshly allocated output
        arenas (the caller's outputs of the submission): returns (table,
        requests in table order, layout, arenas, hit logs)."""
        torch = self.torch
        if not requests:
            raise ValueError("empty batch")
        tuning = tuning or BatchTuning()
        reqs = sorted(requests, key=lambda r: -r.work) if order else list(requests)
        lay = _plan(reqs)
        dev = f"cuda:{self.device}"
        f32 = torch.empty(max(lay.f32, 1), dtype=torch.float32, device=dev)
        i32 = torch.empty(max(lay.i32, 1), dtype=torch.int32, device=dev)
        i64 = torch.empty(max(lay.i64, 1), dtype=torch.int64, device=dev)
        src_host = np.concatenate([r.sources4() for r in reqs]).astype(np.float32)
        src = torch.from_numpy(src_host).to(dev, non_blocking=False)
        table = (_lib.BatchTrace * len(reqs))()
        fp, ip, lp, sp = f32.data_ptr(), i32.data_ptr(), i64.data_ptr(), src.data_ptr()
        logs = {}
        for i, r in enumerate(reqs):
            if r.log_rays > 0 and r.log_bounces > 0:
                # bounces a ray never reaches keep -2 / 0
                shape = (min(r.log_rays, r.num_rays), r.log_bounces)
                logs[i] = (torch.full(shape, -2, dtype=torch.int32, device=dev),
                           torch.zeros(shape, dtype=torch.float32, device=dev),
                           torch.zeros(shape + (3,), dtype=torch.float32, device=dev))
        for i, (r, slot) in enumerate(zip(reqs, lay.slots)):
            t = table[i]
            t.scene = r.scene.handle
            t.params = r.params()
            if tuning.sources_per_lane > 0 and r.num_sources > 1:
                t.params.sources_per_lane = int(tuning.sources_per_lane)
            t.d_sources = sp + slot["sources"][0] * 16
            t.d_diffuse_bank = self.bank(r.num_diffuse_samples).data_ptr()
            o = t.outputs
            o.d_echogram = fp + slot["echogram"][0] * 4
            o.d_ambisonic = (fp + slot["ambisonic"][0] * 4) if "ambisonic" in slot else 0
            o.d_surface_energy = fp + slot["surface_energy"][0] * 4
            o.d_hit_counts = ip + slot["hit_counts"][0] * 4
            o.d_contrib_counts = ip + slot["contrib_counts"][0] * 4
            o.d_actual_bounces = ip + slot["actual_bounces"][0] * 4
            o.d_active_count = ip + slot["active_count"][0] * 4
            o.d_counters = lp + slot["counters"][0] * 8
            if i in logs:
                o.d_log_triangle, o.d_log_t, o.d_log_point = (x.data_ptr() for x in logs[i])
        return table, reqs, lay, (f32, i32, i64, src), logs

    def _finish(self, reqs, lay, arenas, logs, s, copy_back: bool) -> BatchResult:
        torch = self.torch
        host = None
        if copy_back:
            host = tuple(torch.empty(a.shape, dtype=a.dtype, pin_memory=True) for a in arenas[:3])
            with torch.cuda.stream(s):
                for h, a in zip(host, arenas[:3]):
                    h.copy_(a, non_blocking=True)
        event = torch.cuda.Event()
        event.record(s)
        return BatchResult(reqs, lay, arenas, event, self.device, host, logs)

    def submit(self, requests: Sequence[TraceRequest], tuning: BatchTuning | None = None, stream=None,
               order: bool = True, copy_back: bool = False) -> BatchResult:
        """Enqueue one persistent batch on ``stream`` (``aa_batch_submit``).

        ``copy_back`` also queues the copies of every output into pinned host
        memory on the same stream, so that ``done()`` means the results are
        on the host and reading them waits on nothing else.
        """
        if len(requests) > self.max_traces:
            raise ValueError(f"at most {self.max_traces} traces per submission")
        tuning = tuning or BatchTuning()
        table, reqs, lay, arenas, logs = self.pack(requests, tuning, order)
        s = stream if stream is not None else self.torch.cuda.current_stream(self.device)
        ts = tuning.as_struct()
        _lib.check(_lib.load().aa_batch_submit(self._handle, table, len(reqs), ctypes.byref(ts),
                                               ctypes.c_void_p(s.cuda_stream)))
        return self._finish(reqs, lay, arenas, logs, s, copy_back)

    def submit_static(self, requests: Sequence[TraceRequest], tuning: BatchTuning | None = None, stream=None,
                      order: bool = True) -> BatchResult:
        """The same batch through ``aa_batch_submit_static``: the library
        allocates nothing for it (its table is a static device array of
        ``AA_STATIC_BATCH_TRACES`` entries); the outputs are the arenas
        allocated here, as the caller's."""
        tuning = tuning or BatchTuning()
        table, reqs, lay, arenas, logs = self.pack(requests, tuning, order)
        s = stream if stream is not None else self.torch.cuda.current_stream(self.device)
        ts = tuning.as_struct()
        _lib.check(_lib.load().aa_batch_submit_static(self.device, table, len(reqs), ctypes.byref(ts),
                                                      ctypes.c_void_p(s.cuda_stream)))
        return self._finish(reqs, lay, arenas, logs, s, False)


def trace_sequential(requests: Sequence[TraceRequest], stream=None,
                     plain: bool = False) -> tuple[list[dict[str, Any]], float]:
    """One launch per trace, synchronised and copied back per call.

    ``plain`` uses the one-thread-per-ray kernel of the reference GPU backend
    (``aa_trace_plain``), else the persistent single-trace kernel.  Host
    scenes are uploaded (once per content) on device 0 for the call and
    released after it.
    """
    import torch

    out = []
    uploaded: dict[int, DeviceScene] = {}  # host scenes (HostScene) uploaded for this call
    t0 = time.perf_counter()
    for r in requests:
        if not isinstance(r.scene, DeviceScene):
            key = r.scene.content_hash
            if key not in uploaded:
                uploaded[key] = r.scene.upload(0)
            r = dataclasses.replace(r, scene=uploaded[key])
        dev = f"cuda:{r.scene.device}"
        bank = torch.from_numpy(np.ascontiguousarray(diffuse_sample_bank(r.num_diffuse_samples),
                                                     dtype=np.float32)).to(dev)
        src = torch.from_numpy(r.sources4().astype(np.float32)).to(dev)
        logged = r.log_rays > 0 and r.log_bounces > 0
        buf = TraceBuffers.allocate(r.scene.device, r.num_sources, r.num_bins, r.scene.num_surfaces,
                                    r.render_ambisonics, log_rays=min(r.log_rays, r.num_rays) if logged else 0,
                                    log_bounces=r.log_bounces if logged else 0)
        s = stream if stream is not None else torch.cuda.current_stream(r.scene.device)
        launch_trace(r.scene, r.params(), src, bank, buf, stream=s, plain=plain)
        s.synchronize()
        out.append({
            "echogram": buf.echogram.cpu().numpy(),
            "hit_counts": buf.hit_counts.cpu().numpy(),
            "contrib_counts": buf.contrib_counts.cpu().numpy(),
            "surface_energy": buf.surface_energy.cpu().numpy(),
            "actual_bounces": buf.actual_bounces.cpu().numpy(),
            "active_count": buf.active_count.cpu().numpy(),
            "counters": dict(zip(_lib.COUNTER_NAMES, (int(x) for x in buf.counters.cpu().numpy()))),
            "tag": r.tag,
        })
        if r.render_ambisonics:
            out[-1]["ambisonic"] = buf.ambisonic.cpu().numpy()
        if logged:
            out[-1]["log_triangle"] = buf.log_triangle.cpu().numpy()
            out[-1]["log_t"] = buf.log_t.cpu().numpy()
            out[-1]["log_point"] = buf.log_point.cpu().numpy()
    elapsed = time.perf_counter() - t0
    for scene in uploaded.values():
        scene.close()
    return out, elapsed
