"""HIP backend module: the drop-in twin of the reference's CUDA backend.

The reference binds its GPU tracer as a module with two calls,
``cuda_available(device_id)`` and ``trace_energy_field_cuda(**kw)``
(acoustic_agent/steam_rt_cuda.py:33-226).  This module offers
``hip_available`` and ``trace_energy_field_hip`` with the same keyword
arguments and the same raw result keys (echogram, ambisonic, hit_counts,
contrib_counts, surface_energy, actual_bounces, active_count, kernel_time_s,
transfer_time_s, device, device_input_cache), computed by the gfx950 kernels
of ``libaa_tracer.so`` through its C ABI.

Differences that follow from the MI355X design:
  * ``arrays`` (the reference surface arrays) is tessellated to triangles that
    carry their reference surface id; the device scene is cached by content
    (like the reference's device input cache, steam_rt_cuda.py:229-261).
  * ``rng_mode="counter"`` draws the ray set on the device from Philox4x32-10
    instead of uploading B x N workspaces; ``"reference"`` uploads the
    reference's NumPy streams (the default here, for bit-level input parity).
  * ``sources`` traces several source positions (keyframes) that share one
    listener and one ray set in a single pass (the trajectories do not depend
    on the source, steam_rt.py:4750-4849).
PyTorch-ROCm provides device memory and the stream only.
"""
from __future__ import annotations

import atexit
import ctypes
import hashlib
import threading
import time
from collections import OrderedDict
from dataclasses import dataclass
from typing import Any, Mapping

import numpy as np

from . import _lib
from .scene import TriangleMesh, tessellate

_SHUTDOWN = False


@atexit.register
def _mark_shutdown() -> None:
    global _SHUTDOWN
    _SHUTDOWN = True


def _torch():
    import torch

    return torch


def hip_available(device_id: int = 0) -> bool:
    """True iff ``libaa_tracer.so`` loads and ``device_id`` is a HIP device.

    A missing library raises instead of returning False: the package has no
    CPU tracer to fall back to.
    """
    lib = _lib.load()
    count = ctypes.c_int32(0)
    if lib.aa_device_count(ctypes.byref(count)) != _lib.AA_OK:
        return False
    return 0 <= int(device_id) < int(count.value)


def device_info(device_id: int = 0) -> dict[str, Any]:
    lib = _lib.load()
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(64)
    hbm = ctypes.c_uint64(0)
    cus = ctypes.c_int32(0)
    _lib.check(lib.aa_device_info(int(device_id), name, 256, arch, 64, ctypes.byref(hbm), ctypes.byref(cus)))
    return {
        "id": int(device_id),
        "name": name.value.decode(),
        "arch": arch.value.decode(),
        "hbm_bytes": int(hbm.value),
        "compute_units": int(cus.value),
    }


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


class DeviceScene:
    """A triangle scene resident on one device (opaque ``aa_scene_t``)."""

    def __init__(self, mesh: TriangleMesh, reflection: np.ndarray, scattering: np.ndarray, *, device: int = 0,
                 leaf_size: int = 4) -> None:
        lib = _lib.load()
        v = np.ascontiguousarray(mesh.vertices, dtype=np.float32)
        s = np.ascontiguousarray(mesh.surface, dtype=np.int32)
        n = np.ascontiguousarray(mesh.normals, dtype=np.float32)
        refl = np.ascontiguousarray(reflection, dtype=np.float32)
        scat = np.ascontiguousarray(scattering, dtype=np.float32)
        if refl.shape != (mesh.num_surfaces, _lib.AA_NUM_BANDS) or scat.shape != (mesh.num_surfaces,):
            raise ValueError("reflection must be [S,6] and scattering [S] for the mesh's S surfaces")
        handle = ctypes.c_void_p()
        _lib.check(lib.aa_scene_create(int(device), _ptr(v), _ptr(s), _ptr(n), int(v.shape[0]), _ptr(refl), _ptr(scat),
                                       int(mesh.num_surfaces), int(leaf_size), ctypes.byref(handle)))
        self._handle = handle
        self.device = int(device)
        self.mesh = mesh
        info = _lib.SceneInfo()
        _lib.check(lib.aa_scene_info(handle, ctypes.byref(info)))
        self.info = {name: getattr(info, name) for name, _ in _lib.SceneInfo._fields_}

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._handle is None:
            raise RuntimeError("scene has been released")
        return self._handle

    @property
    def num_surfaces(self) -> int:
        return int(self.info["num_surfaces"])

    @property
    def device_bytes(self) -> int:
        return int(self.info["device_bytes"])

    def close(self) -> None:
        if self._handle is not None and not _SHUTDOWN:
            _lib.load().aa_scene_release(self._handle)
        self._handle = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


class HostScene:
    """A triangle scene on the host, identified by its content hash.

    The hash is ``aa_scene_hash`` over the same float32/int32 inputs
    ``aa_scene_create`` takes, so a host scene and its resident device copy
    share one key (a serving cache uploads a host scene once and reuses the
    device copy for every request with the same content).  ``workload``
    optionally keeps the description the scene was compiled from."""

    # upper bound of device bytes per triangle before upload: at most
    # max(1, T - 1) wide nodes of 128 B, two triangle copies of 48 B, 2T - 1
    # binary nodes of 32 B and a 16 B normal
    BOUND_BYTES_PER_TRIANGLE = 128 + 96 + 64 + 16

    def __init__(self, mesh: TriangleMesh, reflection: np.ndarray, scattering: np.ndarray, *, leaf_size: int = 4,
                 workload: Any = None) -> None:
        self.mesh = mesh
        self.reflection = np.ascontiguousarray(reflection, dtype=np.float32)
        self.scattering = np.ascontiguousarray(scattering, dtype=np.float32)
        if self.reflection.shape != (mesh.num_surfaces, _lib.AA_NUM_BANDS) or \
                self.scattering.shape != (mesh.num_surfaces,):
            raise ValueError("reflection must be [S,6] and scattering [S] for the mesh's S surfaces")
        self.leaf_size = int(leaf_size)
        self.workload = workload
        v = np.ascontiguousarray(mesh.vertices, dtype=np.float32)
        s = np.ascontiguousarray(mesh.surface, dtype=np.int32)
        n = np.ascontiguousarray(mesh.normals, dtype=np.float32)
        h = ctypes.c_uint64(0)
        _lib.check(_lib.load().aa_scene_hash(_ptr(v), _ptr(s), _ptr(n), int(v.shape[0]), _ptr(self.reflection),
                                             _ptr(self.scattering), int(mesh.num_surfaces), self.leaf_size,
                                             ctypes.byref(h)))
        self.content_hash = int(h.value)
        self.info = {"num_triangles": int(v.shape[0]), "num_surfaces": int(mesh.num_surfaces),
                     "content_hash": self.content_hash}

    @property
    def num_surfaces(self) -> int:
        return int(self.mesh.num_surfaces)

    @property
    def bound_device_bytes(self) -> int:
        return self.info["num_triangles"] * self.BOUND_BYTES_PER_TRIANGLE + 128 + self.num_surfaces * 28 + 4

    def upload(self, device: int = 0) -> "DeviceScene":
        ds = DeviceScene(self.mesh, self.reflection, self.scattering, device=device, leaf_size=self.leaf_size)
        ds.workload = self.workload
        return ds


def scene_from_arrays(arrays: Mapping[str, Any], *, device: int = 0, max_edge: float | None = None,
                      leaf_size: int = 4) -> DeviceScene:
    mesh = tessellate(arrays, max_edge)
    return DeviceScene(mesh, np.asarray(arrays["reflection"]), np.asarray(arrays["scattering"]), device=device,
                       leaf_size=leaf_size)


# ---------------------------------------------------------------------------
# device caches (scenes by content, ray-set inputs by identity)
# ---------------------------------------------------------------------------

_CACHE_LOCK = threading.RLock()
_SCENE_CACHE: "OrderedDict[tuple, DeviceScene]" = OrderedDict()
_SCENE_CACHE_BYTES = 0
_SCENE_CACHE_LIMIT = 512 * 1024 * 1024
_INPUT_CACHE: "OrderedDict[tuple, tuple[Any, Any, int]]" = OrderedDict()
_INPUT_CACHE_BYTES = 0
_INPUT_CACHE_LIMIT = 512 * 1024 * 1024


def _arrays_key(arrays: Mapping[str, Any], device: int, max_edge: float | None, leaf_size: int) -> tuple:
    h = hashlib.blake2b(digest_size=16)
    for name in ("kinds", "wall_a", "wall_delta", "wall_z", "z_values", "box_center", "box_axis_u", "box_axis_v",
                 "box_half", "box_z", "normals", "reflection", "scattering", "corners"):
        a = np.ascontiguousarray(arrays[name])
        h.update(name.encode())
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return (int(device), h.hexdigest(), max_edge, int(leaf_size))


def cached_scene(arrays: Mapping[str, Any], *, device: int = 0, max_edge: float | None = None, leaf_size: int = 4,
                 stats: dict[str, int] | None = None) -> DeviceScene:
    global _SCENE_CACHE_BYTES
    key = _arrays_key(arrays, device, max_edge, leaf_size)
    with _CACHE_LOCK:
        scene = _SCENE_CACHE.get(key)
        if scene is not None:
            _SCENE_CACHE.move_to_end(key)
            if stats is not None:
                stats["hits"] += 1
            return scene
    scene = scene_from_arrays(arrays, device=device, max_edge=max_edge, leaf_size=leaf_size)
    with _CACHE_LOCK:
        if stats is not None:
            stats["misses"] += 1
        _SCENE_CACHE[key] = scene
        _SCENE_CACHE_BYTES += scene.device_bytes
        while len(_SCENE_CACHE) > 1 and _SCENE_CACHE_BYTES > _SCENE_CACHE_LIMIT:
            _, old = _SCENE_CACHE.popitem(last=False)
            _SCENE_CACHE_BYTES -= old.device_bytes
            old.close()
    return scene


def _device_input(value: np.ndarray, dtype: Any, device: int, stats: dict[str, int]):
    """Upload a read-only host array once (keyed by identity, like steam_rt_cuda.py:229-261)."""
    global _INPUT_CACHE_BYTES
    torch = _torch()
    host = np.asarray(value)
    key = (int(device), int(host.__array_interface__["data"][0]), host.shape, host.strides, host.dtype.str,
           np.dtype(dtype).str)
    with _CACHE_LOCK:
        cached = _INPUT_CACHE.get(key)
        if cached is not None and cached[0] is value:
            _INPUT_CACHE.move_to_end(key)
            stats["hits"] += 1
            return cached[1]
    converted = np.ascontiguousarray(host, dtype=dtype)
    if not converted.flags.writeable:  # the reference marks its arrays read-only; torch wants writable memory
        converted = converted.copy()
    tensor = torch.from_numpy(converted).to(f"cuda:{device}")
    size = int(converted.nbytes)
    with _CACHE_LOCK:
        stats["misses"] += 1
        previous = _INPUT_CACHE.pop(key, None)
        if previous is not None:
            _INPUT_CACHE_BYTES -= previous[2]
        while _INPUT_CACHE and _INPUT_CACHE_BYTES + size > _INPUT_CACHE_LIMIT:
            _, expired = _INPUT_CACHE.popitem(last=False)
            _INPUT_CACHE_BYTES -= expired[2]
        if size <= _INPUT_CACHE_LIMIT and not host.flags.writeable:
            _INPUT_CACHE[key] = (value, tensor, size)
            _INPUT_CACHE_BYTES += size
    return tensor


def clear_caches() -> None:
    global _SCENE_CACHE_BYTES, _INPUT_CACHE_BYTES
    with _CACHE_LOCK:
        for scene in _SCENE_CACHE.values():
            scene.close()
        _SCENE_CACHE.clear()
        _INPUT_CACHE.clear()
        _SCENE_CACHE_BYTES = 0
        _INPUT_CACHE_BYTES = 0


# ---------------------------------------------------------------------------
# trace
# ---------------------------------------------------------------------------


@dataclass
class TraceBuffers:
    """Device outputs of one trace (torch tensors)."""

    echogram: Any
    ambisonic: Any
    hit_counts: Any
    contrib_counts: Any
    surface_energy: Any
    actual_bounces: Any
    active_count: Any
    counters: Any
    log_triangle: Any
    log_t: Any
    log_point: Any = None
    visual: tuple | None = None  # points, surfaces, rays, orders, distances, gains (visual candidates)
    ray_times: Any = None  # [N,2] int64 start / end ticks of the device's 100 MHz clock

    @classmethod
    def allocate(cls, device: int, num_sources: int, num_bins: int, num_surfaces: int, ambisonic: bool,
                 log_rays: int = 0, log_bounces: int = 0, visual_limit: int = 0,
                 num_bounces: int = 0, ray_times_rays: int = 0) -> "TraceBuffers":
        torch = _torch()
        dev = f"cuda:{device}"
        f32, i32 = torch.float32, torch.int32
        V, B = int(visual_limit), max(int(num_bounces), 0)
        visual = (torch.empty((V, B, 3), dtype=f32, device=dev), torch.empty((V, B), dtype=i32, device=dev),
                  torch.empty((V,), dtype=i32, device=dev), torch.empty((V,), dtype=i32, device=dev),
                  torch.empty((V,), dtype=f32, device=dev), torch.empty((V,), dtype=f32, device=dev)) if V > 0 else None
        return cls(
            echogram=torch.empty((num_sources, 6, num_bins), dtype=f32, device=dev),
            ambisonic=torch.empty((num_sources, 6, 4, num_bins), dtype=f32, device=dev) if ambisonic else None,
            hit_counts=torch.empty((num_sources, num_surfaces), dtype=i32, device=dev),
            contrib_counts=torch.empty((num_sources, num_surfaces), dtype=i32, device=dev),
            surface_energy=torch.empty((num_sources, num_surfaces), dtype=f32, device=dev),
            actual_bounces=torch.empty((num_sources,), dtype=i32, device=dev),
            active_count=torch.empty((num_sources,), dtype=i32, device=dev),
            counters=torch.empty((_lib.AA_CNT_COUNT,), dtype=torch.int64, device=dev),
            # bounces a ray never reaches keep -2 / 0
            log_triangle=torch.full((log_rays, log_bounces), -2, dtype=i32, device=dev) if log_rays > 0 else None,
            log_t=torch.zeros((log_rays, log_bounces), dtype=f32, device=dev) if log_rays > 0 else None,
            log_point=torch.zeros((log_rays, log_bounces, 3), dtype=f32, device=dev) if log_rays > 0 else None,
            visual=visual,
            ray_times=torch.zeros((ray_times_rays, 2), dtype=torch.int64, device=dev) if ray_times_rays > 0 else None,
        )

    def as_struct(self) -> _lib.TraceOutputs:
        def p(t):
            return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)

        vis = [p(t) for t in self.visual] if self.visual is not None else [ctypes.c_void_p(0)] * 6
        return _lib.TraceOutputs(p(self.echogram), p(self.ambisonic), p(self.hit_counts), p(self.contrib_counts),
                                 p(self.surface_energy), p(self.actual_bounces), p(self.active_count), p(self.counters),
                                 p(self.log_triangle), p(self.log_t), p(self.log_point), *vis, p(self.ray_times))


def make_params(*, num_rays: int, num_bounces: int, num_bins: int, num_diffuse_samples: int, num_sources: int,
                rng_mode: int, use_bvh: bool, render_ambisonics: bool, seed: int, listener, source_forward,
                bin_dur: float, speed_of_sound: float, max_path_len: float, listener_radius: float,
                source_radius: float, irradiance_min_distance: float, specular_exponent: float,
                dipole_weight: float, dipole_power: float, log_rays: int = 0, log_bounces: int = 0,
                tuning: Mapping[str, int] | None = None, visual_limit: int = 0,
                visual_stride: int = 1) -> _lib.TraceParams:
    t = dict(tuning or {})
    p = _lib.TraceParams()
    p.num_rays = int(num_rays)
    p.num_bounces = int(num_bounces)
    p.num_bins = int(num_bins)
    p.num_diffuse_samples = int(num_diffuse_samples)
    p.num_sources = int(num_sources)
    p.rng_mode = int(rng_mode)
    p.use_bvh = 1 if use_bvh else 0
    p.render_ambisonics = 1 if render_ambisonics else 0
    p.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    for i in range(3):
        p.listener[i] = float(listener[i])
        p.source_forward[i] = float(source_forward[i])
    p.bin_duration = float(bin_dur)
    p.speed_of_sound = float(speed_of_sound)
    p.max_path_length = float(max_path_len)
    p.listener_radius = float(listener_radius)
    p.source_radius = float(source_radius)
    p.irradiance_min_distance = float(irradiance_min_distance)
    p.specular_exponent = float(specular_exponent)
    p.dipole_weight = float(dipole_weight)
    p.dipole_power = float(dipole_power)
    p.block_size = int(t.get("block_size", 0))
    p.lds_histogram = int(t.get("lds_histogram", -1))
    p.workgroups_per_cu = int(t.get("workgroups_per_cu", 0))
    p.sources_per_lane = int(t.get("sources_per_lane", 0))
    p.log_rays = int(log_rays)
    p.log_bounces = int(log_bounces)
    p.visual_limit = max(int(visual_limit), 0)
    p.visual_stride = max(int(visual_stride), 1)
    return p


def launch_trace(scene: DeviceScene, params: _lib.TraceParams, sources_dev, bank_dev, buffers: TraceBuffers, *,
                 directions_dev=None, random_dev=None, indices_dev=None, stream=None, plain: bool = False) -> None:
    """Enqueue one trace on ``stream`` (torch stream or None for the current one).

    ``plain`` runs the one-thread-per-ray kernel of the reference GPU
    backend's execution model (``aa_trace_plain``) instead of the persistent
    tracer; the results are the same.
    """
    torch = _torch()
    lib = _lib.load()
    s = stream if stream is not None else torch.cuda.current_stream(scene.device)
    out = buffers.as_struct()

    def p(t):
        return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)

    fn = lib.aa_trace_plain if plain else lib.aa_trace
    _lib.check(fn(scene.handle, ctypes.byref(params), p(sources_dev), p(directions_dev), p(bank_dev),
                  p(random_dev), p(indices_dev), ctypes.byref(out), ctypes.c_void_p(s.cuda_stream)))


_VISUAL_KEYS = ("visual_hit_points", "visual_surface_indices", "visual_ray_indices", "visual_orders",
                "visual_distances", "visual_gains")


def _visual_result(buffers: TraceBuffers, num_bounces: int) -> dict[str, np.ndarray]:
    """The raw visual keys of steam_rt_cuda.py:209-214 (empty when not requested)."""
    if buffers.visual is None:
        empty = (np.zeros((0, num_bounces, 3), np.float32), np.zeros((0, num_bounces), np.int32),
                 np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), np.zeros(0, np.float32))
        return dict(zip(_VISUAL_KEYS, empty))
    return {k: t.cpu().numpy() for k, t in zip(_VISUAL_KEYS, buffers.visual)}


def trace_energy_field_hip(
    *,
    source: np.ndarray | None,
    listener: np.ndarray,
    directions: np.ndarray | None,
    diffuse_bank: np.ndarray,
    diffuse_random: np.ndarray | None,
    diffuse_indices: np.ndarray | None,
    arrays: Mapping[str, Any] | None,
    use_bvh: bool,
    num_bounces: int,
    num_bins: int,
    bin_dur: float,
    speed_of_sound: float,
    max_path_len: float,
    direct_delay: float | None,
    listener_radius: float,
    source_radius: float,
    irradiance_min_distance: float,
    specular_exponent: float,
    source_forward_vector: np.ndarray,
    dipole_weight: float,
    dipole_power: float,
    visual_candidate_limit: int = 0,
    visual_stride: int = 1,
    render_ambisonics: bool = False,
    device_id: int = 0,
    rng_mode: str = "reference",
    seed: int = 0,
    num_rays: int | None = None,
    scene: DeviceScene | None = None,
    tessellation_max_edge: float | None = None,
    sources: np.ndarray | None = None,
    direct_delays: np.ndarray | None = None,
    log_rays: int = 0,
    log_bounces: int = 0,
    tuning: Mapping[str, int] | None = None,
    plain: bool = False,
    ray_times: bool = False,
) -> dict[str, Any]:
    if not hip_available(device_id):
        raise RuntimeError(f"HIP device {device_id} is not available")
    torch = _torch()
    device = int(device_id)
    cache_stats = {"hits": 0, "misses": 0}
    started = time.perf_counter()
    if scene is None:
        if arrays is None:
            raise ValueError("either arrays or scene is required")
        scene = cached_scene(arrays, device=device, max_edge=tessellation_max_edge, stats=cache_stats)
    if rng_mode not in ("reference", "counter"):
        raise ValueError("rng_mode must be 'reference' or 'counter'")
    if rng_mode == "reference":
        if directions is None or diffuse_random is None or diffuse_indices is None:
            raise ValueError("rng_mode='reference' needs directions, diffuse_random and diffuse_indices")
        n_rays = int(np.asarray(directions).shape[0])
        if np.asarray(directions).shape != (n_rays, 3):
            raise ValueError("directions must be [num_rays, 3]")
        if np.asarray(diffuse_random).shape != (int(num_bounces), n_rays) or \
                np.asarray(diffuse_indices).shape != (int(num_bounces), n_rays):
            raise ValueError("diffuse_random and diffuse_indices must be [num_bounces, num_rays]")
        idx = np.asarray(diffuse_indices)
        if idx.size and (int(idx.min()) < 0 or int(idx.max()) >= int(np.asarray(diffuse_bank).shape[0])):
            raise ValueError("diffuse_indices must index the diffuse bank")
    else:
        if num_rays is None:
            raise ValueError("rng_mode='counter' needs num_rays")
        n_rays = int(num_rays)
    if sources is None:
        if source is None or direct_delay is None:
            raise ValueError("source and direct_delay are required")
        src = np.asarray(source, dtype=np.float64).reshape(1, 3)
        delays = np.asarray([float(direct_delay)])
    else:
        src = np.asarray(sources, dtype=np.float64).reshape(-1, 3)
        delays = np.asarray(direct_delays, dtype=np.float64).reshape(-1) if direct_delays is not None else (
            np.linalg.norm(src - np.asarray(listener, dtype=np.float64)[None, :], axis=1) / float(speed_of_sound))
    Ks = int(src.shape[0])
    if not 1 <= Ks <= _lib.AA_MAX_SOURCES:
        raise ValueError(f"between 1 and {_lib.AA_MAX_SOURCES} sources per trace")
    src4 = np.concatenate([src, delays[:, None]], axis=1).astype(np.float32)
    sources_dev = torch.from_numpy(src4).to(f"cuda:{device}")
    bank_dev = _device_input(diffuse_bank, np.float32, device, cache_stats)
    dirs_dev = rand_dev = idx_dev = None
    if rng_mode == "reference":
        dirs_dev = _device_input(directions, np.float32, device, cache_stats)
        rand_dev = _device_input(diffuse_random, np.float32, device, cache_stats)
        idx_dev = _device_input(diffuse_indices, np.int32, device, cache_stats)
    V = max(int(visual_candidate_limit), 0)
    if V > 0 and (sources is not None and Ks != 1):
        raise ValueError("visual candidates need a single-source trace")
    buffers = TraceBuffers.allocate(device, Ks, int(num_bins), scene.num_surfaces, bool(render_ambisonics),
                                    int(log_rays), int(log_bounces), visual_limit=V, num_bounces=int(num_bounces),
                                    ray_times_rays=n_rays if ray_times else 0)
    params = make_params(
        num_rays=n_rays, num_bounces=num_bounces, num_bins=num_bins,
        num_diffuse_samples=int(np.asarray(diffuse_bank).shape[0]), num_sources=Ks,
        rng_mode=_lib.AA_RNG_ARRAYS if rng_mode == "reference" else _lib.AA_RNG_COUNTER, use_bvh=use_bvh,
        render_ambisonics=render_ambisonics, seed=seed, listener=listener, source_forward=source_forward_vector,
        bin_dur=bin_dur, speed_of_sound=speed_of_sound, max_path_len=max_path_len, listener_radius=listener_radius,
        source_radius=source_radius, irradiance_min_distance=irradiance_min_distance,
        specular_exponent=specular_exponent, dipole_weight=dipole_weight, dipole_power=dipole_power,
        log_rays=min(int(log_rays), n_rays), log_bounces=int(log_bounces), tuning=tuning, visual_limit=V,
        visual_stride=visual_stride)
    stream = torch.cuda.current_stream(device)
    stream.synchronize()
    upload_time = time.perf_counter() - started
    start_ev = torch.cuda.Event(enable_timing=True)
    end_ev = torch.cuda.Event(enable_timing=True)
    start_ev.record(stream)
    launch_trace(scene, params, sources_dev, bank_dev, buffers, directions_dev=dirs_dev, random_dev=rand_dev,
                 indices_dev=idx_dev, stream=stream, plain=plain)
    end_ev.record(stream)
    end_ev.synchronize()
    kernel_time = start_ev.elapsed_time(end_ev) / 1000.0
    download_started = time.perf_counter()
    echogram = buffers.echogram.cpu().numpy()
    result: dict[str, Any] = {
        "echogram": echogram[0] if sources is None else echogram,
        "ambisonic": (buffers.ambisonic.cpu().numpy()[0] if sources is None else buffers.ambisonic.cpu().numpy())
        if buffers.ambisonic is not None else np.zeros((6, 4, int(num_bins)), dtype=np.float32),
        "hit_counts": buffers.hit_counts.cpu().numpy()[0] if sources is None else buffers.hit_counts.cpu().numpy(),
        "contrib_counts": (buffers.contrib_counts.cpu().numpy()[0] if sources is None
                           else buffers.contrib_counts.cpu().numpy()),
        "surface_energy": (buffers.surface_energy.cpu().numpy()[0] if sources is None
                           else buffers.surface_energy.cpu().numpy()),
        "actual_bounces": (int(buffers.actual_bounces.cpu()[0]) if sources is None
                           else buffers.actual_bounces.cpu().numpy()),
        "active_count": int(buffers.active_count.cpu()[0]) if sources is None else buffers.active_count.cpu().numpy(),
        **_visual_result(buffers, int(num_bounces)),
        "counters": dict(zip(_lib.COUNTER_NAMES, (int(x) for x in buffers.counters.cpu().numpy()))),
    }
    if buffers.log_triangle is not None:
        result["log_triangle"] = buffers.log_triangle.cpu().numpy()
        result["log_t"] = buffers.log_t.cpu().numpy()
        result["log_point"] = buffers.log_point.cpu().numpy()
    if buffers.ray_times is not None:
        result["ray_times"] = buffers.ray_times.cpu().numpy()
    result["kernel_time_s"] = kernel_time
    result["transfer_time_s"] = upload_time + (time.perf_counter() - download_started)
    info = device_info(device)
    result["device"] = {"id": device, "name": info["name"], "compute_capability": info["arch"]}
    result["device_input_cache"] = cache_stats
    result["scene"] = dict(scene.info)
    return result
