"""ctypes binding of the oracle library ``liboracle.so``.

TEST INFRASTRUCTURE ONLY.  The oracle is a CPU restatement of the reference
tracer ``_trace_energy_kernel`` (acoustic_agent/steam_rt.py:4708-4886) and
its helpers, in binary64 (the reference Numba path) and binary32 (the
reference CUDA path's arithmetic), over either the reference's analytic
primitives (walls, floor/ceiling polygon, boxes) or their triangle
tessellation.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import it, and only as the checker.

Pins (see tests/test_oracle.py):
  primitives    analytic known answers (exact t in a shoebox, boxes from
                inside/outside, L-polygon inclusion, parallel thresholds)
  BVH           ao_build_bvh vs hand-checked trees; BVH == linear closest hit
  triangles     triangle mode == analytic mode on the same rooms
  deposit       single-ray closed form (bin, six band energies)
  RNG           Philox4x32-10 known-answer vectors; reference-mode streams
                regenerated with NumPy (the reference's call order)
  end to end    Eyring RT60 of a diffuse shoebox (benchmark.py:268-298),
                FP32 vs FP64 within the reference's tolerance
                (tests/test_accelerators.py:31-53), FOA direction invariant
                (tests/test_steam_audio_parity.py:329-357)
  visual        candidate surfaces = the hit log while the ray lives, step
                lengths = logged t (+- the hit offset); a ray traced alone:
                strongest contribution = largest band-summed bin, and its
                path length maps back to that bin
Every function is pinned by at least one of these; none is unpinned.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path
from typing import Any, Mapping

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "liboracle.so"
NB = 6

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64


def build() -> Path:
    subprocess.run(["make", "-C", str(ORACLE_DIR), "-s"], check=True)
    return LIB_PATH


def _structs(real):
    class Scene(ctypes.Structure):
        _fields_ = [
            ("mode", _I32), ("num_surfaces", _I32), ("kinds", _P),
            ("wall_a", _P), ("wall_delta", _P), ("wall_z", _P), ("z_values", _P),
            ("box_center", _P), ("box_axis_u", _P), ("box_axis_v", _P), ("box_half", _P), ("box_z", _P),
            ("normals", _P), ("corners", _P), ("num_corners", _I32),
            ("num_triangles", _I32), ("tri_v", _P), ("tri_surface", _P), ("tri_normal", _P),
            ("use_bvh", _I32), ("num_nodes", _I32), ("bvh_min", _P), ("bvh_max", _P),
            ("bvh_start", _P), ("bvh_count", _P), ("bvh_escape", _P), ("bvh_prims", _P),
            ("reflection", _P), ("scattering", _P),
        ]

    class Params(ctypes.Structure):
        _fields_ = [
            ("num_rays", _I32), ("num_bounces", _I32), ("num_bins", _I32), ("num_diffuse_samples", _I32),
            ("rng_mode", _I32), ("render_ambisonics", _I32), ("shadow_closest", _I32), ("num_threads", _I32),
            ("seed", ctypes.c_uint64),
            ("source", real * 3), ("listener", real * 3), ("source_forward", real * 3),
            ("bin_dur", real), ("c", real), ("max_path_len", real), ("direct_delay", real),
            ("listener_radius", real), ("source_radius", real), ("irradiance_min_distance", real),
            ("specular_exponent", real), ("dipole_weight", real), ("dipole_power", real),
            ("directions", _P), ("diffuse_bank", _P), ("diffuse_random", _P), ("diffuse_indices", _P),
            ("log_rays", _I32), ("log_bounces", _I32), ("ray_begin", _I32), ("ray_end", _I32),
            ("visual_limit", _I32), ("visual_stride", _I32),
        ]

    class Out(ctypes.Structure):
        _fields_ = [
            ("echogram", _P), ("ambisonic", _P), ("hit_counts", _P), ("contrib_counts", _P),
            ("surface_energy", _P), ("log_prim", _P), ("log_surface", _P), ("log_t", _P), ("log_point", _P),
            ("visual_points", _P), ("visual_surfaces", _P), ("visual_rays", _P), ("visual_orders", _P),
            ("visual_distances", _P), ("visual_gains", _P),
            ("actual_bounces", _I64), ("active_count", _I64), ("bounce_queries", _I64), ("shadow_queries", _I64),
            ("nodes_bounce", _I64), ("nodes_shadow", _I64), ("prims_bounce", _I64), ("prims_shadow", _I64),
            ("contributions", _I64), ("escaped", _I64),
        ]

    return Scene, Params, Out


_STRUCTS = {"float64": _structs(ctypes.c_double), "float32": _structs(ctypes.c_float)}
_lib: ctypes.CDLL | None = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(LIB_PATH))
        for suffix, prec in (("f64", "float64"), ("f32", "float32")):
            Scene, Params, Out = _STRUCTS[prec]
            fn = getattr(lib, f"ao_trace_{suffix}")
            fn.restype = _I32
            fn.argtypes = [ctypes.POINTER(Scene), ctypes.POINTER(Params), ctypes.POINTER(Out)]
            fn = getattr(lib, f"ao_closest_hit_{suffix}")
            fn.restype = _I32
            fn.argtypes = [ctypes.POINTER(Scene), _P, _P, ctypes.POINTER(_I32), ctypes.POINTER(_I32), _P, _P]
            fn = getattr(lib, f"ao_diffuse_direction_{suffix}")
            fn.restype = None
            fn.argtypes = [_P, _P, _P]
        lib.ao_build_bvh.restype = _I32
        lib.ao_build_bvh.argtypes = [_P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P]
        lib.ao_philox4x32_10.restype = None
        lib.ao_philox4x32_10.argtypes = [_P, _P, _P]
        lib.ao_counter_direction.restype = None
        lib.ao_counter_direction.argtypes = [ctypes.c_uint64, _I64, _I64, _P]
        lib.ao_counter_bounce.restype = None
        lib.ao_counter_bounce.argtypes = [ctypes.c_uint64, _I64, _I32, _I32, _P, _P]
        _lib = lib
    return _lib


def _p(a: np.ndarray | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(0) if a is None else ctypes.c_void_p(a.ctypes.data)


def philox(counter, key) -> np.ndarray:
    c = np.ascontiguousarray(counter, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    load().ao_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def counter_direction(seed: int, ray: int, num_rays: int) -> np.ndarray:
    out = np.zeros(3)
    load().ao_counter_direction(int(seed), int(ray), int(num_rays), _p(out))
    return out


def counter_bounce(seed: int, ray: int, bounce: int, num_samples: int) -> tuple[int, int]:
    u = np.zeros(1, dtype=np.uint32)
    j = np.zeros(1, dtype=np.int32)
    load().ao_counter_bounce(int(seed), int(ray), int(bounce), int(num_samples), _p(u), _p(j))
    return int(u[0]), int(j[0])


def diffuse_direction(sample, normal, precision: str = "float64") -> np.ndarray:
    """The oracle's diffuse continuation: bank sample ``sample`` turned into
    the frame of the oriented normal ``normal`` (steam_rt.py:4564-4582)."""
    real = np.float64 if precision == "float64" else np.float32
    s = np.ascontiguousarray(sample, dtype=real)
    n = np.ascontiguousarray(normal, dtype=real)
    out = np.zeros(3, dtype=real)
    getattr(load(), "ao_diffuse_direction_" + ("f64" if precision == "float64" else "f32"))(_p(s), _p(n), _p(out))
    return out.astype(np.float64)


def build_bvh(lower: np.ndarray, upper: np.ndarray, leaf_size: int = 4) -> dict[str, Any]:
    """The oracle's own restatement of bvh.py:8-77."""
    lo = np.ascontiguousarray(lower, dtype=np.float64)
    hi = np.ascontiguousarray(upper, dtype=np.float64)
    n = int(lo.shape[0])
    cap = max(1, 2 * n - 1)
    out = {k: np.zeros((cap, 3)) for k in ("bounds_min", "bounds_max")}
    for k in ("start", "count", "escape"):
        out[k] = np.zeros(cap, dtype=np.int64)
    prims = np.zeros(max(n, 1), dtype=np.int64)
    stats = np.zeros(3, dtype=np.int64)
    rc = load().ao_build_bvh(_p(lo), _p(hi), n, int(leaf_size), _p(out["bounds_min"]), _p(out["bounds_max"]),
                             _p(out["start"]), _p(out["count"]), _p(out["escape"]), _p(prims), _p(stats))
    if rc != 0:
        raise ValueError(f"oracle BVH build failed ({rc})")
    nodes = int(stats[0])
    res = {k: v[:nodes].copy() for k, v in out.items()}
    res.update(primitives=prims[:n], node_count=nodes, leaf_count=int(stats[1]), max_depth=int(stats[2]))
    return res


class OracleScene:
    """Arrays of one scene in one precision, kept alive for the C structs."""

    def __init__(self, arrays: Mapping[str, Any], mesh=None, precision: str = "float64", use_bvh: bool = True,
                 leaf_size: int = 4, bvh: Mapping[str, Any] | None = None) -> None:
        self.precision = precision
        real = np.float64 if precision == "float64" else np.float32
        Scene, _, _ = _STRUCTS[precision]
        keep: list[np.ndarray] = []

        def r(name_or_array, dtype=real):
            a = np.ascontiguousarray(arrays[name_or_array] if isinstance(name_or_array, str) else name_or_array,
                                     dtype=dtype)
            keep.append(a)
            return a

        sc = Scene()
        S = int(np.asarray(arrays["kinds"]).shape[0])
        sc.num_surfaces = S
        sc.kinds = _p(r("kinds", np.int64))
        for name in ("wall_a", "wall_delta", "wall_z", "z_values", "box_center", "box_axis_u", "box_axis_v",
                     "box_half", "box_z", "normals", "corners", "reflection", "scattering"):
            setattr(sc, name, _p(r(name)))
        sc.num_corners = int(np.asarray(arrays["corners"]).shape[0])
        if mesh is None:
            sc.mode = 0
            lo = np.asarray(arrays["surface_bounds_min"], dtype=np.float64)
            hi = np.asarray(arrays["surface_bounds_max"], dtype=np.float64)
            self.num_primitives = S
        else:
            sc.mode = 1
            tri_v = r(mesh.vertices)
            sc.tri_v = _p(tri_v)
            sc.tri_surface = _p(r(mesh.surface, np.int32))
            sc.tri_normal = _p(r(mesh.normals))
            sc.num_triangles = int(mesh.vertices.shape[0])
            pad = 2.0e-6 if precision == "float64" else 1.0e-5
            v = tri_v.astype(np.float64).reshape(-1, 3, 3)
            lo, hi = v.min(axis=1) - pad, v.max(axis=1) + pad
            self.num_primitives = sc.num_triangles
        if bvh is None:
            bvh = build_bvh(lo, hi, leaf_size)
        self.bvh = bvh
        sc.use_bvh = 1 if use_bvh else 0
        sc.num_nodes = int(bvh["node_count"])
        sc.bvh_min = _p(r(bvh["bounds_min"]))
        sc.bvh_max = _p(r(bvh["bounds_max"]))
        for name, field in (("start", "bvh_start"), ("count", "bvh_count"), ("escape", "bvh_escape"),
                            ("primitives", "bvh_prims")):
            setattr(sc, field, _p(r(bvh[name], np.int64)))
        self.struct = sc
        self._keep = keep
        self.num_surfaces = S

    def closest_hit(self, origin, direction):
        real = np.float64 if self.precision == "float64" else np.float32
        o = np.ascontiguousarray(origin, dtype=real)
        d = np.ascontiguousarray(direction, dtype=real)
        t = np.zeros(1, dtype=real)
        n = np.zeros(3, dtype=real)
        prim, surf = _I32(0), _I32(0)
        fn = getattr(load(), "ao_closest_hit_" + ("f64" if self.precision == "float64" else "f32"))
        fn(ctypes.byref(self.struct), _p(o), _p(d), ctypes.byref(prim), ctypes.byref(surf), _p(t), _p(n))
        return int(prim.value), int(surf.value), float(t[0]), n.astype(np.float64)


def trace(scene: OracleScene, *, source, listener, num_rays: int, num_bounces: int, num_bins: int, bin_dur: float,
          c: float = 343.0, max_path_len: float | None = None, direct_delay: float | None = None,
          listener_radius: float = 0.1, source_radius: float = 0.1, irradiance_min_distance: float = 1.0,
          specular_exponent: float = 100.0, source_forward=(1.0, 0.0, 0.0), dipole_weight: float = 0.0,
          dipole_power: float = 1.0, diffuse_bank: np.ndarray, rng: str = "reference", directions=None,
          diffuse_random=None, diffuse_indices=None, seed: int = 0, render_ambisonics: bool = False,
          shadow_closest: bool = False, threads: int = 0, log_rays: int = 0, log_bounces: int = 0,
          ray_range: tuple[int, int] | None = None, visual_limit: int = 0,
          visual_stride: int = 1) -> dict[str, Any]:
    prec = scene.precision
    real = np.float64 if prec == "float64" else np.float32
    _, Params, Out = _STRUCTS[prec]
    keep: list[np.ndarray] = []

    def r(a, dtype=real):
        arr = np.ascontiguousarray(a, dtype=dtype)
        keep.append(arr)
        return arr

    p = Params()
    p.num_rays, p.num_bounces, p.num_bins = int(num_rays), int(num_bounces), int(num_bins)
    bank = r(diffuse_bank)
    p.num_diffuse_samples = int(bank.shape[0])
    p.rng_mode = 0 if rng == "reference" else 1
    p.render_ambisonics = 1 if render_ambisonics else 0
    p.shadow_closest = 1 if shadow_closest else 0
    p.num_threads = int(threads)
    p.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    src = np.asarray(source, dtype=np.float64)
    lis = np.asarray(listener, dtype=np.float64)
    for i in range(3):
        p.source[i] = real(src[i])
        p.listener[i] = real(lis[i])
        p.source_forward[i] = real(source_forward[i])
    if max_path_len is None:
        max_path_len = num_bins * bin_dur * c
    if direct_delay is None:
        direct_delay = float(np.linalg.norm(src - lis)) / c
    p.bin_dur, p.c, p.max_path_len, p.direct_delay = real(bin_dur), real(c), real(max_path_len), real(direct_delay)
    p.listener_radius, p.source_radius = real(listener_radius), real(source_radius)
    p.irradiance_min_distance = real(irradiance_min_distance)
    p.specular_exponent, p.dipole_weight, p.dipole_power = real(specular_exponent), real(dipole_weight), real(
        dipole_power)
    p.diffuse_bank = _p(bank)
    if p.rng_mode == 0:
        p.directions = _p(r(directions))
        p.diffuse_random = _p(r(diffuse_random))
        p.diffuse_indices = _p(r(diffuse_indices, np.int32))
    p.log_rays, p.log_bounces = int(log_rays), int(log_bounces)
    if ray_range is not None:
        p.ray_begin, p.ray_end = int(ray_range[0]), int(ray_range[1])
    S = scene.num_surfaces
    out = Out()
    echogram = np.zeros((NB, num_bins), dtype=real)
    amb = np.zeros((NB, 4, num_bins), dtype=real) if render_ambisonics else None
    hits = np.zeros(S, dtype=np.int64)
    contribs = np.zeros(S, dtype=np.int64)
    senergy = np.zeros(S, dtype=real)
    log_prim = np.full((max(log_rays, 0), max(log_bounces, 0)), -2, dtype=np.int32)
    log_surf = np.full_like(log_prim, -2)
    log_t = np.zeros(log_prim.shape)
    log_point = np.zeros(log_prim.shape + (3,))
    out.echogram, out.ambisonic = _p(echogram), _p(amb)
    out.hit_counts, out.contrib_counts, out.surface_energy = _p(hits), _p(contribs), _p(senergy)
    if log_rays > 0 and log_bounces > 0:
        out.log_prim, out.log_surface, out.log_t = _p(log_prim), _p(log_surf), _p(log_t)
        out.log_point = _p(log_point)
    V = max(int(visual_limit), 0)
    p.visual_limit, p.visual_stride = V, max(int(visual_stride), 1)
    vis = {
        "visual_hit_points": np.zeros((V, num_bounces, 3), dtype=real),
        "visual_surface_indices": np.full((V, num_bounces), -1, dtype=np.int32),
        "visual_ray_indices": np.full(V, -1, dtype=np.int32),
        "visual_orders": np.zeros(V, dtype=np.int32),
        "visual_distances": np.zeros(V, dtype=real),
        "visual_gains": np.zeros(V, dtype=real),
    }
    if V > 0:
        out.visual_points, out.visual_surfaces = _p(vis["visual_hit_points"]), _p(vis["visual_surface_indices"])
        out.visual_rays, out.visual_orders = _p(vis["visual_ray_indices"]), _p(vis["visual_orders"])
        out.visual_distances, out.visual_gains = _p(vis["visual_distances"]), _p(vis["visual_gains"])
    fn = getattr(load(), "ao_trace_" + ("f64" if prec == "float64" else "f32"))
    rc = fn(ctypes.byref(scene.struct), ctypes.byref(p), ctypes.byref(out))
    if rc != 0:
        raise ValueError(f"oracle trace failed ({rc})")
    res = {
        "echogram": echogram,
        "ambisonic": amb,
        "hit_counts": hits,
        "contrib_counts": contribs,
        "surface_energy": senergy,
        "actual_bounces": int(out.actual_bounces),
        "active_count": int(out.active_count),
        "counters": {name: int(getattr(out, name)) for name in (
            "bounce_queries", "shadow_queries", "nodes_bounce", "nodes_shadow", "prims_bounce", "prims_shadow",
            "contributions", "escaped")},
    }
    if log_rays > 0 and log_bounces > 0:
        res.update(log_prim=log_prim, log_surface=log_surf, log_t=log_t, log_point=log_point)
    if V > 0:
        res.update(vis)
    return res
