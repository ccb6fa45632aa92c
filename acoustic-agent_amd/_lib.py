"""ctypes binding of ``libaa_tracer.so`` (C ABI in ``include/aa_tracer.h``).

The library is loaded from the package directory only.  If it is missing the
import of this module raises: the tracer has no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
LIB_PATH = PACKAGE_DIR / "libaa_tracer.so"

AA_OK = 0
AA_ERR_INVALID = -1
AA_ERR_DEVICE = -2
AA_ERR_MEMORY = -3
AA_ERR_LIMIT = -4

AA_RNG_ARRAYS = 0
AA_RNG_COUNTER = 1
AA_NUM_BANDS = 6
AA_MAX_SOURCES = 64
AA_CNT_COUNT = 9
AA_STATIC_BATCH_TRACES = 256

COUNTER_NAMES = (
    "bounce_queries",
    "shadow_queries",
    "nodes_bounce",
    "nodes_shadow",
    "triangles_bounce",
    "triangles_shadow",
    "contributions",
    "escaped_rays",
    "rays_claimed",
)


class SceneInfo(ctypes.Structure):
    _fields_ = [
        ("num_triangles", ctypes.c_int32),
        ("num_surfaces", ctypes.c_int32),
        ("num_nodes", ctypes.c_int32),
        ("depth", ctypes.c_int32),
        ("canonical_nodes", ctypes.c_int32),
        ("canonical_depth", ctypes.c_int32),
        ("leaf_size", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("device_bytes", ctypes.c_uint64),
        ("content_hash", ctypes.c_uint64),
    ]


class TraceParams(ctypes.Structure):
    _fields_ = [
        ("num_rays", ctypes.c_int32),
        ("num_bounces", ctypes.c_int32),
        ("num_bins", ctypes.c_int32),
        ("num_diffuse_samples", ctypes.c_int32),
        ("num_sources", ctypes.c_int32),
        ("rng_mode", ctypes.c_int32),
        ("use_bvh", ctypes.c_int32),
        ("render_ambisonics", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("listener", ctypes.c_float * 3),
        ("source_forward", ctypes.c_float * 3),
        ("bin_duration", ctypes.c_float),
        ("speed_of_sound", ctypes.c_float),
        ("max_path_length", ctypes.c_float),
        ("listener_radius", ctypes.c_float),
        ("source_radius", ctypes.c_float),
        ("irradiance_min_distance", ctypes.c_float),
        ("specular_exponent", ctypes.c_float),
        ("dipole_weight", ctypes.c_float),
        ("dipole_power", ctypes.c_float),
        ("block_size", ctypes.c_int32),
        ("lds_histogram", ctypes.c_int32),
        ("workgroups_per_cu", ctypes.c_int32),
        ("sources_per_lane", ctypes.c_int32),
        ("log_rays", ctypes.c_int32),
        ("log_bounces", ctypes.c_int32),
        ("visual_limit", ctypes.c_int32),
        ("visual_stride", ctypes.c_int32),
    ]


class TraceOutputs(ctypes.Structure):
    _fields_ = [
        ("d_echogram", ctypes.c_void_p),
        ("d_ambisonic", ctypes.c_void_p),
        ("d_hit_counts", ctypes.c_void_p),
        ("d_contrib_counts", ctypes.c_void_p),
        ("d_surface_energy", ctypes.c_void_p),
        ("d_actual_bounces", ctypes.c_void_p),
        ("d_active_count", ctypes.c_void_p),
        ("d_counters", ctypes.c_void_p),
        ("d_log_triangle", ctypes.c_void_p),
        ("d_log_t", ctypes.c_void_p),
        ("d_log_point", ctypes.c_void_p),
        ("d_visual_points", ctypes.c_void_p),
        ("d_visual_surfaces", ctypes.c_void_p),
        ("d_visual_rays", ctypes.c_void_p),
        ("d_visual_orders", ctypes.c_void_p),
        ("d_visual_distances", ctypes.c_void_p),
        ("d_visual_gains", ctypes.c_void_p),
        ("d_ray_times", ctypes.c_void_p),
    ]


class BatchTrace(ctypes.Structure):
    _fields_ = [
        ("scene", ctypes.c_void_p),
        ("params", TraceParams),
        ("d_sources", ctypes.c_void_p),
        ("d_directions", ctypes.c_void_p),
        ("d_diffuse_bank", ctypes.c_void_p),
        ("d_diffuse_random", ctypes.c_void_p),
        ("d_diffuse_indices", ctypes.c_void_p),
        ("outputs", TraceOutputs),
    ]


class BatchTuning(ctypes.Structure):
    _fields_ = [
        ("block_size", ctypes.c_int32),
        ("packet_rays", ctypes.c_int32),
        ("workgroups_per_cu", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64

_SIGNATURES = {
    "aa_last_error": (ctypes.c_char_p, []),
    "aa_device_count": (_I32, [ctypes.POINTER(_I32)]),
    "aa_device_info": (
        _I32,
        [_I32, ctypes.c_char_p, _I32, ctypes.c_char_p, _I32, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(_I32)],
    ),
    "aa_build_bvh": (_I32, [_P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P]),
    "aa_build_bvh_sah": (_I32, [_P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P]),
    "aa_scene_create": (_I32, [_I32, _P, _P, _P, _I32, _P, _P, _I32, _I32, ctypes.POINTER(_P)]),
    "aa_scene_info": (_I32, [_P, ctypes.POINTER(SceneInfo)]),
    "aa_scene_hash": (_I32, [_P, _P, _P, _I32, _P, _P, _I32, _I32, ctypes.POINTER(ctypes.c_uint64)]),
    "aa_scene_release": (_I32, [_P]),
    "aa_trace": (_I32, [_P, ctypes.POINTER(TraceParams), _P, _P, _P, _P, _P, ctypes.POINTER(TraceOutputs), _P]),
    "aa_trace_plain": (_I32, [_P, ctypes.POINTER(TraceParams), _P, _P, _P, _P, _P, ctypes.POINTER(TraceOutputs),
                              _P]),
    "aa_stream_sync": (_I32, [_P]),
    "aa_stream_query": (_I32, [_P, ctypes.POINTER(_I32)]),
    "aa_batch_create": (_I32, [_I32, _I32, ctypes.POINTER(_P)]),
    "aa_batch_submit": (_I32, [_P, ctypes.POINTER(BatchTrace), _I32, ctypes.POINTER(BatchTuning), _P]),
    "aa_batch_release": (_I32, [_P]),
    "aa_batch_submit_static": (_I32, [_I32, ctypes.POINTER(BatchTrace), _I32, ctypes.POINTER(BatchTuning), _P]),
    "aa_fdn_impulse": (_I32, [_P, _I32, _P, _I32, _I64, _P, _I32, _P, _P]),
}

_lock = threading.Lock()
_lib: ctypes.CDLL | None = None


class TracerError(RuntimeError):
    """A C ABI call failed; ``code`` is the AA_ERR_* value."""

    def __init__(self, code: int, message: str) -> None:
        super().__init__(message)
        self.code = code


def load() -> ctypes.CDLL:
    """Load ``libaa_tracer.so`` (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(the tracer has no CPU fallback)"
            )
        # PyTorch-ROCm ships its own libamdhip64.so.7 / libhsa-runtime64.so.1.
        # Loading torch first makes this library bind to that same runtime
        # (same SONAME); loading it first would bring a second HIP runtime into
        # the process and torch would then see no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(os.fspath(LIB_PATH))
        for name, (restype, argtypes) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = lib
        return lib


def exported_symbols() -> tuple[str, ...]:
    return tuple(_SIGNATURES)


def check(code: int) -> None:
    """Raise the Python exception the reference raises for this failure."""
    if code == AA_OK:
        return
    lib = load()
    message = (lib.aa_last_error() or b"").decode("utf-8", "replace")
    if code == AA_ERR_INVALID:
        raise ValueError(message)
    raise TracerError(code, message)
