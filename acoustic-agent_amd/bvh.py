"""Canonical BVH builder: the reference's ``build_bvh`` (acoustic_agent/bvh.py:8-77).

Same signature, validation and output dictionary; the construction runs in
the native library (``aa_build_bvh``) so that 50k-primitive scenes build in
milliseconds.
"""
from __future__ import annotations

import ctypes
from typing import Any

import numpy as np

from . import _lib


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def build_bvh(bounds_min: np.ndarray, bounds_max: np.ndarray, *, leaf_size: int = 4, sah: bool = False) -> dict[str, Any]:
    """``sah=True`` builds the device tree (binned SAH) in the same layout."""
    lower = np.ascontiguousarray(bounds_min, dtype=np.float64)
    upper = np.ascontiguousarray(bounds_max, dtype=np.float64)
    if lower.shape != upper.shape or lower.ndim != 2 or lower.shape[1] != 3:
        raise ValueError("BVH bounds must have shape [surface, 3]")
    if lower.shape[0] < 1:
        raise ValueError("BVH requires at least one surface")
    n = int(lower.shape[0])
    cap = max(1, 2 * n - 1)
    node_min = np.zeros((cap, 3))
    node_max = np.zeros((cap, 3))
    start = np.zeros(cap, dtype=np.int64)
    count = np.zeros(cap, dtype=np.int64)
    escape = np.zeros(cap, dtype=np.int64)
    prims = np.zeros(n, dtype=np.int64)
    stats = np.zeros(3, dtype=np.int64)
    lib = _lib.load()
    _lib.check((lib.aa_build_bvh_sah if sah else lib.aa_build_bvh)(_ptr(lower), _ptr(upper), n, max(1, int(leaf_size)), _ptr(node_min), _ptr(node_max),
                                _ptr(start), _ptr(count), _ptr(escape), _ptr(prims), _ptr(stats)))
    nodes = int(stats[0])
    return {
        "bounds_min": node_min[:nodes].copy(),
        "bounds_max": node_max[:nodes].copy(),
        "start": start[:nodes].copy(),
        "count": count[:nodes].copy(),
        "escape": escape[:nodes].copy(),
        "primitives": prims,
        "node_count": nodes,
        "leaf_count": int(stats[1]),
        "max_depth": int(stats[2]),
        "leaf_size": max(1, int(leaf_size)),
    }
