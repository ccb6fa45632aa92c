"""Compact binary scene files (``.aascene``) for the serving scene cache.

A compiled scene (triangles with their reference surface ids and normals,
per-surface materials) is stored so that a server loads it without the scene
compiler, and keys it by the same content hash ``aa_scene_create`` computes
(``aa_scene_hash``).  Layout, little-endian, no padding:

    offset  size  field
    0       8     magic b"AASCENE1"
    8       4     u32 version (1)
    12      4     u32 T, triangles
    16      4     u32 S, surfaces
    20      4     i32 leaf_size of the BVH build
    24      8     u64 content hash: aa_scene_hash of the payload arrays and leaf_size
    32      4     u32 bytes of surface names (UTF-8, newline-separated; may be 0)
    36      4     u32 0 (reserved)
    40      ...   f32 vertices [T,9] | i32 surface [T] | f32 normals [T,3]
                  | f32 reflection [S,6] | f32 scattering [S] | names

``load_scene`` checks the magic, version and exact file length, and
recomputes the hash, so a truncated or corrupted file raises ``ValueError``
instead of producing a different scene.
"""
from __future__ import annotations

import ctypes
import os
import struct
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _lib
from .scene import TriangleMesh

MAGIC = b"AASCENE1"
VERSION = 1
_HEADER = struct.Struct("<8sIIIiQII")  # 40 bytes


def content_hash(vertices: np.ndarray, surface: np.ndarray, normals: np.ndarray, reflection: np.ndarray,
                 scattering: np.ndarray, leaf_size: int = 4) -> int:
    """``aa_scene_hash`` of the float32 / int32 forms of the arrays."""
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 9)
    s = np.ascontiguousarray(surface, dtype=np.int32).reshape(-1)
    n = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
    r = np.ascontiguousarray(reflection, dtype=np.float32).reshape(-1, 6)
    c = np.ascontiguousarray(scattering, dtype=np.float32).reshape(-1)
    if not (v.shape[0] == s.shape[0] == n.shape[0]) or r.shape[0] != c.shape[0]:
        raise ValueError("inconsistent scene array sizes")
    out = ctypes.c_uint64()
    p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    _lib.check(_lib.load().aa_scene_hash(p(v), p(s), p(n), int(v.shape[0]), p(r), p(c), int(r.shape[0]),
                                         int(leaf_size), ctypes.byref(out)))
    return int(out.value)


@dataclass
class SceneFile:
    mesh: TriangleMesh
    reflection: np.ndarray  # [S,6] float32
    scattering: np.ndarray  # [S] float32
    leaf_size: int
    content_hash: int
    surface_names: tuple[str, ...] = ()


def save_scene(path: str | os.PathLike, mesh: TriangleMesh, reflection: np.ndarray, scattering: np.ndarray,
               *, leaf_size: int = 4, surface_names: Sequence[str] = ()) -> int:
    """Write ``path``; returns the content hash."""
    v = np.ascontiguousarray(mesh.vertices, dtype=np.float32).reshape(-1, 9)
    s = np.ascontiguousarray(mesh.surface, dtype=np.int32).reshape(-1)
    n = np.ascontiguousarray(mesh.normals, dtype=np.float32).reshape(-1, 3)
    r = np.ascontiguousarray(reflection, dtype=np.float32).reshape(-1, 6)
    c = np.ascontiguousarray(scattering, dtype=np.float32).reshape(-1)
    if r.shape[0] != mesh.num_surfaces or c.shape[0] != mesh.num_surfaces:
        raise ValueError("reflection must be [S,6] and scattering [S] for the mesh's S surfaces")
    if surface_names and len(surface_names) != mesh.num_surfaces:
        raise ValueError("surface_names must name every surface")
    if any("\n" in name for name in surface_names):
        raise ValueError("surface names must not contain a newline")
    names = "\n".join(surface_names).encode("utf-8")
    h = content_hash(v, s, n, r, c, leaf_size)
    header = _HEADER.pack(MAGIC, VERSION, v.shape[0], r.shape[0], int(leaf_size), h, len(names), 0)
    tmp = f"{os.fspath(path)}.tmp"
    with open(tmp, "wb") as f:
        f.write(header)
        for a in (v, s, n, r, c):
            f.write(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes())
        f.write(names)
    os.replace(tmp, path)
    return h


def load_scene(path: str | os.PathLike) -> SceneFile:
    """Read and verify ``path``."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < _HEADER.size:
        raise ValueError(f"{path}: shorter than a scene header")
    magic, version, T, S, leaf_size, h, name_bytes, _ = _HEADER.unpack_from(data, 0)
    if magic != MAGIC:
        raise ValueError(f"{path}: not a scene file")
    if version != VERSION:
        raise ValueError(f"{path}: unsupported scene file version {version}")
    sizes = (T * 9 * 4, T * 4, T * 3 * 4, S * 6 * 4, S * 4)
    if len(data) != _HEADER.size + sum(sizes) + name_bytes:
        raise ValueError(f"{path}: length {len(data)} does not match its header")
    off = _HEADER.size
    parts = []
    for size, dtype, shape in zip(sizes, ("<f4", "<i4", "<f4", "<f4", "<f4"), ((T, 9), (T,), (T, 3), (S, 6), (S,))):
        parts.append(np.frombuffer(data, dtype=dtype, count=size // 4, offset=off).reshape(shape).copy())
        off += size
    v, s, n, r, c = parts
    names = tuple(data[off:off + name_bytes].decode("utf-8").split("\n")) if name_bytes else ()
    if content_hash(v, s, n, r, c, leaf_size) != h:
        raise ValueError(f"{path}: content hash mismatch (corrupted file)")
    if T and (s.min() < 0 or s.max() >= S):
        raise ValueError(f"{path}: surface id out of range")
    mesh = TriangleMesh(vertices=v.astype(np.float64), surface=s, normals=n.astype(np.float64), num_surfaces=int(S))
    return SceneFile(mesh, r, c, int(leaf_size), int(h), names)


def device_scene_from_file(path: str | os.PathLike, device: int = 0):
    """A resident ``DeviceScene`` of a scene file (its hash is checked against the device's)."""
    from .hip_backend import DeviceScene

    f = load_scene(path)
    ds = DeviceScene(f.mesh, f.reflection, f.scattering, device=device, leaf_size=f.leaf_size)
    if int(ds.info["content_hash"]) != f.content_hash:
        raise RuntimeError(f"{path}: device scene hash differs from the file's")
    return ds
