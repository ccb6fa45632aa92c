"""Scene compiler: room description -> reference surface arrays -> triangles.

``RoomScene`` builds the surface list in the reference order
(acoustic_agent/steam_rt.py:291-302: walls, floor, ceiling, boxes) with the
reference surface semantics:
  walls       steam_rt.py:83-131 (vertical quad, normal (dy, -dx, 0)/|d|)
  floor/ceil  steam_rt.py:134-148 (one shared outline polygon, normal +-z)
  boxes       steam_rt.py:167-183 (oriented solid; size >= 1e-3; z range)
and ``kernel_arrays`` returns the arrays of ``_scene_kernel_arrays``
(steam_rt.py:834-924): kinds, wall_*, z_values, box_*, normals, absorption,
reflection = 1 - absorption, scattering = mean over bands, names, corners,
surface bounds padded by 2e-6 (steam_rt.py:234-262) and the canonical BVH.

``tessellate`` turns those arrays into a triangle soup for the HIP tracer:
every triangle carries the id of the reference surface it covers, and its
normal is that surface's normal (for boxes: the face's outward axis), so
closest-hit surface ids and reflection geometry are those of the reference
primitives.  Optional midpoint refinement (1 -> 4 triangles) bounds the edge
length; coplanar sub-triangles keep the surface id.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Mapping, Sequence

import numpy as np

from .models import FREQUENCY_BANDS

NUM_BANDS = len(FREQUENCY_BANDS)
BVH_BOUNDS_EPS = 2.0e-6  # steam_rt.py:43
EPS = 1e-6


@dataclass(frozen=True)
class Material:
    """Per-surface acoustic data: six-band absorption and scalar scattering."""

    absorption: tuple[float, ...]
    scattering: float = 0.1

    def __post_init__(self) -> None:
        if len(self.absorption) != NUM_BANDS:
            raise ValueError("absorption needs six bands")


@dataclass(frozen=True)
class WallSpec:
    a: tuple[float, float]
    b: tuple[float, float]
    z_min: float = 0.0
    z_max: float | None = None
    name: str = "wall"
    material: Material | None = None


@dataclass(frozen=True)
class BoxSpec:
    center: tuple[float, float]
    size: tuple[float, float, float]
    rotation_deg: float = 0.0
    z_center: float | None = None
    name: str = "box"
    material: Material | None = None


@dataclass(frozen=True)
class RoomSpec:
    corners: tuple[tuple[float, float], ...]
    height: float
    wall_material: Material
    floor_material: Material | None = None
    ceiling_material: Material | None = None
    walls: tuple[WallSpec, ...] = ()  # empty: one wall per outline edge
    boxes: tuple[BoxSpec, ...] = ()
    name: str = "room"
    metadata: Mapping[str, Any] = field(default_factory=dict)


def point_in_polygon(x: float, y: float, corners: np.ndarray) -> bool:
    """Crossing-number test of steam_rt.py:4258-4274."""
    inside = False
    j = corners.shape[0] - 1
    for i in range(corners.shape[0]):
        xi, yi = corners[i]
        xj, yj = corners[j]
        if (yi > y) != (yj > y):
            denom = yj - yi
            if abs(denom) > 1e-12:
                if x < (xj - xi) * (y - yi) / denom + xi:
                    inside = not inside
        j = i
    return inside


class RoomScene:
    """Surface list of one room in the reference order and layout."""

    def __init__(self, spec: RoomSpec) -> None:
        self.spec = spec
        corners = np.asarray(spec.corners, dtype=np.float64)
        if corners.ndim != 2 or corners.shape[1] != 2 or corners.shape[0] < 3:
            raise ValueError("corners must be an [P,2] polygon with P >= 3")
        self.corners = corners
        self.height = float(spec.height)
        floor = spec.floor_material or spec.wall_material
        ceiling = spec.ceiling_material or spec.wall_material
        surfaces: list[dict[str, Any]] = []
        walls = spec.walls or tuple(
            WallSpec(tuple(corners[i]), tuple(corners[(i + 1) % len(corners)]), name=f"wall_{i}")
            for i in range(len(corners))
        )
        for w in walls:
            a = np.asarray(w.a, dtype=np.float64)
            b = np.asarray(w.b, dtype=np.float64)
            z_min = max(0.0, float(w.z_min))
            z_max = min(self.height, float(self.height if w.z_max is None else w.z_max))
            if float(np.linalg.norm(b - a)) <= EPS or z_max - z_min <= EPS:
                continue
            seg = b - a
            n = float(np.hypot(seg[1], seg[0]))
            normal = np.asarray([seg[1] / n, -seg[0] / n, 0.0]) if n > EPS else np.asarray([1.0, 0.0, 0.0])
            surfaces.append({"kind": 0, "name": w.name, "a": a, "b": b, "z": (z_min, z_max), "normal": normal,
                             "material": w.material or spec.wall_material})
        surfaces.append({"kind": 1, "name": "floor", "z_value": 0.0, "normal": np.asarray([0.0, 0.0, 1.0]),
                         "material": floor})
        surfaces.append({"kind": 1, "name": "ceiling", "z_value": self.height, "normal": np.asarray([0.0, 0.0, -1.0]),
                         "material": ceiling})
        for box in spec.boxes:
            center = np.asarray(box.center, dtype=np.float64)
            if not point_in_polygon(float(center[0]), float(center[1]), corners):
                continue  # steam_rt.py:636-637
            size = np.maximum(np.asarray(box.size, dtype=np.float64), 1e-3)
            angle = math.radians(float(box.rotation_deg))
            axis_u = np.asarray([math.cos(angle), math.sin(angle)])
            axis_v = np.asarray([-math.sin(angle), math.cos(angle)])
            zc = float(size[2] * 0.5 if box.z_center is None else box.z_center)
            z_min = float(max(0.0, zc - size[2] * 0.5))
            z_max = float(max(z_min + 1e-3, zc + size[2] * 0.5))
            surfaces.append({"kind": 2, "name": box.name, "center": center, "axis_u": axis_u, "axis_v": axis_v,
                             "half": (float(size[0] * 0.5), float(size[1] * 0.5)), "z": (z_min, z_max),
                             "normal": np.asarray([0.0, 0.0, 1.0]), "material": box.material or spec.wall_material})
        self.surfaces = surfaces
        self._arrays: dict[str, Any] | None = None

    @property
    def num_surfaces(self) -> int:
        return len(self.surfaces)

    def surface_bounds(self) -> tuple[np.ndarray, np.ndarray]:
        """Per-surface AABBs padded by 2e-6 (steam_rt.py:234-262)."""
        lo = np.zeros((len(self.surfaces), 3))
        hi = np.zeros((len(self.surfaces), 3))
        for i, s in enumerate(self.surfaces):
            if s["kind"] == 0:
                lo[i] = (min(s["a"][0], s["b"][0]), min(s["a"][1], s["b"][1]), s["z"][0])
                hi[i] = (max(s["a"][0], s["b"][0]), max(s["a"][1], s["b"][1]), s["z"][1])
            elif s["kind"] == 2:
                ext = np.abs(s["axis_u"]) * s["half"][0] + np.abs(s["axis_v"]) * s["half"][1]
                lo[i] = (s["center"][0] - ext[0], s["center"][1] - ext[1], s["z"][0])
                hi[i] = (s["center"][0] + ext[0], s["center"][1] + ext[1], s["z"][1])
            else:
                lo[i] = (self.corners[:, 0].min(), self.corners[:, 1].min(), s["z_value"])
                hi[i] = (self.corners[:, 0].max(), self.corners[:, 1].max(), s["z_value"])
        return lo - BVH_BOUNDS_EPS, hi + BVH_BOUNDS_EPS

    def kernel_arrays(self) -> dict[str, Any]:
        """The reference's ``_scene_kernel_arrays`` dictionary for this room."""
        if self._arrays is not None:
            return self._arrays
        from .bvh import build_bvh

        S = len(self.surfaces)
        kinds = np.zeros(S, dtype=np.int64)
        wall_a = np.zeros((S, 2))
        wall_b = np.zeros((S, 2))
        wall_z = np.zeros((S, 2))
        wall_z[:, 1] = self.height
        z_values = np.zeros(S)
        box_center = np.zeros((S, 2))
        box_axis_u = np.zeros((S, 2))
        box_axis_v = np.zeros((S, 2))
        box_half = np.zeros((S, 2))
        box_z = np.zeros((S, 2))
        normals = np.zeros((S, 3))
        absorption = np.zeros((S, NUM_BANDS))
        scattering = np.zeros(S)
        names = []
        for i, s in enumerate(self.surfaces):
            names.append(s["name"])
            normals[i] = s["normal"]
            absorption[i] = s["material"].absorption
            scattering[i] = float(np.mean(np.atleast_1d(s["material"].scattering)))
            kinds[i] = s["kind"]
            if s["kind"] == 0:
                wall_a[i] = s["a"]
                wall_b[i] = s["b"]
                wall_z[i] = s["z"]
            elif s["kind"] == 2:
                box_center[i] = s["center"]
                box_axis_u[i] = s["axis_u"]
                box_axis_v[i] = s["axis_v"]
                box_half[i] = s["half"]
                box_z[i] = s["z"]
            else:
                z_values[i] = s["z_value"]
        lo, hi = self.surface_bounds()
        bvh = build_bvh(lo, hi)
        arrays = {
            "kinds": kinds,
            "wall_a": wall_a,
            "wall_delta": wall_b - wall_a,
            "wall_z": wall_z,
            "z_values": z_values,
            "box_center": box_center,
            "box_axis_u": box_axis_u,
            "box_axis_v": box_axis_v,
            "box_half": box_half,
            "box_z": box_z,
            "normals": normals,
            "absorption": absorption,
            "reflection": 1.0 - absorption,
            "scattering": scattering,
            "names": tuple(names),
            "corners": self.corners.copy(),
            "height": self.height,
            "surface_bounds_min": lo,
            "surface_bounds_max": hi,
            "bvh_bounds_min": bvh["bounds_min"],
            "bvh_bounds_max": bvh["bounds_max"],
            "bvh_start": bvh["start"],
            "bvh_count": bvh["count"],
            "bvh_escape": bvh["escape"],
            "bvh_primitives": bvh["primitives"],
            "bvh_node_count": int(bvh["node_count"]),
            "bvh_leaf_count": int(bvh["leaf_count"]),
            "bvh_max_depth": int(bvh["max_depth"]),
            "bvh_leaf_size": int(bvh["leaf_size"]),
        }
        for value in arrays.values():
            if isinstance(value, np.ndarray):
                value.setflags(write=False)
        self._arrays = arrays
        return arrays


@dataclass(frozen=True)
class TriangleMesh:
    vertices: np.ndarray  # [T,9] float64: v0, v1, v2
    surface: np.ndarray   # [T] int32 reference surface id
    normals: np.ndarray   # [T,3] float64 surface normal (unoriented)
    num_surfaces: int

    @property
    def num_triangles(self) -> int:
        return int(self.vertices.shape[0])

    def bounds(self, pad: float) -> tuple[np.ndarray, np.ndarray]:
        v = self.vertices.reshape(-1, 3, 3)
        return v.min(axis=1) - pad, v.max(axis=1) + pad


def _ear_clip(poly: np.ndarray) -> list[tuple[int, int, int]]:
    """Triangulate a simple polygon (either winding) by ear clipping."""
    n = poly.shape[0]
    area = 0.5 * float(np.sum(poly[:, 0] * np.roll(poly[:, 1], -1) - np.roll(poly[:, 0], -1) * poly[:, 1]))
    idx = list(range(n)) if area > 0 else list(range(n - 1, -1, -1))

    def cross(o, a, b):
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])

    def inside(p, a, b, c):
        return cross(a, b, p) >= -1e-12 and cross(b, c, p) >= -1e-12 and cross(c, a, p) >= -1e-12

    tris: list[tuple[int, int, int]] = []
    guard = 0
    while len(idx) > 3 and guard < 10 * n * n:
        guard += 1
        m = len(idx)
        clipped = False
        for k in range(m):
            i0, i1, i2 = idx[(k - 1) % m], idx[k], idx[(k + 1) % m]
            a, b, c = poly[i0], poly[i1], poly[i2]
            if cross(a, b, c) <= 1e-12:
                continue  # reflex or degenerate
            if any(inside(poly[j], a, b, c) for j in idx if j not in (i0, i1, i2)):
                continue
            tris.append((i0, i1, i2))
            idx.pop(k)
            clipped = True
            break
        if not clipped:
            raise ValueError("polygon is not simple; cannot triangulate")
    tris.append((idx[0], idx[1], idx[2]))
    return tris


def _refine_all(tris: np.ndarray, max_edge: float | None) -> tuple[np.ndarray, np.ndarray]:
    """Midpoint-refine [n,3,3] triangles until every edge <= max_edge.

    Returns the refined triangles and, for each, the index of its parent.
    """
    parent = np.arange(tris.shape[0])
    if max_edge is None or tris.shape[0] == 0:
        return tris, parent
    e = np.stack([np.linalg.norm(tris[:, 1] - tris[:, 0], axis=1), np.linalg.norm(tris[:, 2] - tris[:, 1], axis=1),
                  np.linalg.norm(tris[:, 0] - tris[:, 2], axis=1)], axis=1).max(axis=1)
    levels = np.where(e <= max_edge, 0, np.ceil(np.log2(np.maximum(e, 1e-300) / max_edge))).astype(np.int64)
    out_t, out_p = [], []
    for level in np.unique(levels):
        sel = np.flatnonzero(levels == level)
        cur, cur_p = tris[sel], parent[sel]
        for _ in range(int(level)):
            a, b, c = cur[:, 0], cur[:, 1], cur[:, 2]
            ab, bc, ca = 0.5 * (a + b), 0.5 * (b + c), 0.5 * (c + a)
            cur = np.stack([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1), np.stack([ca, bc, c], 1),
                            np.stack([ab, bc, ca], 1)], axis=1).reshape(-1, 3, 3)
            cur_p = np.repeat(cur_p, 4)
        out_t.append(cur)
        out_p.append(cur_p)
    t = np.concatenate(out_t)
    p = np.concatenate(out_p)
    order = np.argsort(p, kind="stable")  # keep surfaces contiguous and in input order
    return t[order], p[order]


def tessellate(arrays: Mapping[str, Any], max_edge: float | None = None) -> TriangleMesh:
    """Triangles of the reference surface arrays (see module docstring)."""
    kinds = np.asarray(arrays["kinds"])
    S = int(kinds.shape[0])
    corners = np.asarray(arrays["corners"], dtype=np.float64)
    floor_tris = _ear_clip(corners)
    verts: list[np.ndarray] = []
    surf: list[int] = []
    norms: list[np.ndarray] = []

    def add(si: int, a, b, c, normal) -> None:
        verts.append(np.concatenate([np.asarray(a, float), np.asarray(b, float), np.asarray(c, float)]))
        surf.append(si)
        norms.append(np.asarray(normal, dtype=np.float64))

    for si in range(S):
        kind = int(kinds[si])
        normal = np.asarray(arrays["normals"][si], dtype=np.float64)
        if kind == 0:
            a = np.asarray(arrays["wall_a"][si], float)
            b = a + np.asarray(arrays["wall_delta"][si], float)
            z0, z1 = (float(x) for x in arrays["wall_z"][si])
            p0, p1, p2, p3 = (a[0], a[1], z0), (b[0], b[1], z0), (b[0], b[1], z1), (a[0], a[1], z1)
            add(si, p0, p1, p2, normal)
            add(si, p0, p2, p3, normal)
        elif kind == 1:
            z = float(arrays["z_values"][si])
            for i0, i1, i2 in floor_tris:
                add(si, (*corners[i0], z), (*corners[i1], z), (*corners[i2], z), normal)
        else:
            c = np.asarray(arrays["box_center"][si], float)
            u = np.asarray(arrays["box_axis_u"][si], float)
            v = np.asarray(arrays["box_axis_v"][si], float)
            hw, hd = (float(x) for x in arrays["box_half"][si])
            z0, z1 = (float(x) for x in arrays["box_z"][si])

            def corner(su: float, sv: float, z: float) -> tuple[float, float, float]:
                p = c + su * hw * u + sv * hd * v
                return (float(p[0]), float(p[1]), z)

            u3 = np.asarray([u[0], u[1], 0.0])
            v3 = np.asarray([v[0], v[1], 0.0])
            faces = (
                ((1, -1), (1, 1), u3),     # +u face
                ((-1, 1), (-1, -1), -u3),  # -u face
                ((1, 1), (-1, 1), v3),     # +v face
                ((-1, -1), (1, -1), -v3),  # -v face
            )
            for (s0u, s0v), (s1u, s1v), n in faces:
                q0, q1 = corner(s0u, s0v, z0), corner(s1u, s1v, z0)
                q2, q3 = corner(s1u, s1v, z1), corner(s0u, s0v, z1)
                add(si, q0, q1, q2, n)
                add(si, q0, q2, q3, n)
            b00, b10, b11, b01 = corner(-1, -1, z0), corner(1, -1, z0), corner(1, 1, z0), corner(-1, 1, z0)
            add(si, b00, b11, b10, (0.0, 0.0, -1.0))
            add(si, b00, b01, b11, (0.0, 0.0, -1.0))
            t00, t10, t11, t01 = corner(-1, -1, z1), corner(1, -1, z1), corner(1, 1, z1), corner(-1, 1, z1)
            add(si, t00, t10, t11, (0.0, 0.0, 1.0))
            add(si, t00, t11, t01, (0.0, 0.0, 1.0))
    coarse = np.asarray(verts, dtype=np.float64).reshape(-1, 3, 3)
    refined, parent = _refine_all(coarse, max_edge)
    return TriangleMesh(
        vertices=refined.reshape(-1, 9),
        surface=np.asarray(surf, dtype=np.int32)[parent],
        normals=np.asarray(norms, dtype=np.float64).reshape(-1, 3)[parent],
        num_surfaces=S,
    )


def _refined_count(longest: np.ndarray, max_edge: float) -> int:
    levels = np.where(longest <= max_edge, 0, np.ceil(np.log2(np.maximum(longest, 1e-300) / max_edge)))
    return int(np.sum(4.0 ** levels))


def tessellate_to_count(arrays: Mapping[str, Any], target: int) -> tuple[TriangleMesh, float | None]:
    """Tessellate with the refinement edge whose triangle count is nearest ``target``."""
    coarse = tessellate(arrays, None)
    if coarse.num_triangles >= target:
        return coarse, None
    v = coarse.vertices.reshape(-1, 3, 3)
    longest = np.stack([np.linalg.norm(v[:, 1] - v[:, 0], axis=1), np.linalg.norm(v[:, 2] - v[:, 1], axis=1),
                        np.linalg.norm(v[:, 0] - v[:, 2], axis=1)], axis=1).max(axis=1)
    lo, hi = 1e-3, float(longest.max()) * 1.01
    best_edge, best_err = None, abs(coarse.num_triangles - target)
    for _ in range(60):
        mid = math.sqrt(lo * hi)
        n = _refined_count(longest, mid)
        if abs(n - target) < best_err:
            best_edge, best_err = mid, abs(n - target)
        if n > target:
            lo = mid
        else:
            hi = mid
        if hi / lo < 1.0005:
            break
    if best_edge is None:
        return coarse, None
    return tessellate(arrays, best_edge), best_edge
