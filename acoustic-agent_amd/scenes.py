"""Seeded synthetic scenes of the benchmark workloads.

Materials: per-surface six-band absorption ~ U[0.05, 0.4] and scalar
scattering ~ U[0.05, 0.3] from ``numpy.random.default_rng(seed)``.

* ``shoebox``: a rectangular room (4 walls, floor, ceiling; 12 triangles).
* ``l_plan``: a 12 x 9 x 2.8 m L-shaped outline split into three rooms by
  interior walls with 0.8 m door gaps and lintels from 2.1 m to the ceiling
  (the door pattern of acoustic_agent/floorplan.py:606-624), furnished with
  seeded oriented boxes (sides U[0.3, 2] m, rotation U[0, 90) deg) placed
  inside a room, clear of the door zones and of the source and listener.
* ``random_scene``: shoeboxes for tiny triangle budgets, furnished L plans
  refined to a target triangle count otherwise.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .mic import circular_array
from .scene import BoxSpec, Material, RoomScene, RoomSpec, TriangleMesh, WallSpec, point_in_polygon, tessellate, \
    tessellate_to_count


def random_material(rng: np.random.Generator) -> Material:
    absorption = tuple(float(x) for x in rng.uniform(0.05, 0.4, 6))
    return Material(absorption=absorption, scattering=float(rng.uniform(0.05, 0.3)))


@dataclass
class Workload:
    """A scene with its source / listener placement."""

    scene: RoomScene
    mesh: TriangleMesh
    source: np.ndarray
    listener: np.ndarray
    seed: int
    max_edge: float | None


def shoebox(size=(5.0, 4.0, 3.0), seed: int = 0) -> RoomScene:
    rng = np.random.default_rng(seed)
    lx, ly, h = (float(v) for v in size)
    corners = ((0.0, 0.0), (lx, 0.0), (lx, ly), (0.0, ly))
    walls = tuple(WallSpec(corners[i], corners[(i + 1) % 4], name=f"wall_{i}", material=random_material(rng))
                  for i in range(4))
    spec = RoomSpec(corners=corners, height=h, wall_material=walls[0].material, walls=walls,
                    floor_material=random_material(rng), ceiling_material=random_material(rng), name="shoebox")
    return RoomScene(spec)


L_OUTLINE = ((0.0, 0.0), (12.0, 0.0), (12.0, 5.0), (7.0, 5.0), (7.0, 9.0), (0.0, 9.0))
# rooms as axis-aligned rectangles (x0, y0, x1, y1)
L_ROOMS = ((0.0, 0.0, 7.0, 5.0), (0.0, 5.0, 7.0, 9.0), (7.0, 0.0, 12.0, 5.0))
# doors: (wall axis, fixed coordinate, lo, hi)
L_DOORS = (("y", 5.0, 3.0, 3.8), ("x", 7.0, 2.0, 2.8))
L_SOURCE = (2.0, 2.0, 1.5)
L_LISTENER = (10.0, 2.5, 1.4)


def _rect_corners(cx: float, cy: float, w: float, d: float, angle: float) -> np.ndarray:
    u = np.asarray([math.cos(angle), math.sin(angle)])
    v = np.asarray([-math.sin(angle), math.cos(angle)])
    c = np.asarray([cx, cy])
    return np.asarray([c + su * w * 0.5 * u + sv * d * 0.5 * v for su, sv in ((-1, -1), (1, -1), (1, 1), (-1, 1))])


def l_plan(seed: int = 0, num_boxes: int = 60, height: float = 2.8, door_height: float = 2.1,
           source=L_SOURCE, listener=L_LISTENER, max_tries: int = 20000, keep_out=()) -> RoomScene:
    rng = np.random.default_rng(seed)
    walls: list[WallSpec] = []
    for i in range(len(L_OUTLINE)):
        a, b = L_OUTLINE[i], L_OUTLINE[(i + 1) % len(L_OUTLINE)]
        walls.append(WallSpec(a, b, name=f"wall_{i}", material=random_material(rng)))
    for di, (axis, fixed, lo, hi) in enumerate(L_DOORS):
        span = (0.0, 7.0) if axis == "y" else (0.0, 5.0)

        def pt(s: float) -> tuple[float, float]:
            return (s, fixed) if axis == "y" else (fixed, s)

        m = random_material(rng)
        walls.append(WallSpec(pt(span[0]), pt(lo), name=f"interior_{di}_a", material=m))
        walls.append(WallSpec(pt(hi), pt(span[1]), name=f"interior_{di}_b", material=m))
        walls.append(WallSpec(pt(lo), pt(hi), z_min=door_height, z_max=height, name=f"interior_{di}_lintel",
                              material=m))
    outline = np.asarray(L_OUTLINE)
    keep_out = [np.asarray(source[:2]), np.asarray(listener[:2])] + [np.asarray(p[:2]) for p in keep_out]
    door_zones = []
    for axis, fixed, lo, hi in L_DOORS:
        mid = 0.5 * (lo + hi)
        door_zones.append(np.asarray((mid, fixed) if axis == "y" else (fixed, mid)))
    boxes: list[BoxSpec] = []
    tries = 0
    while len(boxes) < num_boxes and tries < max_tries:
        tries += 1
        room = L_ROOMS[int(rng.integers(0, len(L_ROOMS)))]
        w, d, h = (float(x) for x in rng.uniform(0.3, 2.0, 3))
        h = min(h, height - 0.3)
        angle = math.radians(float(rng.uniform(0.0, 90.0)))
        cx = float(rng.uniform(room[0], room[2]))
        cy = float(rng.uniform(room[1], room[3]))
        rect = _rect_corners(cx, cy, w, d, angle)
        margin = 0.05
        if (rect[:, 0].min() < room[0] + margin or rect[:, 0].max() > room[2] - margin or
                rect[:, 1].min() < room[1] + margin or rect[:, 1].max() > room[3] - margin):
            continue
        radius = 0.5 * math.hypot(w, d)
        if any(np.linalg.norm(np.asarray([cx, cy]) - p) < radius + 0.6 for p in keep_out):
            continue
        if any(np.linalg.norm(np.asarray([cx, cy]) - p) < radius + 1.0 for p in door_zones):
            continue
        if not point_in_polygon(cx, cy, outline):
            continue
        boxes.append(BoxSpec(center=(cx, cy), size=(w, d, h), rotation_deg=math.degrees(angle), z_center=0.5 * h,
                             name=f"box_{len(boxes)}", material=random_material(rng)))
    spec = RoomSpec(corners=L_OUTLINE, height=height, wall_material=walls[0].material, walls=tuple(walls),
                    floor_material=random_material(rng), ceiling_material=random_material(rng), boxes=tuple(boxes),
                    name="l_plan")
    return RoomScene(spec)


def shoebox_workload(seed: int = 0) -> Workload:
    """5 x 4 x 3 m shoebox, source (1.2, 1.1, 1.5), listener (3.6, 2.7, 1.4)."""
    scene = shoebox((5.0, 4.0, 3.0), seed)
    return Workload(scene, tessellate(scene.kernel_arrays()), np.asarray((1.2, 1.1, 1.5)),
                    np.asarray((3.6, 2.7, 1.4)), seed, None)


def l_plan_workload(seed: int = 0, target_triangles: int = 20000, num_boxes: int = 60) -> Workload:
    """Furnished three-room L plan tessellated to about ``target_triangles``."""
    scene = l_plan(seed, num_boxes=num_boxes)
    mesh, edge = tessellate_to_count(scene.kernel_arrays(), target_triangles)
    return Workload(scene, mesh, np.asarray(L_SOURCE), np.asarray(L_LISTENER), seed, edge)


def random_scene(seed: int, target_triangles: int) -> Workload:
    """Shoebox (target < 100 triangles) or furnished L plan near the target."""
    rng = np.random.default_rng(seed)
    if target_triangles < 100:
        size = (float(rng.uniform(3.0, 10.0)), float(rng.uniform(3.0, 8.0)), float(rng.uniform(2.4, 4.0)))
        scene = shoebox(size, seed)
        src = np.asarray((size[0] * 0.3, size[1] * 0.35, 1.5))
        lis = np.asarray((size[0] * 0.7, size[1] * 0.6, 1.4))
        return Workload(scene, tessellate(scene.kernel_arrays()), src, lis, seed, None)
    num_boxes = int(min(60, max(0, (target_triangles - 40) // 12)))
    scene = l_plan(seed, num_boxes=num_boxes)
    mesh, edge = tessellate_to_count(scene.kernel_arrays(), target_triangles)
    return Workload(scene, mesh, np.asarray(L_SOURCE), np.asarray(L_LISTENER), seed, edge)


def catmull_rom(points: np.ndarray, samples: int) -> np.ndarray:
    """Uniform Catmull-Rom spline through ``points`` (end points repeated)."""
    p = np.asarray(points, dtype=np.float64)
    p = np.concatenate([p[:1], p, p[-1:]])
    segs = p.shape[0] - 3
    out = []
    for k in range(samples):
        u = k / max(samples - 1, 1) * segs
        i = min(int(u), segs - 1)
        t = u - i
        p0, p1, p2, p3 = p[i], p[i + 1], p[i + 2], p[i + 3]
        t2, t3 = t * t, t * t * t
        out.append(0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t2 + (-p0 + 3 * p1 - 3 * p2 + p3) * t3))
    return np.asarray(out)


def moving_source_path(seed: int, keyframes: int = 64, fps: float = 32.0, start=L_SOURCE,
                       bounds=(0.6, 0.6, 6.4, 4.4)) -> np.ndarray:
    """Seeded source path: a Catmull-Rom spline through random control points
    inside ``bounds`` (room 1 of the L plan), resampled to ``keyframes``
    positions ``1/fps`` apart at a constant speed drawn from U[0.5, 1.5] m/s."""
    rng = np.random.default_rng(seed)
    speed = float(rng.uniform(0.5, 1.5))
    length = speed * (keyframes - 1) / fps
    pts = [np.asarray(start, dtype=np.float64)]
    while True:
        nxt = pts[-1] + np.concatenate([rng.normal(0.0, 1.0, 2), [0.0]])
        nxt[0] = min(max(nxt[0], bounds[0]), bounds[2])
        nxt[1] = min(max(nxt[1], bounds[1]), bounds[3])
        pts.append(nxt)
        dense = catmull_rom(np.asarray(pts), 50 * len(pts))
        arc = np.concatenate([[0.0], np.cumsum(np.linalg.norm(np.diff(dense, axis=0), axis=1))])
        if arc[-1] >= length or len(pts) > 64:
            break
    keep = np.concatenate([[True], np.diff(arc) > 1e-12])  # np.interp needs strictly increasing abscissae
    arc, dense = arc[keep], dense[keep]
    target = np.linspace(0.0, min(length, arc[-1]), keyframes)
    out = np.stack([np.interp(target, arc, dense[:, a]) for a in range(3)], axis=1)
    out[:, 2] = start[2]
    return out


def mic_array_workload(seed: int, target_triangles: int, num_mics: int = 4, radius: float = 0.05) -> tuple[Workload, np.ndarray]:
    """A scene of the batched mix with its microphone-ring channel positions."""
    w = random_scene(seed, target_triangles)
    return w, circular_array(w.listener, num_mics, radius)


def keyframe_workload(seed: int, target_triangles: int = 20000, keyframes: int = 64) -> tuple[Workload, np.ndarray]:
    """A furnished L plan whose furniture keeps clear of a moving source path."""
    path = moving_source_path(seed, keyframes)
    scene = l_plan(seed, keep_out=[tuple(p) for p in path[:: max(1, keyframes // 16)]] + [tuple(path[-1])])
    mesh, edge = tessellate_to_count(scene.kernel_arrays(), target_triangles)
    return Workload(scene, mesh, path[0], np.asarray(L_LISTENER), seed, edge), path
