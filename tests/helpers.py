"""Shared inputs for the parity tests: one workload, one ray set, run through
both the HIP tracer (C ABI via hip_backend) and the oracle."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from acoustic_agent_amd import sampling, scenes
from oracle import oracle as O

C = 343.0


@dataclass
class Case:
    workload: scenes.Workload
    num_rays: int
    num_bounces: int
    num_bins: int
    bin_dur: float = 1e-3
    seed: int = 0
    rng: str = "reference"
    ambisonics: bool = False
    log_bounces: int = 0
    sources: np.ndarray | None = None
    extra: dict = field(default_factory=dict)

    @property
    def arrays(self):
        return self.workload.scene.kernel_arrays()

    @property
    def duration(self) -> float:
        return self.num_bins * self.bin_dur

    def ray_set(self):
        bank = sampling.diffuse_sample_bank(128)
        if self.rng == "reference":
            dirs = sampling.sphere_samples(self.num_rays, self.seed)
            U, J = sampling.diffuse_random_sequence(self.num_rays, self.num_bounces, 128, self.seed)
            return dirs, bank, U, J
        return None, bank, None, None

    def direct_delay(self, source=None) -> float:
        s = self.workload.source if source is None else source
        return float(np.linalg.norm(np.asarray(s) - self.workload.listener)) / C


def run_gpu(case: Case, *, use_bvh: bool = True, tuning=None, log_rays: int = 0, scene=None, source=None,
            plain: bool = False, visual_limit: int = 0, visual_stride: int = 1, ray_times: bool = False):
    from acoustic_agent_amd import hip_backend

    dirs, bank, U, J = case.ray_set()
    w = case.workload
    src = w.source if source is None else np.asarray(source)
    kw = dict(source=src, direct_delay=case.direct_delay(src))
    if case.sources is not None and source is None:
        kw = dict(source=None, direct_delay=None, sources=case.sources)
    return hip_backend.trace_energy_field_hip(
        listener=w.listener, directions=dirs, diffuse_bank=bank, diffuse_random=U, diffuse_indices=J,
        arrays=case.arrays, use_bvh=use_bvh, num_bounces=case.num_bounces, num_bins=case.num_bins,
        bin_dur=case.bin_dur, speed_of_sound=C, max_path_len=case.duration * C, listener_radius=0.1,
        source_radius=0.1, irradiance_min_distance=1.0, specular_exponent=100.0,
        source_forward_vector=np.asarray(case.extra.get("source_forward", (1.0, 0.0, 0.0))),
        dipole_weight=float(case.extra.get("dipole_weight", 0.0)),
        dipole_power=float(case.extra.get("dipole_power", 1.0)),
        render_ambisonics=case.ambisonics, device_id=0, rng_mode=case.rng, seed=case.seed,
        num_rays=case.num_rays, scene=scene, tessellation_max_edge=w.max_edge, log_rays=log_rays,
        log_bounces=case.log_bounces if log_rays else 0, tuning=tuning, plain=plain,
        visual_candidate_limit=visual_limit, visual_stride=visual_stride, ray_times=ray_times, **kw)


def run_oracle(case: Case, precision: str = "float32", *, triangles: bool = True, log_rays: int = 0,
               source=None, threads: int = 0, use_bvh: bool = True, ray_range=None, visual_limit: int = 0,
               visual_stride: int = 1):
    dirs, bank, U, J = case.ray_set()
    w = case.workload
    src = w.source if source is None else np.asarray(source)
    scene = O.OracleScene(case.arrays, w.mesh if triangles else None, precision, use_bvh=use_bvh)
    return O.trace(scene, source=src, listener=w.listener, num_rays=case.num_rays, num_bounces=case.num_bounces,
                   num_bins=case.num_bins, bin_dur=case.bin_dur, c=C, max_path_len=case.duration * C,
                   direct_delay=case.direct_delay(src), diffuse_bank=bank, rng=case.rng, directions=dirs,
                   source_forward=case.extra.get("source_forward", (1.0, 0.0, 0.0)),
                   dipole_weight=float(case.extra.get("dipole_weight", 0.0)),
                   dipole_power=float(case.extra.get("dipole_power", 1.0)),
                   diffuse_random=U, diffuse_indices=J, seed=case.seed, render_ambisonics=case.ambisonics,
                   log_rays=log_rays, log_bounces=case.log_bounces if log_rays else 0, threads=threads,
                   ray_range=ray_range, visual_limit=visual_limit, visual_stride=visual_stride)


def rel_l1(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def band_rel(a, b) -> np.ndarray:
    """Per-band relative difference of the total energy."""
    a = np.asarray(a, dtype=np.float64).sum(axis=-1)
    b = np.asarray(b, dtype=np.float64).sum(axis=-1)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


# ---------------------------------------------------------------------------
# hit logs: where two traces of the same ray set part ways
# ---------------------------------------------------------------------------

# Grazing tolerance.  A binary32 position of a room-sized scene (|x| < 16 m)
# has an ulp of at most 9.5e-7 m, so two traces of one ray that agree up to
# rounding can differ by about a micrometre per bounce; a binary32 trace and a
# binary64 trace of a long path also drift apart (each specular reflection off
# a rotated box face turns the rounding of the direction into a position
# offset that grows with the path).  A divergence at bounce j is grazing if
# the hit point lies within GRAZING_M * (j + 1) + 2 * separation of an edge of
# its triangle or of the other trace's triangle, where the separation is the
# measured distance between the two traces' rays at that bounce.
GRAZING_M = 1e-6


def _point_segment(p, a, b) -> float:
    ab = b - a
    s = float(np.clip(np.dot(p - a, ab) / max(np.dot(ab, ab), 1e-300), 0.0, 1.0))
    return float(np.linalg.norm(p - (a + s * ab)))


def _point_line(p, a, b) -> float:
    d = b - a
    n = float(np.linalg.norm(d))
    return float(np.linalg.norm(np.cross(p - a, d)) / n) if n > 0 else float(np.linalg.norm(p - a))


def edge_distance(p, tri) -> float:
    """Distance from ``p`` to the nearest edge of triangle ``tri`` [3,3]."""
    return min(_point_segment(p, tri[i], tri[(i + 1) % 3]) for i in range(3))


def triangle_distance(p, tri) -> float:
    """Distance from ``p`` to triangle ``tri`` [3,3] (interior or boundary)."""
    a, b, c = tri
    n = np.cross(b - a, c - a)
    nn = float(np.dot(n, n))
    if nn > 0.0:
        q = p - np.dot(p - a, n) / nn * n  # projection onto the plane
        w = [np.dot(np.cross(tri[(i + 1) % 3] - tri[i], q - tri[i]), n) for i in range(3)]
        if all(x >= 0 for x in w) or all(x <= 0 for x in w):
            return float(np.linalg.norm(p - q))
    return edge_distance(p, tri)


def _plane_crossing(origin, through, tri):
    """Where the line from ``origin`` through ``through`` meets the plane of
    ``tri`` (None if parallel)."""
    d = through - origin
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    den = float(np.dot(n, d))
    if abs(den) <= 1e-300 * max(float(np.dot(n, n)), 1.0):
        return None
    return origin + float(np.dot(n, tri[0] - origin)) / den * d


def _origins(tri, points, normals, listener, j):
    """Origin of segment j of one logged ray: the listener, or the hit point
    of j - 1 moved 1 cm along its triangle's normal, turned against the
    arriving segment (steam_rt.py:4795 offset, 4395-4403 orientation)."""
    if j == 0:
        return listener
    p = points[j - 1]
    before = points[j - 2] if j > 1 else listener
    n = np.asarray(normals[tri[j - 1]], dtype=np.float64)
    if np.dot(n, p - before) > 0:
        n = -n
    return p + 0.01 * n


def _unit(d):
    return d / max(float(np.linalg.norm(d)), 1e-300)


def hit_log_divergences(gpu_tri, gpu_point, ref_prim, ref_point, mesh, listener, tol_m: float = GRAZING_M,
                        on_triangle_m: float = 1e-4, max_angle: float = 1e-2, min_segment_m: float = 1e-3) -> dict:
    """Compare two hit logs [rays, bounces] of triangle indices (-1: no hit,
    -2: bounce not reached) with their hit points [rays, bounces, 3] (a miss
    logs origin + 1e30 direction).  For every ray whose sequences differ, take
    the first differing bounce j, the origins of segment j in both traces
    (the previous hit point moved 1 cm along its oriented normal) and their
    distance sep.  Then:
      termination  one trace ended the ray at bounce j (an end test on the
                   previous, identical segment decided differently);
      otherwise the divergence must be explained by the geometry: every
      logged hit point lies on its triangle (within on_triangle_m), the two
      segments point the same way (within max_angle, checked when both are
      longer than min_segment_m), and one of
        * a trace that hit triangle T at p did so within tol = tol_m * (j + 1)
          + 2 * s of an edge of T or of the other trace's triangle, s being
          the larger of sep and the distance between p and the other trace's
          ray in T's plane;
        * the common hit at j - 1 was within tol_m * j + 2 * sep of an edge
          of its triangle (the segment starts at a grazing hit);
        * a trace's hit at j lies within 2 * sep of its origin (the other
          trace may start on the far side of that triangle);
      explained with sep <= tol_m * (j + 1) it is grazing (the traces agree up
      to rounding), with a larger sep it is drift (binary32 against binary64
      over a long path: the traces had moved apart first);
      other        anything else: a real disagreement.
    Returns the counts, the first-divergent-bounce histogram, the largest
    origin separation among drifted rays and the largest explained distance
    relative to its tolerance."""
    g = np.asarray(gpu_tri)
    r = np.asarray(ref_prim)
    gp = np.asarray(gpu_point, dtype=np.float64)
    rp = np.asarray(ref_point, dtype=np.float64)
    assert g.shape == r.shape and gp.shape[:2] == g.shape and rp.shape[:2] == r.shape, (g.shape, r.shape)
    v = np.asarray(mesh.vertices, dtype=np.float64).reshape(-1, 3, 3)
    normals = np.asarray(mesh.normals, dtype=np.float64)
    lis = np.asarray(listener, dtype=np.float64)
    diff = g != r
    rays = np.flatnonzero(diff.any(axis=1))
    first = diff[rays].argmax(axis=1) if rays.size else np.zeros(0, dtype=int)
    out = {"rays": int(g.shape[0]), "same": int(g.shape[0] - rays.size), "grazing": 0, "drift": 0,
           "termination": 0, "other": [], "first_bounce_hist": {}, "max_drift_m": 0.0, "worst_grazing_ratio": 0.0}
    for ray, j in zip(rays.tolist(), first.tolist()):
        out["first_bounce_hist"][j] = out["first_bounce_hist"].get(j, 0) + 1
        a, b = int(r[ray, j]), int(g[ray, j])
        if a == -2 or b == -2:
            out["termination"] += 1
            continue
        o_g = _origins(g[ray], gp[ray], normals, lis, j)
        o_r = _origins(r[ray], rp[ray], normals, lis, j)
        sep = float(np.linalg.norm(o_g - o_r))
        seg_g, seg_r = gp[ray, j] - o_g, rp[ray, j] - o_r
        if any(triangle_distance(p, v[t]) > on_triangle_m for p, t in ((rp[ray, j], a), (gp[ray, j], b)) if t >= 0):
            out["other"].append((ray, j, a, b, float("inf")))
            continue
        if min(np.linalg.norm(seg_g), np.linalg.norm(seg_r)) > min_segment_m and \
                float(np.dot(_unit(seg_g), _unit(seg_r))) < np.cos(max_angle):
            out["other"].append((ray, j, a, b, float("inf")))
            continue
        ratios = []
        if j > 0 and r[ray, j - 1] >= 0:  # the segment starts at a grazing hit
            ratios.append(edge_distance(rp[ray, j - 1], v[r[ray, j - 1]]) / (tol_m * j + 2.0 * sep))
        for p, own, other, o_own, o_other, hit_other in ((rp[ray, j], a, b, o_r, o_g, gp[ray, j]),
                                                        (gp[ray, j], b, a, o_g, o_r, rp[ray, j])):
            if own < 0:
                continue
            q = _plane_crossing(o_other, hit_other, v[own])
            s = max(sep, float(np.linalg.norm(q - p))) if q is not None else sep
            d = edge_distance(p, v[own])
            if other >= 0:
                d = min(d, triangle_distance(p, v[other]))
            ratios.append(d / (tol_m * (j + 1) + 2.0 * s))
            if sep > 0.0:  # a hit closer to its origin than the traces are apart
                ratios.append(float(np.linalg.norm(p - o_own)) / (2.0 * sep))
        best = min(ratios) if ratios else float("inf")
        if best <= 1.0:
            if sep <= tol_m * (j + 1):
                out["grazing"] += 1
            else:
                out["drift"] += 1
                out["max_drift_m"] = max(out["max_drift_m"], sep)
            out["worst_grazing_ratio"] = max(out["worst_grazing_ratio"], best)
        else:
            out["other"].append((ray, j, a, b, best))
    out["first_bounce_hist"] = dict(sorted(out["first_bounce_hist"].items()))
    return out


def oracle_for_request(req, source_index: int = 0, precision: str = "float32", log: bool = False,
                       threads: int = 0):
    """The oracle trace of one source of a batch ``TraceRequest`` (counter
    RNG, the request's own parameters), on the request's host scene."""
    w = req.scene.workload
    src = np.asarray(req.sources, dtype=np.float64).reshape(-1, 3)[source_index]
    scene = O.OracleScene(w.scene.kernel_arrays(), w.mesh, precision)
    n = req.num_rays
    return O.trace(scene, source=src, listener=req.listener, num_rays=n, num_bounces=req.num_bounces,
                   num_bins=req.num_bins, bin_dur=req.bin_dur, c=req.c, max_path_len=req.duration_s * req.c,
                   direct_delay=float(np.linalg.norm(src - np.asarray(req.listener, dtype=np.float64))) / req.c,
                   listener_radius=req.listener_radius, source_radius=req.source_radius,
                   irradiance_min_distance=req.irradiance_min_distance, specular_exponent=req.specular_exponent,
                   source_forward=req.source_forward, dipole_weight=req.dipole_weight,
                   dipole_power=req.dipole_power, diffuse_bank=sampling.diffuse_sample_bank(req.num_diffuse_samples),
                   rng="counter", seed=req.seed, threads=threads, log_rays=n if log else 0,
                   log_bounces=req.num_bounces if log else 0)
