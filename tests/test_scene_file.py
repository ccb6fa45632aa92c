"""Scene files and the scene content hash (host logic, no GPU)."""
from __future__ import annotations

import numpy as np
import pytest

from acoustic_agent_amd import scenes
from acoustic_agent_amd.scene import TriangleMesh
from acoustic_agent_amd.scene_file import MAGIC, content_hash, load_scene, save_scene


def _fnv1a(chunks) -> int:
    h = 1469598103934665603
    for b in chunks:
        for byte in bytes(b):
            h ^= byte
            h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _tiny():
    v = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 0, 1, 1, 0, 1, 0, 1, 1]], dtype=np.float32)
    s = np.array([0, 1], dtype=np.int32)
    n = np.array([[0, 0, 1], [0, 0, -1]], dtype=np.float32)
    r = np.linspace(0.1, 0.9, 12, dtype=np.float32).reshape(2, 6)
    c = np.array([0.05, 0.3], dtype=np.float32)
    return v, s, n, r, c


def test_scene_hash_is_fnv1a_of_the_inputs():
    v, s, n, r, c = _tiny()
    want = _fnv1a([v.tobytes(), s.tobytes(), n.tobytes(), r.tobytes(), c.tobytes(), np.int32(4).tobytes()])
    assert content_hash(v, s, n, r, c, 4) == want
    assert content_hash(v, s, n, r, c, 2) != want
    r2 = r.copy()
    r2[1, 5] = np.nextafter(r2[1, 5], 1.0)
    assert content_hash(v, s, n, r2, c, 4) != want


def test_scene_file_round_trip(tmp_path):
    w = scenes.l_plan_workload(0, target_triangles=3000)
    a = w.scene.kernel_arrays()
    names = tuple(f"surface_{i}" for i in range(w.mesh.num_surfaces))
    path = tmp_path / "plan.aascene"
    h = save_scene(path, w.mesh, a["reflection"], a["scattering"], surface_names=names)
    f = load_scene(path)
    assert f.content_hash == h == content_hash(w.mesh.vertices, w.mesh.surface, w.mesh.normals, a["reflection"],
                                               a["scattering"], 4)
    np.testing.assert_array_equal(f.mesh.vertices, w.mesh.vertices.astype(np.float32))
    np.testing.assert_array_equal(f.mesh.surface, w.mesh.surface)
    np.testing.assert_array_equal(f.mesh.normals, w.mesh.normals.astype(np.float32))
    np.testing.assert_array_equal(f.reflection, np.asarray(a["reflection"], dtype=np.float32))
    np.testing.assert_array_equal(f.scattering, np.asarray(a["scattering"], dtype=np.float32))
    assert f.surface_names == names and f.leaf_size == 4 and f.mesh.num_surfaces == w.mesh.num_surfaces
    assert path.stat().st_size == 40 + w.mesh.num_triangles * 52 + w.mesh.num_surfaces * 28 + len("\n".join(names))


def test_scene_file_rejects_damage(tmp_path):
    v, s, n, r, c = _tiny()
    mesh = TriangleMesh(vertices=v.astype(np.float64), surface=s, normals=n.astype(np.float64), num_surfaces=2)
    path = tmp_path / "tiny.aascene"
    save_scene(path, mesh, r, c)
    good = path.read_bytes()
    assert good.startswith(MAGIC) and load_scene(path).mesh.num_triangles == 2
    for bad in (good[:-3], b"NOTASCN1" + good[8:], good[:8] + b"\x02" + good[9:],
                good[:60] + bytes([good[60] ^ 1]) + good[61:]):
        path.write_bytes(bad)
        with pytest.raises(ValueError):
            load_scene(path)
    with pytest.raises(ValueError):
        save_scene(path, mesh, r[:1], c[:1])
    with pytest.raises(ValueError):  # names are newline-separated in the file
        save_scene(path, mesh, r, c, surface_names=("wall", "floor\ntile"))
