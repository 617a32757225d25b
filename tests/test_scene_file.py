"""Scene files (vkr_scene_save / vkr_scene_load, host only): the bytes after
the 64-byte header are the std430 arrays exactly as uploaded, a file
round-trips bit for bit, and truncated or foreign files are rejected."""
import struct

import numpy as np
import pytest

import scenes
import vkr


def _same(a, b):
    return a.dtype == b.dtype and a.shape == b.shape and a.tobytes() == b.tobytes()


def test_round_trip_bit_exact(tmp_path):
    wl = scenes.soup100k(seed=3, n_tris=500, width=64, height=48, n_spheres=16)
    sc = wl.scene
    mats = np.zeros(3, vkr.MATERIAL)
    mats[1]["albedo"] = (0.25, 0.5, np.float32(np.nan))   # NaN payloads survive too
    sc.materials = mats
    sc.flags = vkr.SCENE_TRI_MATERIAL_INDEX
    p = str(tmp_path / "s.vkrs")
    sc.save(p)
    got = vkr.Scene.load(p)
    assert _same(got.tris, np.ascontiguousarray(sc.tris, vkr.TRIANGLE))
    assert _same(got.spheres, np.ascontiguousarray(sc.spheres, vkr.SPHERE))
    assert _same(got.planes, np.ascontiguousarray(sc.planes, vkr.PLANE))
    assert _same(got.materials, mats)
    assert np.asarray(got.tri_material, vkr.MATERIAL).tobytes() == np.asarray(sc.tri_material, vkr.MATERIAL).tobytes()
    assert got.flags == sc.flags


def test_layout_is_header_then_std430_arrays(tmp_path):
    wl = scenes.reference_scene(8, 8)
    p = tmp_path / "ref.vkrs"
    wl.scene.save(str(p), max_leaf=3, max_depth=40, traversal_cost=2.5)
    raw = p.read_bytes()
    magic, version, flags, nt, ns, npl, nm, htm, leaf, depth, cost = struct.unpack_from("<8s9If", raw, 0)
    assert magic == b"VKRSCENE" and version == 1 and flags == 0
    assert (nt, ns, npl, nm, htm) == (1, 4, 5, 0, 1)
    assert (leaf, depth, cost) == (3, 40, 2.5)
    assert raw[48:64] == bytes(16)
    body = raw[64:]
    tm = np.asarray(wl.scene.tri_material, vkr.MATERIAL).reshape(1).tobytes()
    expect = (tm + np.ascontiguousarray(wl.scene.tris, vkr.TRIANGLE).tobytes() +
              np.ascontiguousarray(wl.scene.spheres, vkr.SPHERE).tobytes() +
              np.ascontiguousarray(wl.scene.planes, vkr.PLANE).tobytes())
    assert body == expect
    # The reference scene's sphere radii sit at offset 60 of each 64-byte record (std430).
    sph = body[48 + 48:48 + 48 + 4 * 64]
    radii = [struct.unpack_from("<f", sph, 64 * i + 60)[0] for i in range(4)]
    assert radii == [16.0, 12.0, 24.0, 11.0]


def test_bad_files_rejected(tmp_path):
    wl = scenes.reference_scene(8, 8)
    p = tmp_path / "ok.vkrs"
    wl.scene.save(str(p))
    raw = p.read_bytes()
    cases = {"truncated": raw[:-1], "trailing": raw + b"\0", "magic": b"X" + raw[1:],
             "version": raw[:8] + struct.pack("<I", 2) + raw[12:], "empty": b""}
    for name, data in cases.items():
        q = tmp_path / f"{name}.vkrs"
        q.write_bytes(data)
        with pytest.raises(vkr.VkrError):
            vkr.Scene.load(str(q))
    with pytest.raises(vkr.VkrError):
        vkr.Scene.load(str(tmp_path / "missing.vkrs"))
