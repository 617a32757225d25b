"""The same-BVH CPU baseline (oracle/bvh_accel.cpp) against the linear scan.

The accelerated oracle traces the product's BVH (vk-renderer_amd/csrc/bvh.cpp)
on the CPU and replays the reference's scan-order acceptance rule
(Assets/Tracer.comp:374-431) over the hits near the closest one; the shadow
rays use the any-hit form.  It must give exactly the linear scan's results,
or the CPU baseline it times would not be the same computation.
"""
import numpy as np
import pytest

import oracle_ffi
import scenes
import vkr

OUT_KEYS = ("rgba8", "radiance", "prim_id", "hit_t", "ray_count")


def _small_soup(seed=0, n_tris=3000, width=96, height=54):
    return scenes.soup100k(seed=seed, n_tris=n_tris, width=width, height=height, spp=2, max_depth=4,
                           n_spheres=48)


CASES = {
    "soup3000": lambda: _small_soup(),
    "many_emitters": lambda: scenes.many_emitters(),
    "far_sphere": lambda: scenes.far_sphere(res=64),
    "heightfield": lambda: scenes.heightfield(n=32, extent=64.0, n_spheres=64, width=80, height=60, spp=2, max_depth=4),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_accel_frame_equals_scan(name):
    w = CASES[name]()
    kw = dict(spp=w.spp, max_depth=w.max_depth)
    ref = oracle_ffi.trace(w.scene, w.frame, w.width, w.height, **kw)
    acc = oracle_ffi.trace(w.scene, w.frame, w.width, w.height, accel=True, **kw)
    for k in OUT_KEYS:
        assert np.array_equal(ref[k], acc[k]), f"{name}: {k} differs at {np.argwhere(ref[k] != acc[k])[:5]}"


def test_accel_spheres64_pixels_equal_scan():
    w = scenes.spheres64()
    px = np.random.default_rng(5).choice(w.width * w.height, 4096, replace=False).astype(np.uint32)
    ref = oracle_ffi.trace(w.scene, w.frame, w.width, w.height, spp=w.spp, max_depth=w.max_depth, pixels=px)
    acc = oracle_ffi.trace(w.scene, w.frame, w.width, w.height, spp=w.spp, max_depth=w.max_depth, pixels=px,
                           accel=True)
    for k in OUT_KEYS:
        assert np.array_equal(ref[k], acc[k]), k


def test_accel_trace_ray_equals_scan():
    """Single closest-hit queries: random origins and directions, random
    t_max (some below the nearest hit), and rays grazing coincident
    triangles, where the scan's `t < best + EPSILON` rule picks a later
    primitive than the nearest one."""
    lib = oracle_ffi.load()
    rng = np.random.default_rng(11)
    w = _small_soup(seed=3, n_tris=2000)
    # Near-coincident copies: the same triangles shifted by less than EPSILON.
    tris = w.scene.tris
    dup = tris[:300].copy()
    for k in ("v0", "v1", "v2"):
        dup[k] = dup[k] + np.float32(4e-4)
    scene = vkr.Scene(tris=np.concatenate([tris, dup]), spheres=w.scene.spheres, planes=w.scene.planes)
    plain = oracle_ffi.OracleScene(scene)
    fast = oracle_ffi.OracleScene(scene, accel=True)
    P = oracle_ffi.C.POINTER
    fp = P(oracle_ffi.C.c_float)
    n_diff, n_hit = 0, 0
    for i in range(6000):
        if i % 3 == 0:   # aim at a duplicated triangle's centroid
            t = tris[rng.integers(300)]
            c = (t["v0"] + t["v1"] + t["v2"]) / np.float32(3)
            o = (rng.uniform(-80, 80, 3)).astype(np.float32)
            d = (c - o).astype(np.float32)
        else:
            o = rng.uniform(-60, 60, 3).astype(np.float32)
            d = rng.normal(size=3).astype(np.float32)
        d = (d / np.float32(np.sqrt(np.float32(np.dot(d, d))))).astype(np.float32)
        tmax = np.float32(rng.choice([1e30, rng.uniform(1.0, 120.0)]))
        res = []
        for sc in (plain, fast):
            prim = oracle_ffi.C.c_uint32(0)
            tout = oracle_ffi.C.c_float(0)
            h = lib.vkro_trace_ray(oracle_ffi.C.byref(sc.c), o.ctypes.data_as(fp), d.ctypes.data_as(fp),
                                   float(tmax), 1e-3, oracle_ffi.C.byref(prim), oracle_ffi.C.byref(tout))
            res.append((h, prim.value, np.float32(tout.value).view(np.uint32)))
        n_hit += res[0][0]
        n_diff += res[0] != res[1]
    assert n_hit > 1000
    assert n_diff == 0


def test_accel_nan_ray_misses_like_scan():
    lib = oracle_ffi.load()
    w = _small_soup(seed=1, n_tris=200)
    fp = oracle_ffi.C.POINTER(oracle_ffi.C.c_float)
    o = np.array([np.nan, 0.0, 0.0], np.float32)
    d = np.array([0.0, 0.0, 1.0], np.float32)
    out = []
    for accel in (False, True):
        sc = oracle_ffi.OracleScene(w.scene, accel=accel)
        prim = oracle_ffi.C.c_uint32(0)
        tout = oracle_ffi.C.c_float(0)
        h = lib.vkro_trace_ray(oracle_ffi.C.byref(sc.c), o.ctypes.data_as(fp), d.ctypes.data_as(fp), 1e30, 1e-3,
                               oracle_ffi.C.byref(prim), oracle_ffi.C.byref(tout))
        out.append((h, prim.value))
    assert out[0] == out[1]


def test_accel_many_coincident_candidates():
    """400 copies of one triangle, all hit at the same t: the scan's
    `t < best + EPSILON` makes the last copy win, and every copy is a
    candidate (more than any fixed-size list holds)."""
    lib = oracle_ffi.load()
    tri = scenes.tris_from(np.array([[-1, -1, 0]], np.float32), np.array([[1, -1, 0]], np.float32),
                           np.array([[0, 1, 0]], np.float32))
    rng = np.random.default_rng(2)
    other = scenes.random_soup(rng, 200, lo=30.0, hi=50.0)   # off the rays' path
    scene = vkr.Scene(tris=np.concatenate([other[:100], np.repeat(tri, 400), other[100:]]))
    fp = oracle_ffi.C.POINTER(oracle_ffi.C.c_float)
    d = np.array([0.0, 0.0, -1.0], np.float32)   # towards the front face
    origins = [np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), 10.0 + k], np.float32) for k in range(50)]
    out = []
    for accel in (False, True):
        sc = oracle_ffi.OracleScene(scene, accel=accel)
        res = []
        for o in origins:
            prim = oracle_ffi.C.c_uint32(0)
            tout = oracle_ffi.C.c_float(0)
            h = lib.vkro_trace_ray(oracle_ffi.C.byref(sc.c), o.ctypes.data_as(fp), d.ctypes.data_as(fp), 1e30, 1e-3,
                                   oracle_ffi.C.byref(prim), oracle_ffi.C.byref(tout))
            res.append((h, prim.value, tout.value))
        out.append(res)
    assert out[0] == out[1]
    assert all(h == 1 and p == 100 + 399 for h, p, _ in out[0]), out[0][:3]


def test_accel_soup100k_sample_equals_scan():
    """The benchmark frame (100k triangles: the builder's multithreaded path,
    which small scenes never reach) on 40 pixels: the same-BVH trace equals
    the linear scan, so the full-size GPU checks that use it compare against
    the reference's semantics, not against the product's builder."""
    w = scenes.soup100k()
    px = np.random.default_rng(9).choice(w.width * w.height, 40, replace=False).astype(np.uint32)
    ref = oracle_ffi.trace(w.scene, w.frame, w.width, w.height, spp=w.spp, max_depth=w.max_depth, pixels=px)
    acc = oracle_ffi.trace(w.scene, w.frame, w.width, w.height, spp=w.spp, max_depth=w.max_depth, pixels=px,
                           accel=True)
    for k in OUT_KEYS:
        assert np.array_equal(ref[k], acc[k]), k


def test_accel_heightfield_sample_equals_scan():
    """The 2.09M-triangle heightfield (8 spp, 8 bounces) on 24 pixels: the
    same-BVH trace equals the linear scan at full size."""
    w = scenes.heightfield()
    px = np.random.default_rng(10).choice(w.width * w.height, 24, replace=False).astype(np.uint32)
    ref = oracle_ffi.trace(w.scene, w.frame, w.width, w.height, spp=w.spp, max_depth=w.max_depth, pixels=px)
    acc = oracle_ffi.trace(w.scene, w.frame, w.width, w.height, spp=w.spp, max_depth=w.max_depth, pixels=px,
                           accel=True)
    assert int(ref["ray_count"].sum()) > 0
    for k in OUT_KEYS:
        assert np.array_equal(ref[k], acc[k]), k


def test_accel_chain_beyond_the_window():
    """Ten parallel copies of a triangle 0.9 EPSILON apart along the ray, in
    scan order from the nearest: the scan accepts each in turn (t < best +
    EPSILON) and ends 8.1 EPSILON past the nearest hit, beyond the 4 EPSILON
    window the replay starts with; the window is widened until the chain ends
    inside it."""
    lib = oracle_ffi.load()
    k = np.arange(10, dtype=np.float32)
    z = (-k * np.float32(0.9e-3)).astype(np.float32)
    v0 = np.stack([np.full(10, -2.0), np.full(10, -2.0), z], axis=1).astype(np.float32)
    v1 = np.stack([np.full(10, 2.0), np.full(10, -2.0), z], axis=1).astype(np.float32)
    v2 = np.stack([np.full(10, 0.0), np.full(10, 2.0), z], axis=1).astype(np.float32)
    # Wound to face a ray going down -z.
    tris = scenes.tris_from(v0, v1, v2)
    other = scenes.random_soup(np.random.default_rng(6), 100, lo=30.0, hi=50.0)
    scene = vkr.Scene(tris=np.concatenate([tris, other]))
    fp = oracle_ffi.C.POINTER(oracle_ffi.C.c_float)
    d = np.array([0.0, 0.0, -1.0], np.float32)
    out = []
    for accel in (False, True):
        sc = oracle_ffi.OracleScene(scene, accel=accel)
        res = []
        for x in (-0.3, 0.0, 0.2):
            o = np.array([x, 0.1, 10.0], np.float32)
            prim = oracle_ffi.C.c_uint32(0)
            tout = oracle_ffi.C.c_float(0)
            h = lib.vkro_trace_ray(oracle_ffi.C.byref(sc.c), o.ctypes.data_as(fp), d.ctypes.data_as(fp), 1e30, 1e-3,
                                   oracle_ffi.C.byref(prim), oracle_ffi.C.byref(tout))
            res.append((h, prim.value, tout.value))
        out.append(res)
    assert all(h == 1 and p == 9 for h, p, _ in out[0]), out[0]
    assert out[0] == out[1]
