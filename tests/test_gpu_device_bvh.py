"""GPU parity of scenes whose triangle subtree is built on the device
(VKR_SCENE_DEVICE_BVH: a linear BVH over Morton codes, lbvh.hip).

The tree differs from the host SAH tree but the traversal replays the
reference's scan-order acceptance rule (Assets/Tracer.comp:374-431) over the
hits it finds, so every image must equal the oracle's linear scan, as for the
host-built tree.  Full-size workloads are checked against the oracle tracing
the host builder's tree (bit-equal to the scan, tests/test_oracle_accel.py) on
large pixel samples.
"""
import numpy as np
import pytest

import oracle_ffi
import scenes
import vkr
from gpu_util import compare, gpu_trace

pytestmark = pytest.mark.gpu

BAR = 0.9999


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def ctx(torch_cuda):
    c = vkr.Context()
    yield c
    c.close()


def _report(name, st):
    print(f"\n[{name}] " + " ".join(f"{k}={v:.6f}" if isinstance(v, float) else f"{k}={v}" for k, v in st.items()),
          flush=True)


def _assert_bar(st):
    assert st["prim_id"] >= BAR, st
    assert st["hit_t_1e-5"] >= BAR, st
    assert st["radiance_1e-4"] >= BAR, st
    assert st["ray_count"] >= BAR, st


def _soup(n_tris=3000, seed=0, **kw):
    return scenes.soup100k(seed=seed, n_tris=n_tris, width=kw.get("width", 96), height=kw.get("height", 54),
                           spp=2, max_depth=4, n_spheres=48)


def _tris_only():
    """Triangles and a plane, no sphere: the device tree is the whole BVH."""
    w = _soup(seed=4)
    sc = vkr.Scene(tris=w.scene.tris, planes=w.scene.planes)
    return scenes.Workload("tris_only", sc, w.frame, w.width, w.height, w.spp, w.max_depth)


def _non_finite():
    """Some triangles with NaN or infinite vertices: never hit, and their
    leaves must not disturb the boxes of the others."""
    w = _soup(seed=5)
    tris = w.scene.tris.copy()
    tris["v0"][::37, 0] = np.nan
    tris["v2"][5::53, 1] = np.inf
    tris["v1"][9::61, 2] = -np.inf
    sc = vkr.Scene(tris=tris, spheres=w.scene.spheres, planes=w.scene.planes)
    return scenes.Workload("non_finite", sc, w.frame, w.width, w.height, w.spp, w.max_depth)


def _two_tris():
    """The smallest device build: two triangles (one internal node) + a light."""
    M = scenes.reference_materials()
    tris = scenes.tris_from(np.array([[-5, -5, 0], [-5, 5, 0]], np.float32),
                            np.array([[5, -5, 0], [5, 5, 0]], np.float32),
                            np.array([[-5, 5, 0], [5, -5, 0]], np.float32))
    sph = scenes.spheres_from([M["light"]], [(0.0, 0.0, -20.0)], [4.0])
    cam = vkr.camera_from_angles((0.0, 0.0, -30.0), 0.0, 90.0)
    sc = vkr.Scene(tris=tris, spheres=sph)
    return scenes.Workload("two_tris", sc, scenes.frame(cam, 1.0), 64, 64, 2, 3)


def _coincident():
    """Four copies of one triangle: every centre is the same point, so the
    centre bounds are empty, every Morton key is equal and the tree splits by
    index alone.  The scan's `t < best + EPSILON` makes the last copy win every
    hit (four copies: the kernel keeps up to four candidates within 4 EPSILON
    of the nearest hit, more are a documented near-tie)."""
    M = scenes.reference_materials()
    one = scenes.tris_from(np.array([[-4, -4, 0]], np.float32), np.array([[-4, 4, 0]], np.float32),
                           np.array([[4, -4, 0]], np.float32))
    sph = scenes.spheres_from([M["light"], M["matte_white"]], [(0.0, 12.0, -8.0), (6.0, 6.0, -3.0)], [4.0, 1.5])
    cam = vkr.camera_from_angles((0.0, 0.0, -10.0), 0.0, 90.0)
    sc = vkr.Scene(tris=np.repeat(one, 4), spheres=sph)
    return scenes.Workload("coincident", sc, scenes.frame(cam, 1.0), 48, 48, 2, 3)


def _all_non_finite():
    """Only non-finite triangles (never hit) beside the spheres64 scene: the
    whole triangle subtree gets boxes no ray enters."""
    w = scenes.spheres64()
    tris = scenes.random_soup(np.random.default_rng(4), 100, lo=-5.0, hi=5.0)
    tris["v1"][:, 0] = np.nan
    sc = vkr.Scene(tris=tris, spheres=w.scene.spheres, planes=w.scene.planes)
    return scenes.Workload("all_non_finite", sc, w.frame, 128, 128, w.spp, w.max_depth)


CASES = {
    "soup3000": _soup,
    "many_emitters": scenes.many_emitters,
    "far_sphere": lambda: scenes.far_sphere(res=96),
    "tris_only": _tris_only,
    "non_finite": _non_finite,
    "two_tris": _two_tris,
    "coincident": _coincident,
    "all_non_finite": _all_non_finite,
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_bvh_full_frame(ctx, torch_cuda, name):
    wl = CASES[name]()
    print(f"\n[device_bvh {name}] {len(wl.scene.tris)} tris {len(wl.scene.spheres)} spheres "
          f"{wl.width}x{wl.height}", flush=True)
    sid = ctx.upload(wl.scene, device_bvh=True)
    info = ctx.info(sid)
    print(f"[device_bvh {name}] {info}", flush=True)
    assert info["builder"] == vkr.BUILDER_DEVICE_LBVH
    n_t, n_s = len(wl.scene.tris), len(wl.scene.spheres)
    assert info["n_leaves"] >= 1 and info["n_nodes"] >= 1
    g = gpu_trace(ctx, sid, wl)
    r = oracle_ffi.trace(wl.scene, wl.frame, wl.width, wl.height, wl.spp, wl.max_depth)
    st = compare(g, r)
    _report(f"device_bvh {name}", st)
    _assert_bar(st)
    ctx.destroy_scene(sid)


@pytest.mark.parametrize("max_leaf", [2, 4, 8])
def test_device_bvh_leaf_runs(ctx, torch_cuda, max_leaf):
    """Subtrees of at most max_leaf triangles collapsed into one leaf (a run
    of records): fewer live nodes, the same image."""
    wl = _soup(seed=7)
    sid = ctx.upload(wl.scene, max_leaf=max_leaf, device_bvh=True)
    info = ctx.info(sid)
    print(f"\n[device_bvh leaf {max_leaf}] {info}", flush=True)
    assert info["builder"] == vkr.BUILDER_DEVICE_LBVH
    n_t = len(wl.scene.tris)
    assert info["n_leaves"] < n_t and info["n_nodes"] < n_t
    g = gpu_trace(ctx, sid, wl)
    r = oracle_ffi.trace(wl.scene, wl.frame, wl.width, wl.height, wl.spp, wl.max_depth)
    st = compare(g, r)
    _report(f"device_bvh leaf {max_leaf}", st)
    _assert_bar(st)
    ctx.destroy_scene(sid)


def test_device_bvh_depth_cap_falls_back_to_host(ctx, torch_cuda):
    """A depth cap the Morton tree does not meet (3000 triangles in leaves of
    two, no sphere root: 11 levels at best, the Morton tree needs more) but
    the host builder does: the host builder runs instead, and the image is
    the same."""
    wl = _tris_only()
    sid = ctx.upload(wl.scene, max_leaf=2, max_depth=12, device_bvh=True)
    info = ctx.info(sid)
    print(f"\n[device_bvh fallback] {info}", flush=True)
    assert info["builder"] == vkr.BUILDER_HOST_SAH and info["depth"] <= 12
    g = gpu_trace(ctx, sid, wl)
    r = oracle_ffi.trace(wl.scene, wl.frame, wl.width, wl.height, wl.spp, wl.max_depth)
    st = compare(g, r)
    _report("device_bvh fallback", st)
    _assert_bar(st)
    ctx.destroy_scene(sid)


def _full_size(ctx, wl, n_pix, seed):
    sid_h = ctx.upload(wl.scene)
    sid_d = ctx.upload(wl.scene, device_bvh=True)
    ih, idv = ctx.info(sid_h), ctx.info(sid_d)
    print(f"\n[{wl.name}] host {ih}\n[{wl.name}] device {idv}", flush=True)
    assert idv["builder"] == vkr.BUILDER_DEVICE_LBVH
    gh = gpu_trace(ctx, sid_h, wl)
    gd = gpu_trace(ctx, sid_d, wl)
    st = compare(gd, gh)
    _report(f"{wl.name} device vs host tree", st)
    _assert_bar(st)
    W, H = wl.width, wl.height
    pix = np.random.default_rng(seed).choice(W * H, n_pix, replace=False).astype(np.uint32)
    r = oracle_ffi.trace(wl.scene, wl.frame, W, H, wl.spp, wl.max_depth, pixels=pix, accel=True)
    st = compare(gd, r, pixels=pix)
    _report(f"{wl.name} device tree vs oracle ({n_pix} px)", st)
    _assert_bar(st)
    ctx.destroy_scene(sid_h)
    ctx.destroy_scene(sid_d)


def test_device_bvh_soup100k(ctx, torch_cuda):
    """1920x1080, 100k triangles + 257 spheres: the device-built tree gives the
    host-built tree's image, and the oracle's on 8192 pixels."""
    _full_size(ctx, scenes.soup100k(), 8192, 21)


def test_device_bvh_heightfield(ctx, torch_cuda):
    """1920x1080, 2.09M-triangle heightfield + 4097 spheres, 8 spp, 8 bounces."""
    _full_size(ctx, scenes.heightfield(), 2048, 22)
