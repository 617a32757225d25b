"""The GPU's work counters against the oracle tracing the same BVH.

The benchmark's algorithmic bytes (roofline.achieved) are computed from the
kernel's debug counters: node visits, triangle, sphere and plane tests.  The
oracle's same-BVH trace (oracle/bvh_accel.cpp, bit-equal to the reference's
linear scan) walks the host builder's tree with the same near-first order,
the same slab test and the same cut-off distance, so for the same rays it
counts the same events.  Scenes without emitters are used: with emitters the
kernel skips shadow rays whose contribution is provably zero, and the oracle
traces every one (as the reference does), so the counts differ by design.
"""
import numpy as np
import pytest

import oracle_ffi
import scenes
import vkr
from gpu_util import compare, gpu_trace

pytestmark = pytest.mark.gpu


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


def _dark(wl):
    """The workload with every emission set to zero (no next-event rays)."""
    sph = wl.scene.spheres.copy()
    sph["mat"]["emissive"] = 0.0
    pln = wl.scene.planes.copy()
    sc = vkr.Scene(tris=wl.scene.tris, spheres=sph, planes=pln)
    return scenes.Workload(wl.name + "_dark", sc, wl.frame, wl.width, wl.height, wl.spp, wl.max_depth)


CASES = {
    "soup3000": lambda: _dark(scenes.soup100k(n_tris=3000, width=128, height=72, spp=2, max_depth=4, n_spheres=48)),
    "spheres64": lambda: _dark(scenes.spheres64()),
    "heightfield": lambda: _dark(scenes.heightfield(n=96, extent=128.0, n_spheres=256, width=128, height=72,
                                                    spp=2, max_depth=4)),
}


@pytest.mark.parametrize("pool", [False, True], ids=["single_ray", "ray_pool"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_work_counters_equal_oracle(ctx, torch_cuda, name, pool):
    wl = CASES[name]()
    print(f"\n[counters {name} pool={pool}] {len(wl.scene.tris)} tris {len(wl.scene.spheres)} spheres "
          f"{wl.width}x{wl.height}x{wl.spp}", flush=True)
    sid = ctx.upload(wl.scene)
    flags = vkr.TRACE_RAY_POOL if pool else vkr.TRACE_NO_RAY_POOL
    ctx.set_flags(vkr.CTX_COUNT_WORK)
    try:
        g = gpu_trace(ctx, sid, wl, flags=flags)
        cnt = ctx.counters()
    finally:
        ctx.set_flags(0)
    osc = oracle_ffi.OracleScene(wl.scene, accel=True)
    osc.count_work()
    r = oracle_ffi.trace(wl.scene, wl.frame, wl.width, wl.height, wl.spp, wl.max_depth, oscene=osc)
    work = osc.work()
    print(f"[counters {name} pool={pool}] gpu {cnt}\n[counters {name} pool={pool}] oracle {work}", flush=True)
    st = compare(g, r)
    assert st["prim_id"] == 1.0 and st["radiance_bits"] == 1.0, st
    assert work["rays"] == cnt["rays"] == cnt["rays_traced"] == int(r["ray_count"].sum())
    assert cnt["shadow_rays"] == 0
    for k in ("nodes_visited", "tris_tested", "spheres_tested", "planes_tested"):
        assert cnt[k] == work[k], (k, cnt[k], work[k])
    ctx.destroy_scene(sid)
