"""GPU: the batched workload itself -- 64 independent scenes (random soups of
1e3..1e6 triangles, log-uniform, plus 257 spheres), images of 256^2, 512^2,
1280x720 or 1920x1080, 1..8 spp, 4 bounces -- traced as ONE graph-captured
persistent launch.  Every job must equal its own single-frame trace bit for
bit (Assets/Tracer.comp:557-593 per job), and pixel samples of the smallest
and the largest scene must equal the oracle."""
import numpy as np
import pytest

import oracle_ffi
import scenes
import vkr
from gpu_util import compare

pytestmark = pytest.mark.gpu

KEYS = ("rgba8", "radiance", "prim_id", "hit_t", "ray_count")


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _bufs(torch, wl):
    dev = torch.device("cuda:0")
    n = wl.width * wl.height
    return {"rgba8": torch.zeros((n, 4), dtype=torch.uint8, device=dev),
            "radiance": torch.zeros((n, 4), dtype=torch.float32, device=dev),
            "prim_id": torch.zeros(n, dtype=torch.int32, device=dev),
            "hit_t": torch.zeros(n, dtype=torch.float32, device=dev),
            "ray_count": torch.zeros(n, dtype=torch.int32, device=dev)}


def _same(torch, x, y):
    if x.dtype == torch.float32:
        x, y = x.view(torch.int32), y.view(torch.int32)
    return torch.equal(x, y)


def test_batch64_graph_matches_single_frames_and_oracle(torch_cuda):
    torch = torch_cuda
    wls = scenes.batch_jobs(64)
    tris = [len(w.scene.tris) for w in wls]
    print(f"\n[c4] {len(wls)} scenes, {sum(tris)} triangles ({min(tris)}..{max(tris)}), "
          f"{sum(w.width * w.height for w in wls)} pixels, spp {sorted({w.spp for w in wls})}", flush=True)
    with vkr.Context() as ctx:
        sids = [ctx.upload(wl.scene) for wl in wls]
        single, rays, overflow = [], 0, 0
        for wl, sid in zip(wls, sids):
            b = _bufs(torch, wl)
            ctx.trace(sid, wl.frame, wl.params(), **b)
            c = ctx.counters()
            rays += c["rays"]
            overflow += c["window_overflow"]
            single.append(b)
        print(f"[c4] single frames done: {rays} rays, {overflow} window overflows", flush=True)
        batched = [_bufs(torch, wl) for wl in wls]
        jobs = [ctx.job(sid, wl.frame, wl.params(), **b) for wl, sid, b in zip(wls, sids, batched)]
        gid = ctx.batch_graph(jobs)
        ctx.graph_launch(gid)
        ms = ctx.last_kernel_ms()
        c = ctx.counters()
        print(f"[c4] batched graph: {ms:.1f} ms, {c['rays']} rays", flush=True)
        assert c["rays"] == rays
        for i, (s, b) in enumerate(zip(single, batched)):
            for k in KEYS:
                assert _same(torch, s[k], b[k]), (i, k)
        ctx.graph_destroy(gid)

        # Oracle pixel samples: the smallest scene against the literal scan,
        # both against the oracle tracing the host builder's tree (bit-equal
        # to the scan, tests/test_oracle_accel.py).
        rng = np.random.default_rng(64)
        for name, i, accel, npix in (("smallest/scan", int(np.argmin(tris)), False, 512),
                                     ("smallest/tree", int(np.argmin(tris)), True, 2048),
                                     ("largest/tree", int(np.argmax(tris)), True, 2048)):
            wl = wls[i]
            pix = rng.choice(wl.width * wl.height, npix, replace=False).astype(np.uint32)
            r = oracle_ffi.trace(wl.scene, wl.frame, wl.width, wl.height, wl.spp, wl.max_depth, pixels=pix,
                                 accel=accel)
            g = {k: v.cpu().numpy() for k, v in batched[i].items()}
            g["prim_id"] = g["prim_id"].view(np.uint32)
            g["ray_count"] = g["ray_count"].view(np.uint32)
            st = compare(g, r, pixels=pix)
            print(f"[c4 {name}] job {i}: {tris[i]} tris {wl.width}x{wl.height} spp {wl.spp}: " +
                  " ".join(f"{k}={v}" for k, v in st.items()), flush=True)
            assert st["prim_id"] == 1.0 and st["hit_t_bits"] == 1.0 and st["ray_count"] == 1.0, st
            assert st["radiance_1e-4"] == 1.0 and st["rgba8"] == 1.0, st
        for sid in sids:
            ctx.destroy_scene(sid)
