"""Benchmark: Mrays/s (all bounces) of the gfx950 path kernel on one MI355X.

Headline workload (one "step" = one frame): 1920x1080, 4 spp, max 4
bounces, 100k random triangles (vertices U[-50,50]^3, edges U[0.1,2]) + 256
random spheres + 1 emissive light sphere + a ground plane, seed 0, FP32.
Rays counted are every trace_ray call of the reference program (camera,
bounce and next-event shadow rays), read from the kernel's own counters.

Prints ONE JSON line.  Besides the headline it reports:
  roofline      the dominant kernel (k_path) against HBM: measured HBM bytes
                per launch (rocprof FETCH_SIZE + WRITE_SIZE, calibrated) over
                the event-timed kernel; the algorithmic request bytes (BVH
                nodes, primitive records, output), the measured L2 rate,
                VALU issue utilisation and the node-loop SIMD efficiency
                beside it, since the kernel is bound by issue and latency;
  cpu_baseline  the CPU restatement of the reference tracing the same BVH as
                the GPU (oracle/bvh_accel.cpp, results identical to the
                linear scan) on the host threads this process may use (at
                most 16, a GPU's share of its host) on a pixel sample of the
                same frame; cpu_scan_baseline is the literal linear scan;
  batch64       64 scenes (1e3..1e6 triangles, 256^2..1080p, 1..8 spp): one
                plain launch per scene, one persistent launch per scene, and
                ONE batched persistent launch replayed from a captured
                hipGraph, at the default and at the autotuned shape;
  heightfield   the 2.09M-triangle heightfield, 1080p, 8 spp, 8 bounces;
  adjoint       the 256^2 / 16 spp adjoint workload: forward, adjoint on its
                own, and the forward + adjoint pair of a differentiable render
                step (the forward keeps the adjoint's first pass).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "vk-renderer_amd"), os.path.join(ROOT, "tests")]

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 78.6   # non-packed FP32 FMA: 256 CU x 128 FLOP/clk x 2.4 GHz
# FP32 operations per unit of work (slab test of one box, one Moller-Trumbore
# triangle test, one sphere test), for the secondary compute figure.
BOX_FLOP, TRI_FLOP, SPHERE_FLOP = 20, 50, 20
# Request bytes per unit of work as this build lays them out: a 64-byte node
# (two child boxes), a 48-byte triangle record, a 16-byte sphere record plus
# its 4-byte scan key, a plane's normal and offset, an rgba8 pixel.  The
# std430 sizes of a Sphere and a Plane are 64 bytes each.
NODE_B, TRI_B, SPHERE_B, PLANE_B, OUT_B = 64, 48, 20, 16, 4
STD430_SPHERE_B = STD430_PLANE_B = 64
CPU_THREAD_CAP = 16       # a GPU's share of its host's threads


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sched", type=int, default=0, help="0 auto, 1 plain, 2 persistent")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--cpu-pixels", type=int, default=393216, help="pixel sample of the same-BVH CPU baseline")
    ap.add_argument("--cpu-scan-pixels", type=int, default=64, help="pixel sample of the linear-scan CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batch64 section of the default run")
    ap.add_argument("--no-heightfield", action="store_true", help="skip the heightfield section of the default run")
    ap.add_argument("--batch-steps", type=int, default=2, help="timed repetitions per batch64 variant")
    ap.add_argument("--workload", default="soup100k", choices=["soup100k", "batch64", "bvhbuild"])
    ap.add_argument("--scenes", type=int, default=64, help="batch64: number of scenes")
    ap.add_argument("--device-bvh", action="store_true", help="trace scenes whose triangle BVH was built on the GPU")
    ap.add_argument("--tris", type=int, default=100_000, help="soup workload: number of triangles (the headline is 100000)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


METRIC = "Mrays/s (all bounces) at 1080p 100k-tri, plain vs batched vs adjoint, 1 GPU"


def host_cpu():
    """Host threads this process may use (capped at a GPU's share), the
    machine's logical CPU count and its CPU model."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return min(CPU_THREAD_CAP, avail), os.cpu_count(), avail, model


def bench_adjoint(ctx, torch, repeats=5):
    """Adjoint workload (256x256, 16 spp, 4 bounces, 32 diffuse/metal spheres
    + 1 light) with a dense dLoss/dradiance, each figure over `repeats` calls:
    the forward radiance pass; the adjoint on its own (both passes: calls
    alternate between two frame seeds, so none finds its first pass kept);
    the forward that keeps the adjoint's first pass (VKR_TRACE_RECORD_ADJOINT)
    and the adjoint that then runs only its second pass; and the pair a
    differentiable render runs per step (recording forward, then adjoint)."""
    import numpy as np

    import scenes
    import vkr

    wl = scenes.adjoint_scene()
    sid = ctx.upload(wl.scene)
    n = wl.width * wl.height
    rad = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    g4 = torch.ones((n, 4), dtype=torch.float32, device="cuda")
    d = torch.zeros(7 * len(wl.scene.spheres) + 1, dtype=torch.float32, device="cuda")
    p = wl.params()
    pr = wl.params(flags=vkr.TRACE_RECORD_ADJOINT)
    frames = [wl.frame.copy(), wl.frame.copy()]
    frames[1]["seed"] = frames[0]["seed"] + np.float32(1.0)
    it = {"k": 0}

    def alternating():
        it["k"] ^= 1
        ctx.adjoint(sid, frames[it["k"]], p, g4, d)

    def pair():
        ctx.trace(sid, wl.frame, pr, radiance=rad)
        ctx.adjoint(sid, wl.frame, p, g4, d)

    res = {}
    for name, fn in (("forward", lambda: ctx.trace(sid, wl.frame, p, radiance=rad)),
                     ("adjoint", alternating),
                     ("forward_recording", lambda: ctx.trace(sid, wl.frame, pr, radiance=rad)),
                     ("adjoint_recorded", lambda: ctx.adjoint(sid, wl.frame, p, g4, d)),
                     ("forward_backward", pair)):
        fn()
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(repeats):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - ts) / repeats
    ctx.trace(sid, wl.frame, p, radiance=rad)
    torch.cuda.synchronize()
    rays = ctx.counters()["rays"]
    ctx.destroy_scene(sid)
    ms = {k: round(v * 1e3, 3) for k, v in res.items()}
    return {"workload": "adjoint_256x256_16spp_4bounce_33spheres", "forward_ms": ms["forward"],
            "adjoint_ms": ms["adjoint"], "adjoint_over_forward": round(res["adjoint"] / res["forward"], 3),
            "forward_recording_ms": ms["forward_recording"], "adjoint_recorded_ms": ms["adjoint_recorded"],
            "forward_backward_ms": ms["forward_backward"],
            "forward_backward_over_forward": round(res["forward_backward"] / res["forward"], 3),
            "forward_mrays": round(rays / res["forward"] / 1e6, 2),
            "adjoint_mrays": round(rays / res["adjoint"] / 1e6, 2)}


def batch_section(ctx, torch, n_scenes=64, steps=2, block=0, chunk=0, device_bvh=False, tune=True):
    """64 scenes (1e3..1e6 triangles log-uniform, 256^2..1080p, 1..8 spp, 4
    bounces) traced: one plain launch per scene, one persistent launch per
    scene, and ONE batched persistent launch (eager, and replayed from a
    captured hipGraph), at the default shape and at the batch autotuner's.
    Returns (result dict, seconds per batch of the tuned graph)."""
    import scenes
    import vkr

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    wls = scenes.batch_jobs(n_scenes)
    hsids = [ctx.upload(wl.scene) for wl in wls]
    ctx.destroy_scene(ctx.upload(wls[0].scene, device_bvh=True))   # loads the build kernels
    dsids = [ctx.upload(wl.scene, device_bvh=True) for wl in wls]
    build_ms = {"host_sah": round(sum(ctx.info(s)["build_ms"] for s in hsids), 2),
                "device": round(sum(ctx.info(s)["build_ms"] for s in dsids), 2)}
    for s in (hsids if device_bvh else dsids):
        ctx.destroy_scene(s)
    sids = dsids if device_bvh else hsids
    log(f"batch: built {len(wls)} scenes in {time.perf_counter() - t0:.1f} s; BVH build ms (sum) {build_ms}; "
        f"tris {sum(len(w.scene.tris) for w in wls)}, pixels*spp {sum(w.width * w.height * w.spp for w in wls)}")
    outs = [torch.zeros((wl.width * wl.height, 4), dtype=torch.uint8, device=dev) for wl in wls]

    def per_scene(sched):
        for wl, sid, o in zip(wls, sids, outs):
            ctx.trace(sid, wl.frame, wl.params(sched=sched), rgba8=o)

    jobs = [ctx.job(sid, wl.frame, wl.params(), rgba8=o) for wl, sid, o in zip(wls, sids, outs)]
    ctx.trace_batch(jobs, block=block or 128, chunk=chunk or 8)
    torch.cuda.synchronize()
    cnt = ctx.counters()
    rays, traced = cnt["rays"], cnt["rays_traced"]
    g_default = ctx.batch_graph(jobs, block=block or 128, chunk=chunk or 8)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - ts) / steps

    res = {}
    res["plain_per_scene"] = timed(lambda: per_scene(vkr.SCHED_PLAIN))
    log(f"batch: plain per scene {res['plain_per_scene']:.3f} s")
    res["persistent_per_scene"] = timed(lambda: per_scene(vkr.SCHED_PERSISTENT))
    log(f"batch: persistent per scene {res['persistent_per_scene']:.3f} s")
    res["batched"] = timed(lambda: ctx.trace_batch(jobs, block=block or 128, chunk=chunk or 8))
    res["batched_graph"] = timed(lambda: ctx.graph_launch(g_default))
    log(f"batch: batched {res['batched']:.3f} s, graph {res['batched_graph']:.3f} s")
    ctx.graph_destroy(g_default)
    tuned = None
    if tune and not (block or chunk):
        ts = time.perf_counter()
        tuned = ctx.autotune_batch(jobs, repeats=1)
        tune_s = time.perf_counter() - ts
        g_tuned = ctx.batch_graph(jobs)          # block = chunk = 0: the tuned shape
        res["batched_graph_tuned"] = timed(lambda: ctx.graph_launch(g_tuned))
        ctx.graph_destroy(g_tuned)
        log(f"batch: autotuned {tuned['block']}/{tuned['chunk']}/{tuned['stack']} in {tune_s:.1f} s, "
            f"graph {res['batched_graph_tuned']:.3f} s")
    best = res.get("batched_graph_tuned", res["batched_graph"])
    mr = {k: round(rays / v / 1e6, 2) for k, v in res.items()}
    out = {
        "workload": f"batch{len(wls)}_1e3-1e6tris_256sq-1080p_1-8spp_4bounce",
        "rays_per_batch": rays, "rays_traversed_per_batch": traced,
        "seconds_per_batch": {k: round(v, 4) for k, v in res.items()},
        "mrays": mr,
        "speedup_batched_vs_plain_per_scene": round(res["plain_per_scene"] / best, 3),
        "speedup_batched_vs_persistent_per_scene": round(res["persistent_per_scene"] / best, 3),
        "tuned_shape": None if tuned is None else {k: tuned[k] for k in ("block", "chunk", "stack")},
        "tuned_vs_default_graph_ms": None if tuned is None else
        [round(res["batched_graph_tuned"] * 1e3, 2), round(res["batched_graph"] * 1e3, 2)],
        "autotune_table": None if tuned is None else tuned["table"],
        "bvh_build_ms_all_scenes": build_ms, "bvh": "device" if device_bvh else "host_sah",
    }
    for s in sids:
        ctx.destroy_scene(s)
    return out, best


def heightfield_section(ctx, torch, steps=3):
    """2.09M-triangle fBm heightfield + 4096 metal/dielectric spheres + 1
    light, 1920x1080, 8 spp, 8 bounces: event-timed frames."""
    import scenes

    wl = scenes.heightfield()
    sid = ctx.upload(wl.scene)
    info = ctx.info(sid)
    rgba = torch.zeros((wl.width * wl.height, 4), dtype=torch.uint8, device="cuda")
    p = wl.params()
    ctx.trace(sid, wl.frame, p, rgba8=rgba)
    torch.cuda.synchronize()
    ms = []
    for _ in range(steps):
        ctx.trace(sid, wl.frame, p, rgba8=rgba)
        ms.append(ctx.last_kernel_ms())
    cnt = ctx.counters()
    ctx.destroy_scene(sid)
    k = sum(ms) / len(ms)
    return {"workload": "heightfield2093058_1920x1080_8spp_8bounce", "n_tris": info["n_tris"],
            "n_spheres": info["n_spheres"], "build_ms_host_sah": round(info["build_ms"], 2),
            "kernel_ms": round(k, 3), "rays_per_frame": cnt["rays"], "rays_traversed_per_frame": cnt["rays_traced"],
            "mrays": round(cnt["rays"] / k / 1e3, 2), "traversed_mrays": round(cnt["rays_traced"] / k / 1e3, 2)}


def bench_batch(a):
    """--workload batch64: the batch section on its own, as the JSON line."""
    import torch

    import vkr

    ctx = vkr.Context()
    res, best = batch_section(ctx, torch, a.scenes, max(1, a.steps), a.block, a.chunk, a.device_bvh)
    out = {
        "metric": METRIC,
        "value": res["mrays"]["batched_graph_tuned" if "batched_graph_tuned" in res["mrays"] else "batched_graph"],
        "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps, "warmup": 1,
        "ms_per_step": round(best * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded PCG64 scenes)",
        "config": {"workload": res["workload"], "block": a.block, "chunk": a.chunk},
        "batch64": res,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_bvh_build(a):
    """BVH build on upload, host SAH vs device (VKR_SCENE_DEVICE_BVH), on the
    100k-triangle soup and the 2.09M-triangle heightfield: build ms (median
    of `steps` uploads after one warm-up upload) and the frame time traced
    on each tree."""
    import torch

    import scenes
    import vkr

    ctx = vkr.Context()
    res = {}
    for wl in (scenes.soup100k(), scenes.heightfield()):
        r = {}
        for dev_b in (False, True):
            name = "device" if dev_b else "host_sah"
            ctx.destroy_scene(ctx.upload(wl.scene, device_bvh=dev_b))   # warm-up (loads the build kernels)
            ms = []
            for _ in range(max(1, a.steps)):
                sid = ctx.upload(wl.scene, device_bvh=dev_b)
                ms.append(ctx.info(sid)["build_ms"])
                ctx.destroy_scene(sid)
            sid = ctx.upload(wl.scene, device_bvh=dev_b)
            info = ctx.info(sid)
            rgba = torch.zeros((wl.width * wl.height, 4), dtype=torch.uint8, device="cuda")
            p = wl.params()
            ctx.trace(sid, wl.frame, p, rgba8=rgba)
            torch.cuda.synchronize()
            ts = time.perf_counter()
            for _ in range(3):
                ctx.trace(sid, wl.frame, p, rgba8=rgba)
            torch.cuda.synchronize()
            frame_ms = (time.perf_counter() - ts) / 3 * 1e3
            ctx.destroy_scene(sid)
            r[name] = {"build_ms": round(sorted(ms)[len(ms) // 2], 3), "frame_ms": round(frame_ms, 3),
                       "nodes": info["n_nodes"], "depth": info["depth"], "builder": info["builder"]}
        r["build_speedup"] = round(r["host_sah"]["build_ms"] / r["device"]["build_ms"], 1)
        res[wl.name] = r
        log(f"{wl.name}: {r}")
    out = {"metric": "BVH build ms on upload (host SAH vs device)", "value": res["heightfield2093058_1920x1080"]
           ["device"]["build_ms"], "unit": "ms", "n_gpus": 1, "steps": a.steps, "warmup": 1,
           "higher_is_better": False, "data": "synthetic (seeded PCG64 scenes)",
           "config": {"workload": "bvhbuild_soup100k_heightfield2m"}, "results": res}
    print(json.dumps(out), flush=True)
    ctx.close()


def load_profile(name):
    """A committed rocprof summary of the headline kernel (profiles/*.json)."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1",
                          "n_gpus": a.gpus}), flush=True)
        sys.exit(2)
    if a.workload in ("batch64", "bvhbuild"):
        import torch
        if not torch.cuda.is_available():
            raise SystemExit("bench.py needs a GPU")
        (bench_batch if a.workload == "batch64" else bench_bvh_build)(a)
        return

    import numpy as np
    import torch

    import scenes
    import vkr

    vkr.load()
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    wl = scenes.soup100k(n_tris=a.tris)
    W, H = wl.width, wl.height
    ctx = vkr.Context()
    bvh_build_ms = {}
    for dev_b in (False, True, True):   # the first device build also loads the build kernels
        s_ = ctx.upload(wl.scene, device_bvh=dev_b)
        bvh_build_ms["device" if dev_b else "host_sah"] = round(ctx.info(s_)["build_ms"], 3)
        ctx.destroy_scene(s_)
    sid = ctx.upload(wl.scene, device_bvh=a.device_bvh)
    info = ctx.info(sid)
    log(f"scene: {info}; BVH build ms {bvh_build_ms}")
    rgba = torch.zeros((W * H, 4), dtype=torch.uint8, device=dev)
    p = wl.params(sched=a.sched, block=a.block, chunk=a.chunk)

    # Work-counting pass (separate kernel variant, not timed).
    ctx.set_flags(vkr.CTX_COUNT_WORK)
    ctx.trace(sid, wl.frame, p, rgba8=rgba)
    work = ctx.counters()
    ctx.set_flags(0)
    log(f"work per frame: {work}")

    for _ in range(a.warmup):
        ctx.trace(sid, wl.frame, p, rgba8=rgba)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        evs[i][0].record(stream)
        ctx.trace(sid, wl.frame, p, rgba8=rgba)   # launched on torch's current stream
        evs[i][1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    last_kernel_ms = ctx.last_kernel_ms()
    cnt = ctx.counters()
    launch_ms = [s.elapsed_time(e) for s, e in evs]
    kern_ms = float(np.mean(launch_ms))
    elapsed = t1 - t0
    rays = cnt["rays"]
    value = rays * a.steps / elapsed / 1e6
    traversed_mrays = cnt["rays_traced"] * a.steps / elapsed / 1e6
    log(f"rays/frame {rays} shadow {cnt['shadow_rays']} overflow {cnt['window_overflow']} "
        f"launch ms mean {kern_ms:.3f} min {min(launch_ms):.3f} last kernel (lib events) {last_kernel_ms:.3f}")

    # Request bytes of the traversal and shading (work counters of the same
    # kernel's counting variant): mostly L2 hits, so not an HBM figure.
    req_bytes = (work["nodes_visited"] * NODE_B + work["tris_tested"] * TRI_B +
                 work["spheres_tested"] * SPHERE_B + work["planes_tested"] * PLANE_B + W * H * OUT_B)
    req_bytes_std430 = (work["nodes_visited"] * NODE_B + work["tris_tested"] * TRI_B +
                        work["spheres_tested"] * STD430_SPHERE_B + work["planes_tested"] * STD430_PLANE_B +
                        W * H * OUT_B)
    kern_s = kern_ms * 1e-3
    flops = (work["nodes_visited"] * 2 * BOX_FLOP + work["tris_tested"] * TRI_FLOP +
             work["spheres_tested"] * SPHERE_FLOP)
    # The rocprof files describe the 100k-triangle frame on the host
    # builder's tree with the default launch only.
    measured = a.tris == 100_000 and not a.device_bvh and not (a.sched or a.block or a.chunk)
    tp = load_profile("traffic_soup100k.json") if measured else None
    lp = load_profile("l2_soup100k.json") if measured else None
    traffic = tp.get("hbm_bytes_per_launch") if tp else None
    l2_bytes = (lp["l2_read_bytes_per_launch"] + lp["l2_write_bytes_per_launch"]) if lp else None
    hbm_gbs = traffic / kern_s / 1e9 if traffic else None
    node_eff = work["nodes_visited"] / (64.0 * work["wave_node_iters"]) if work["wave_node_iters"] else None
    roofline = {
        "bound": "hbm",
        # measured HBM bytes per launch over this run's event-timed kernel
        "achieved": round(hbm_gbs, 1) if hbm_gbs else None,
        "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(hbm_gbs / HBM_PEAK_GBS, 4) if hbm_gbs else None,
        "traffic": traffic,
        "achieved_basis": "rocprof FETCH_SIZE + WRITE_SIZE per dispatch (calibrated on a copy) / kernel time",
        "kernel_ms": round(kern_ms, 4),
        # The kernel is not HBM-bound: it is bound by VALU issue and the
        # latency of dependent node fetches (L2 hits).
        "limiter": "VALU issue and dependent-fetch latency",
        "valu_issue_util": tp.get("valu_issue_util") if tp else None,
        "node_loop_simd_eff": round(node_eff, 4) if node_eff else None,
        "leaf_simd_eff": round((work["tris_tested"] + work["spheres_tested"]) / (64.0 * work["wave_leaf_iters"]), 4)
        if work["wave_leaf_iters"] else None,
        "node_visits_per_traversed_ray": round(work["nodes_visited"] / max(1, work["rays_traced"]), 2),
        "l2_gbs_measured": round(l2_bytes / kern_s / 1e9, 1) if l2_bytes else None,
        "l2_hit_rate": tp.get("l2_hit_rate") if tp else None,
        # Algorithmic request bytes per launch (node 64 B, triangle 48 B,
        # sphere 20 B, plane 16 B, rgba8 4 B; with std430 spheres and planes
        # of 64 B in request_bytes_std430_units): served mostly by L2.
        "request_bytes_per_launch": int(req_bytes),
        "request_gbs": round(req_bytes / kern_s / 1e9, 1),
        "request_bytes_std430_units": int(req_bytes_std430),
        "fp32_tflops_est": round(flops / kern_s / 1e12, 3),
        "fp32_peak_tflops": FP32_PEAK_TFLOPS,
    }

    variants = {}
    if not a.no_variants:
        for name, sched in (("plain", vkr.SCHED_PLAIN), ("persistent", vkr.SCHED_PERSISTENT)):
            q = wl.params(sched=sched, block=a.block, chunk=a.chunk)
            ctx.trace(sid, wl.frame, q, rgba8=rgba)
            torch.cuda.synchronize()
            ts = time.perf_counter()
            for _ in range(3):
                ctx.trace(sid, wl.frame, q, rgba8=rgba)
            torch.cuda.synchronize()
            variants[name] = round(ctx.counters()["rays"] * 3 / (time.perf_counter() - ts) / 1e6, 2)
        log(f"variants Mrays/s: {variants}")
    ctx.destroy_scene(sid)

    adjoint = None
    if not a.no_adjoint:
        adjoint = bench_adjoint(ctx, torch)
        log(f"adjoint: {adjoint}")

    heightfield = None
    if not a.no_heightfield:
        heightfield = heightfield_section(ctx, torch)
        log(f"heightfield: {heightfield}")

    batch = None
    if not a.no_batch:
        batch, _ = batch_section(ctx, torch, steps=max(1, a.batch_steps))
        log(f"batch64: { {k: v for k, v in batch.items() if k != 'autotune_table'} }")

    cpu = cpu_scan = None
    if not a.no_cpu_baseline:
        import oracle_ffi
        threads, nproc, avail, model = host_cpu()

        def cpu_leg(n_pix, accel, what):
            pix = np.random.default_rng(0).choice(W * H, n_pix, replace=False).astype(np.uint32)
            osc = oracle_ffi.OracleScene(wl.scene, accel=accel)   # BVH build outside the timed region
            ts = time.perf_counter()
            r = oracle_ffi.trace(wl.scene, wl.frame, W, H, wl.spp, wl.max_depth, pixels=pix, threads=threads,
                                 oscene=osc)
            dt = time.perf_counter() - ts
            n_rays = int(r["ray_count"].sum())
            v = n_rays / dt / 1e6
            return {"value": round(v, 6), "unit": "Mrays/s", "cores": threads, "kind": "port",
                    "sample": f"{n_pix} random pixels of the same frame (all {wl.spp} spp x {wl.max_depth} "
                              f"bounces, {n_rays} rays, {dt:.1f} s), {what}",
                    "host_logical_cpus": nproc, "cpus_available": avail, "cpu_model": model,
                    "threads_note": f"all {avail} CPUs available to this process, capped at {CPU_THREAD_CAP} "
                                    f"(a GPU's share of its host)",
                    "gpu_over_cpu": round(value / v, 1),
                    # the CPU traces every shadow ray; the GPU skips those whose
                    # contribution is exactly 0, so compare traversed rays too
                    "gpu_traversed_mrays": round(traversed_mrays, 2),
                    "gpu_traversed_over_cpu": round(traversed_mrays / v, 1)}

        cpu = cpu_leg(a.cpu_pixels, True, "the oracle tracing the same BVH as the GPU (scalar, one ray at a "
                                         "time, every shadow ray traced)")
        log(f"cpu baseline (same BVH): {cpu}")
        if a.cpu_scan_pixels > 0:
            cpu_scan = cpu_leg(a.cpu_scan_pixels, False, "reference linear scan over every primitive")
            log(f"cpu baseline (linear scan): {cpu_scan}")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded PCG64 scene, FrameData.seed 0)",
        "config": {"workload": f"soup{wl.scene.tris.shape[0] // 1000}k_1920x1080_4spp_4bounce", "width": W, "height": H,
                   "spp": wl.spp,
                   "max_depth": wl.max_depth, "n_tris": info["n_tris"], "n_spheres": info["n_spheres"],
                   "n_planes": info["n_planes"], "n_emitters": info["n_emitters"],
                   "rays_per_frame": rays, "rays_traversed_per_frame": cnt["rays_traced"],
                   "traversed_mrays": round(traversed_mrays, 2),
                   "window_overflow_per_frame": cnt["window_overflow"],
                   "sched": a.sched, "block": a.block, "chunk": a.chunk,
                   "bvh": "device" if a.device_bvh else "host_sah"},
        "bvh_build_ms": bvh_build_ms,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_scan_baseline": cpu_scan,
        "variants_mrays": variants,
        "batch64": batch,
        "heightfield": heightfield,
        "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
