"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars (from the path's contract): primitive id bit-exact, hit t within 1e-5
relative, radiance within 1e-4 per channel, on at least 99.99% of pixels;
RNG and building blocks bit-exact except the documented double-rounding
cases (about one call in 2^28).
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


def _assert_bar(st, ray_count=True):
    # A whole frame compared with the oracle must have no closest-hit ray
    # beyond the replay window (the documented near-tie case).
    assert st.get("window_overflow", 0) == 0, st
    assert st["prim_id"] >= BAR, st
    assert st["hit_t_1e-5"] >= BAR, st
    assert st["radiance_1e-4"] >= BAR, st
    if ray_count:
        assert st["ray_count"] >= BAR, st


# --------------------------------------------------------------------------
# Building blocks
# --------------------------------------------------------------------------
def _eval(ctx, torch, op, inp):
    dev = torch.device("cuda:0")
    i = torch.from_numpy(np.ascontiguousarray(inp, np.float32)).to(dev)
    o = torch.zeros((inp.shape[0], 8), dtype=torch.float32, device=dev)
    ctx.debug_eval(op, i, o)
    torch.cuda.synchronize()
    return o.cpu().numpy()


def test_rand_stream_bit_exact(ctx, torch_cuda):
    rng = np.random.default_rng(1)
    n = 200_000
    inp = np.zeros((n, 16), np.float32)
    inp[:, 0] = rng.random(n)
    inp[:, 1] = rng.random(n)
    inp[:, 2] = (rng.random(n) * 2000.0).astype(np.float32)
    inp[:10, 0] = 0.0  # the reference's x = 0 column: salt = inf / NaN
    inp[:5, 2] = np.inf
    got = _eval(ctx, torch_cuda, "rand", inp)
    lib = oracle_ffi.load()
    ref = np.zeros((n, 8), np.float32)
    buf = (oracle_ffi.C.c_float * 8)()
    for k in range(n):
        lib.vkro_rand_sequence(float(inp[k, 0]), float(inp[k, 1]), float(inp[k, 2]), 8, buf)
        ref[k] = np.frombuffer(buf, np.float32)
    same = (got.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(got) & np.isnan(ref))
    frac = same.mean()
    print(f"\n[rand] bit-exact fraction {frac:.9f} ({(~same).sum()} of {same.size})")
    assert frac >= 1.0 - 1e-6


def test_primitive_tests_bit_exact(ctx, torch_cuda):
    rng = np.random.default_rng(2)
    n = 100_000
    lib = oracle_ffi.load()
    o = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    # triangles around the origin so a good share is hit
    v0 = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    v1 = (v0 + rng.uniform(-3, 3, (n, 3))).astype(np.float32)
    v2 = (v0 + rng.uniform(-3, 3, (n, 3))).astype(np.float32)
    d_to = (v0 - o)
    d[: n // 2] = (d_to[: n // 2] / np.linalg.norm(d_to[: n // 2], axis=1, keepdims=True)).astype(np.float32)
    inp = np.zeros((n, 16), np.float32)
    inp[:, 0:3], inp[:, 3:6], inp[:, 6:9], inp[:, 9:12], inp[:, 12:15] = o, d, v0, v1, v2
    inp[:, 15] = 1e-3
    got = _eval(ctx, torch_cuda, "tri", inp)[:, 0]
    tri = np.zeros(1, vkr.TRIANGLE)
    ref = np.zeros(n, np.float32)
    for k in range(n):
        tri["v0"], tri["v1"], tri["v2"] = v0[k], v1[k], v2[k]
        ref[k] = lib.vkro_tri_intersect(oracle_ffi.f3(o[k]), oracle_ffi.f3(d[k]), tri.ctypes.data, 1e-3)
    assert (got.view(np.uint32) == ref.view(np.uint32)).mean() == 1.0
    assert (ref > 0).mean() > 0.05

    c = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    r = rng.uniform(0.1, 3, n).astype(np.float32)
    inp = np.zeros((n, 16), np.float32)
    inp[:, 0:3], inp[:, 3:6], inp[:, 6:9], inp[:, 9] = o, d, c, r
    got = _eval(ctx, torch_cuda, "sphere", inp)[:, 0]
    sph = np.zeros(1, vkr.SPHERE)
    ref = np.zeros(n, np.float32)
    for k in range(n):
        sph["P"], sph["r"] = c[k], r[k]
        ref[k] = lib.vkro_sphere_intersect(oracle_ffi.f3(o[k]), oracle_ffi.f3(d[k]), sph.ctypes.data)
    assert (got.view(np.uint32) == ref.view(np.uint32)).mean() == 1.0

    N = rng.normal(size=(n, 3)).astype(np.float32)
    ln = rng.uniform(-5, 5, n).astype(np.float32)
    N[:100] = 0.0  # degenerate
    inp = np.zeros((n, 16), np.float32)
    inp[:, 0:3], inp[:, 3:6], inp[:, 6:9], inp[:, 9] = o, d, N, ln
    got = _eval(ctx, torch_cuda, "plane", inp)[:, 0]
    pl = np.zeros(1, vkr.PLANE)
    ref = np.zeros(n, np.float32)
    for k in range(n):
        pl["N"], pl["len"] = N[k], ln[k]
        ref[k] = lib.vkro_plane_intersect(oracle_ffi.f3(o[k]), oracle_ffi.f3(d[k]), pl.ctypes.data)
    same = (got.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(got) & np.isnan(ref))
    assert same.mean() == 1.0


def test_shading_blocks_bit_exact(ctx, torch_cuda):
    rng = np.random.default_rng(3)
    n = 50_000
    lib = oracle_ffi.load()

    def unit(k):
        v = rng.normal(size=(k, 3))
        return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)

    Nn, H, V, L = unit(n), unit(n), unit(n), unit(n)
    rough = rng.uniform(0, 1, n).astype(np.float32)
    cs = rng.uniform(-0.2, 1.0, n).astype(np.float32)
    f0 = rng.uniform(0, 1, n).astype(np.float32)
    inp = np.zeros((n, 16), np.float32)
    inp[:, 0:3], inp[:, 3:6], inp[:, 6:9], inp[:, 9:12] = Nn, H, V, L
    inp[:, 12], inp[:, 13], inp[:, 14] = rough, cs, f0
    got = _eval(ctx, torch_cuda, "brdf", inp)
    ref = np.zeros((n, 6), np.float32)
    fo = (oracle_ffi.C.c_float * 3)()
    for k in range(n):
        f3 = oracle_ffi.f3
        ref[k, 0] = lib.vkro_distribution_ggx(f3(Nn[k]), f3(H[k]), float(rough[k]))
        ref[k, 1] = lib.vkro_geometry_smith(f3(Nn[k]), f3(V[k]), f3(L[k]), float(rough[k]))
        lib.vkro_fresnel_schlick(float(cs[k]), f3([f0[k]] * 3), fo)
        ref[k, 2:5] = np.frombuffer(fo, np.float32)
        ref[k, 5] = lib.vkro_schlick(float(cs[k]), float(rough[k]))
    g = got[:, :6]
    same = (g.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(g) & np.isnan(ref))
    assert same.mean() >= 1.0 - 1e-6, same.mean()

    d = unit(n)
    phi = rng.uniform(0, 6.2831855, n).astype(np.float32)
    ca = rng.uniform(0, 1, n).astype(np.float32)
    sa = np.sqrt(1 - ca * ca).astype(np.float32)
    inp = np.zeros((n, 16), np.float32)
    inp[:, 0:3], inp[:, 3], inp[:, 4], inp[:, 5] = d, phi, sa, ca
    got = _eval(ctx, torch_cuda, "jitter", inp)[:, :3]
    ref = np.zeros((n, 3), np.float32)
    for k in range(n):
        lib.vkro_jitter(oracle_ffi.f3(d[k]), float(phi[k]), float(sa[k]), float(ca[k]), fo)
        ref[k] = np.frombuffer(fo, np.float32)
    same = (got.view(np.uint32) == ref.view(np.uint32))
    assert same.mean() >= 1.0 - 1e-5, same.mean()


# --------------------------------------------------------------------------
# Whole frames
# --------------------------------------------------------------------------
def test_reference_scene_full_frame(ctx, torch_cuda):
    """The reference's own scene, camera and constants: 1024^2, 4 spp, 4 bounces."""
    wl = scenes.reference_scene(1024, 1024)
    sid = ctx.upload(wl.scene)
    g = gpu_trace(ctx, sid, wl)
    r = oracle_ffi.trace(wl.scene, wl.frame, wl.width, wl.height, wl.spp, wl.max_depth)
    st = compare(g, r)
    _report("reference_scene", st)
    _assert_bar(st)
    assert st["rgba8"] >= BAR
    rays = ctx.counters()["rays"]
    assert rays == int(r["ray_count"].sum())


def test_spheres64_full_frame(ctx, torch_cuda):
    wl = scenes.spheres64()
    sid = ctx.upload(wl.scene)
    g = gpu_trace(ctx, sid, wl)
    r = oracle_ffi.trace(wl.scene, wl.frame, wl.width, wl.height, wl.spp, wl.max_depth)
    st = compare(g, r)
    _report("spheres64", st)
    _assert_bar(st)


def test_soup_pixel_subset(ctx, torch_cuda):
    """100k-triangle soup + 257 spheres at 1920x1080 (the benchmark frame):
    the oracle's linear scan on 100 pixels (corners and centre included), and
    the oracle tracing the host builder's tree (bit-equal to the scan,
    tests/test_oracle_accel.py) on 16384 more."""
    wl = scenes.soup100k()
    sid = ctx.upload(wl.scene)
    info = ctx.info(sid)
    print(f"\n[soup100k] bvh {info}", flush=True)
    g = gpu_trace(ctx, sid, wl)
    rng = np.random.default_rng(7)
    W, H = wl.width, wl.height
    pix = rng.choice(W * H, 96, replace=False).astype(np.uint32)
    pix = np.concatenate([pix, np.array([0, W - 1, (H // 2) * W + W // 2, (H - 1) * W + W - 1], np.uint32)])
    r = oracle_ffi.trace(wl.scene, wl.frame, W, H, wl.spp, wl.max_depth, pixels=pix)
    st = compare(g, r, pixels=pix)
    _report("soup100k_subset", st)
    assert st["prim_id"] == 1.0 and st["radiance_1e-4"] == 1.0
    pix = np.random.default_rng(17).choice(W * H, 16384, replace=False).astype(np.uint32)
    r = oracle_ffi.trace(wl.scene, wl.frame, W, H, wl.spp, wl.max_depth, pixels=pix, accel=True)
    st = compare(g, r, pixels=pix)
    _report("soup100k_16384px", st)
    _assert_bar(st)


def test_far_sphere_imprecise_hits(ctx, torch_cuda):
    """Rays from 2000 units away: the reference's sphere test accepts rays
    that pass outside the sphere (rounding in its discriminant); the traversal
    widens the sphere boxes per ray so it finds the same hits."""
    wl = scenes.far_sphere()
    sid = ctx.upload(wl.scene)
    g = gpu_trace(ctx, sid, wl)
    r = oracle_ffi.trace(wl.scene, wl.frame, wl.width, wl.height, wl.spp, wl.max_depth)
    # Float64 distance from each primary ray to the sphere's centre.
    W, H = wl.width, wl.height
    y, x = np.divmod(np.arange(W * H), W)
    cam = wl.frame["camera"]
    d = (cam["dir"][None].astype(np.float64) + cam["right"][None] * (2.0 * x / W - 1.0)[:, None]
         + cam["up"][None] * (2.0 * y / H - 1.0)[:, None])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    oc = cam["pos"].astype(np.float64) - np.array(wl.scene.spheres[0]["P"], np.float64)
    b = d @ oc
    p = np.sqrt(np.maximum(oc @ oc - b * b, 0.0))
    outside = (r["prim_id"] == (1 << 30)) & (p > 1.0001)
    st = compare(g, r)
    _report("far_sphere", st)
    print(f"[far_sphere] oracle hits on rays passing outside the sphere: {int(outside.sum())}", flush=True)
    assert outside.sum() > 100
    assert st["window_overflow"] == 0
    assert st["prim_id"] == 1.0 and st["hit_t_bits"] == 1.0 and st["radiance_bits"] == 1.0 and st["ray_count"] == 1.0


def test_heightfield_pixel_subset(ctx, torch_cuda):
    """2.09M-triangle fBm heightfield + 4096 metal/dielectric spheres, 8 spp,
    8 bounces at 1920x1080: the oracle's linear scan on 24 pixels, the oracle
    over the host builder's tree on 4096 more."""
    wl = scenes.heightfield()
    sid = ctx.upload(wl.scene)
    info = ctx.info(sid)
    print(f"\n[heightfield] bvh {info}", flush=True)
    g = gpu_trace(ctx, sid, wl)
    cnt = ctx.counters()
    rng = np.random.default_rng(11)
    W, H = wl.width, wl.height
    pix = rng.choice(W * H, 24, replace=False).astype(np.uint32)
    r = oracle_ffi.trace(wl.scene, wl.frame, W, H, wl.spp, wl.max_depth, pixels=pix)
    st = compare(g, r, pixels=pix)
    _report("heightfield_subset", st)
    print(f"[heightfield] counters {cnt}", flush=True)
    assert st["prim_id"] == 1.0 and st["radiance_1e-4"] == 1.0
    # The oracle tracing the host builder's tree (bit-equal to the scan).
    pix = np.random.default_rng(12).choice(W * H, 4096, replace=False).astype(np.uint32)
    r = oracle_ffi.trace(wl.scene, wl.frame, W, H, wl.spp, wl.max_depth, pixels=pix, accel=True)
    st = compare(g, r, pixels=pix)
    _report("heightfield_4096px", st)
    _assert_bar(st)


def test_schedules_identical(ctx, torch_cuda):
    """Plain and persistent launches, every block size: identical bits."""
    wl = scenes.reference_scene(256, 192, spp=2, max_depth=4)
    sid = ctx.upload(wl.scene)
    base = gpu_trace(ctx, sid, wl, sched=vkr.SCHED_PLAIN, block=128)
    for sched in (vkr.SCHED_PLAIN, vkr.SCHED_PERSISTENT):
        for block in (64, 128, 256):
            for chunk in (1, 32):
                g = gpu_trace(ctx, sid, wl, sched=sched, block=block, chunk=chunk)
                for k in ("rgba8", "radiance", "prim_id", "hit_t", "ray_count"):
                    assert np.array_equal(g[k].view(np.uint8), base[k].view(np.uint8)), (sched, block, chunk, k)


def test_ray_pool_variants_identical(ctx, torch_cuda):
    """The pooled kernel (several shadow rays per lane and step, traced from a
    wave-shared pool) and the single-ray kernel give identical bits, on a
    scene with 18 emitters and on one with a single emitter."""
    for wl in (scenes.spheres64(), scenes.reference_scene(256, 192, spp=2, max_depth=4)):
        sid = ctx.upload(wl.scene)
        a = gpu_trace(ctx, sid, wl, flags=vkr.TRACE_RAY_POOL)
        b = gpu_trace(ctx, sid, wl, flags=vkr.TRACE_NO_RAY_POOL)
        for k in ("rgba8", "radiance", "prim_id", "hit_t", "ray_count"):
            assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), (wl.name, k)
    r = oracle_ffi.trace(wl.scene, wl.frame, wl.width, wl.height, wl.spp, wl.max_depth)
    st = compare(a, r)
    assert st["prim_id"] == 1.0 and st["radiance_bits"] == 1.0, st


def test_emitter_windows_full_frame(ctx, torch_cuda):
    """The pooled kernel evaluates next-event samples in windows of 32
    emitters; 69, 33 and 32 emitters give three, two and one window per
    diffuse hit.  Whole frames against the oracle, every bit."""
    for n, seed, n_em in ((140, 0, 69), (66, 3, 33), (60, 5, 32)):
        wl = scenes.many_emitters(n, seed)
        sid = ctx.upload(wl.scene)
        assert ctx.info(sid)["n_emitters"] == n_em
        g = gpu_trace(ctx, sid, wl)
        r = oracle_ffi.trace(wl.scene, wl.frame, wl.width, wl.height, wl.spp, wl.max_depth)
        st = compare(g, r)
        _report(wl.name, st)
        assert st["prim_id"] == 1.0 and st["radiance_bits"] == 1.0 and st["ray_count"] == 1.0, st
        assert st["rgba8"] == 1.0, st


def test_plain_launch_split(ctx, torch_cuda):
    """One pixel per lane above 2^21 pixels runs as several launches (bounded
    ray records); the image is identical to the persistent schedule's."""
    wl = scenes.reference_scene(2048, 1088, spp=1, max_depth=2)
    sid = ctx.upload(wl.scene)
    a = gpu_trace(ctx, sid, wl, sched=vkr.SCHED_PLAIN)
    b = gpu_trace(ctx, sid, wl, sched=vkr.SCHED_PERSISTENT)
    for k in ("rgba8", "radiance", "prim_id", "hit_t", "ray_count"):
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


def test_odd_sizes_and_x0_column(ctx, torch_cuda):
    for (w, h) in ((1, 1), (7, 5), (33, 17)):
        wl = scenes.reference_scene(w, h, spp=3, max_depth=4)
        sid = ctx.upload(wl.scene)
        g = gpu_trace(ctx, sid, wl)
        r = oracle_ffi.trace(wl.scene, wl.frame, w, h, wl.spp, wl.max_depth)
        st = compare(g, r)
        assert st["prim_id"] == 1.0 and st["radiance_1e-4"] == 1.0 and st["rgba8"] == 1.0, (w, h, st)


def test_whitted_full_frame(ctx, torch_cuda):
    wl = scenes.reference_scene(512, 512)
    g = gpu_trace(ctx, 0, wl, kernel=vkr.KERNEL_WHITTED)
    r = oracle_ffi.trace(wl.scene, wl.frame, wl.width, wl.height, kernel=oracle_ffi.KERNEL_WHITTED)
    st = compare(g, r)
    _report("whitted", st)
    _assert_bar(st)
    assert st["rgba8"] >= BAR


def test_filter_matches_oracle(ctx, torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(9)
    sw, sh, dw, dh = 256, 256, 320, 240
    cur = rng.integers(0, 256, (sw * sh, 4), dtype=np.uint8)
    prev = cur.copy()
    prev[: sw * sh // 2] = rng.integers(0, 256, (sw * sh // 2, 4), dtype=np.uint8)
    dev = torch.device("cuda:0")
    tc, tp = torch.from_numpy(cur).to(dev), torch.from_numpy(prev).to(dev)
    to = torch.zeros((dw * dh, 4), dtype=torch.uint8, device=dev)
    ctx.filter(tc, tp, sw, sh, to, dw, dh)
    torch.cuda.synchronize()
    ref = np.zeros((dw * dh, 4), np.uint8)
    oracle_ffi.load().vkro_filter(cur.ctypes.data, prev.ctypes.data, sw, sh, ref.ctypes.data, dw, dh)
    assert np.array_equal(to.cpu().numpy(), ref)
