"""GPU: the path-replay adjoint against central finite differences with
common random numbers and frozen branch decisions (h = 1e-3)."""
import numpy as np
import pytest

import scenes
import vkr

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


def _forward(ctx, torch, sid, wl, theta, flags=0, decisions=None):
    ctx.set_sphere_params(sid, theta)
    n = wl.width * wl.height
    rad = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    ctx.trace(sid, wl.frame, wl.params(flags=flags), radiance=rad, decisions=decisions)
    torch.cuda.synchronize()
    return rad[:, :3].double().cpu().numpy()


def _setup(ctx, torch, w=256, spp=16):
    wl = scenes.adjoint_scene(width=w, height=w, spp=spp)
    sid = ctx.upload(wl.scene)
    theta = scenes.sphere_params(wl.scene.spheres, light_scale=1.0)
    dec = torch.zeros(w * w * spp, dtype=torch.int16, device="cuda")
    base = _forward(ctx, torch, sid, wl, theta, vkr.TRACE_RECORD_DECISIONS, dec)
    return wl, sid, theta, dec, base


def _fd(ctx, torch, sid, wl, theta, dec, i, weights, h=1e-3):
    h = h * max(1.0, abs(float(theta[i])))
    tp, tm = theta.copy(), theta.copy()
    tp[i] += h
    tm[i] -= h
    lp = _forward(ctx, torch, sid, wl, tp, vkr.TRACE_REPLAY_DECISIONS, dec)
    lm = _forward(ctx, torch, sid, wl, tm, vkr.TRACE_REPLAY_DECISIONS, dec)
    return float(np.sum(weights * (lp - lm)) / (float(tp[i]) - float(tm[i])))


def test_adjoint_matches_finite_differences(ctx, torch_cuda):
    torch = torch_cuda
    import diff
    wl, sid, theta, dec, base = _setup(ctx, torch)
    n_sph = len(wl.scene.spheres)
    rng = np.random.default_rng(0)
    target = _forward(ctx, torch, sid, wl, (theta * (1 + rng.uniform(-0.2, 0.2, theta.size))).astype(np.float32))
    # L2 loss against the perturbed-parameter target: dLoss/dL = 2 (L - L*)
    t = torch.tensor(theta, device="cuda", requires_grad=True)
    rad = diff.render(t, ctx, sid, wl.frame, wl.params())
    loss = ((rad.double() - torch.from_numpy(target).cuda()) ** 2).sum()
    loss.backward()
    g = t.grad.double().cpu().numpy()
    weights = 2.0 * (base - target)
    # Every parameter that can move the image: albedo and roughness of the 32
    # spheres, the light's emission and the light scale.  The light's own
    # albedo (1,1,1) is left out: with mask (1,1,1) the max3 of Russian
    # roulette (:545) has an exact three-way tie there, a kink where a
    # central difference averages two one-sided slopes.
    idx = [7 * s + k for s in range(n_sph - 1) for k in (0, 1, 2, 6)]
    idx += [7 * (n_sph - 1) + k for k in (3, 4, 5)] + [7 * n_sph]
    fd = np.array([_fd(ctx, torch, sid, wl, theta, dec, i, weights) for i in idx])
    ga = g[idx]
    scale = np.abs(fd).max()
    rel = np.abs(ga - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    worst = np.argsort(-rel)[:8]
    print("\n[adjoint] max|fd| %.4g  median rel %.3g  max rel %.3g" % (scale, np.median(rel), rel.max()))
    for k in worst:
        print(f"  param {idx[k]:4d} adjoint {ga[k]: .6e} fd {fd[k]: .6e} rel {rel[k]:.3g}")
    big = np.abs(fd) > 1e-3 * scale
    assert big.sum() >= 40
    # aggregate relative error of the gradient vector at h = 1e-3
    agg = np.linalg.norm(ga - fd) / np.linalg.norm(fd)
    print(f"[adjoint] aggregate relative error {agg:.3g}; "
          f"{(rel[big] <= 1e-3).mean() * 100:.1f}% of {big.sum()} parameters within 1e-3")
    assert agg <= 1e-3
    assert (rel[big] <= 1e-3).mean() >= 0.95
    # The rest sit near kinks of the estimate (clamp(acc) saturating at 1,
    # :441, or max3 changing channel, :545) that a step of 1e-3 straddles;
    # their difference quotient converges to the adjoint as h shrinks.
    for k in np.where(big & (rel > 1e-3))[0]:
        i = idx[k]
        fd_small = _fd(ctx, torch, sid, wl, theta, dec, i, weights, h=1e-3 / 16)
        r2 = abs(ga[k] - fd_small) / abs(fd_small)
        print(f"  param {i}: rel {rel[k]:.3g} at h=1e-3 -> {r2:.3g} at h=6.25e-5")
        assert r2 < rel[k] and r2 <= 5e-3
    # zero-emission spheres: the structural zero (no emitter, :462)
    assert np.all(g[[7 * s + 3 for s in range(n_sph - 1)]] == 0.0)


def test_adjoint_odd_size_matches_finite_differences(ctx, torch_cuda):
    """45x29 (partial 8x8 tiles) at 3 spp: the second pass takes one work item
    per (pixel, sample), a tile's 64 pixels sample by sample, and must still
    give every pixel's samples their own stream positions."""
    torch = torch_cuda
    w, h, spp = 45, 29, 3
    wl = scenes.adjoint_scene(width=w, height=h, spp=spp)
    sid = ctx.upload(wl.scene)
    theta = scenes.sphere_params(wl.scene.spheres, light_scale=1.0)
    dec = torch.zeros(w * h * spp, dtype=torch.int16, device="cuda")
    base = _forward(ctx, torch, sid, wl, theta, vkr.TRACE_RECORD_DECISIONS, dec)
    weights = np.random.default_rng(3).uniform(-1.0, 1.0, base.shape)
    g4 = torch.zeros((w * h, 4), dtype=torch.float32, device="cuda")
    g4[:, :3] = torch.from_numpy(weights.astype(np.float32)).cuda()
    d = torch.zeros(theta.size, dtype=torch.float32, device="cuda")
    ctx.set_sphere_params(sid, theta)
    ctx.adjoint(sid, wl.frame, wl.params(), g4, d)
    torch.cuda.synchronize()
    g = d.double().cpu().numpy()
    n_sph = len(wl.scene.spheres)
    idx = [7 * s + k for s in range(n_sph - 1) for k in (0, 1, 2, 6)]
    idx += [7 * (n_sph - 1) + k for k in (3, 4, 5)] + [7 * n_sph]
    fd = np.array([_fd(ctx, torch, sid, wl, theta, dec, i, weights) for i in idx])
    ga = g[idx]
    scale = np.abs(fd).max()
    big = np.abs(fd) > 1e-3 * scale
    agg = np.linalg.norm(ga[big] - fd[big]) / np.linalg.norm(fd[big])
    rel = np.abs(ga - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    print(f"\n[adjoint {w}x{h}x{spp}] {big.sum()} parameters, aggregate relative error {agg:.3g}, "
          f"{(rel[big] <= 1e-3).mean() * 100:.1f}% within 1e-3")
    assert big.sum() >= 20
    assert agg <= 1e-3
    ctx.destroy_scene(sid)


def test_adjoint_cost(ctx, torch_cuda):
    torch = torch_cuda
    import time
    wl = scenes.adjoint_scene()
    sid = ctx.upload(wl.scene)
    n = wl.width * wl.height
    rad = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    g4 = torch.ones((n, 4), dtype=torch.float32, device="cuda")
    d = torch.zeros(7 * len(wl.scene.spheres) + 1, dtype=torch.float32, device="cuda")

    frames = [wl.frame, _other(wl.frame)]
    k = [0]

    def fwd():
        ctx.trace(sid, wl.frame, wl.params(), radiance=rad)

    def adj():
        # Alternating frame seeds: every call runs both passes.
        k[0] ^= 1
        ctx.adjoint(sid, frames[k[0]], wl.params(), g4, d)

    res = {}
    for name, fn in (("forward", fwd), ("adjoint", adj)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / 5
    ratio = res["adjoint"] / res["forward"]
    print(f"\n[adjoint cost] forward {res['forward'] * 1e3:.2f} ms adjoint {res['adjoint'] * 1e3:.2f} ms ratio {ratio:.2f}")
    assert ratio <= 3.0


def test_adjoint_device_tree_matches_host_tree(ctx, torch_cuda):
    """The adjoint replays paths through the scene's BVH: a triangle subtree
    built on the device (VKR_SCENE_DEVICE_BVH) gives the same paths, so the
    same gradient (up to the order of the float additions that reduce it)."""
    torch = torch_cuda
    wl0 = scenes.adjoint_scene(width=128, height=128, spp=8)
    tris = scenes.random_soup(np.random.default_rng(3), 3000, lo=-15.0, hi=15.0)
    sc = vkr.Scene(tris=tris, spheres=wl0.scene.spheres, planes=wl0.scene.planes)
    wl = scenes.Workload("adjoint_tris", sc, wl0.frame, 128, 128, 8, wl0.max_depth)
    theta = scenes.sphere_params(sc.spheres, light_scale=1.0)
    n = wl.width * wl.height
    g4 = torch.ones((n, 4), dtype=torch.float32, device="cuda")
    grads, rads = [], []
    for dev in (False, True):
        sid = ctx.upload(sc, device_bvh=dev)
        assert ctx.info(sid)["builder"] == (vkr.BUILDER_DEVICE_LBVH if dev else vkr.BUILDER_HOST_SAH)
        rads.append(_forward(ctx, torch, sid, wl, theta))
        d = torch.zeros(7 * len(sc.spheres) + 1, dtype=torch.float32, device="cuda")
        ctx.adjoint(sid, wl.frame, wl.params(), g4, d)
        torch.cuda.synchronize()
        grads.append(d.double().cpu().numpy())
        ctx.destroy_scene(sid)
    assert np.array_equal(rads[0], rads[1])
    gh, gd = grads
    agg = np.linalg.norm(gd - gh) / np.linalg.norm(gh)
    print(f"\n[adjoint host vs device tree] |g| {np.linalg.norm(gh):.4g} aggregate relative difference {agg:.3g}")
    assert np.linalg.norm(gh) > 0 and agg <= 1e-5


def test_adjoint_is_linear_in_the_output_gradient(ctx, torch_cuda):
    """At the full adjoint workload (256^2, 16 spp, 4 bounces): the parameter
    gradient is linear in the image gradient, adjoint(2 g1) = 2 adjoint(g1)
    and adjoint(g1 + g2) = adjoint(g1) + adjoint(g2), up to the order of the
    float additions that reduce it (which varies between calls)."""
    torch = torch_cuda
    wl = scenes.adjoint_scene()
    sid = ctx.upload(wl.scene)
    n = wl.width * wl.height
    gen = torch.Generator(device="cuda").manual_seed(5)
    g1 = torch.rand((n, 4), device="cuda", generator=gen) - 0.5
    g2 = torch.rand((n, 4), device="cuda", generator=gen) - 0.5

    def adj(g):
        d = torch.zeros(7 * len(wl.scene.spheres) + 1, dtype=torch.float32, device="cuda")
        ctx.adjoint(sid, wl.frame, wl.params(), g.contiguous(), d)
        torch.cuda.synchronize()
        return d.double().cpu().numpy()

    a1, a2, a12, a1x2, a1b = adj(g1), adj(g2), adj(g1 + g2), adj(2.0 * g1), adj(g1)
    scale = np.linalg.norm(a1) + np.linalg.norm(a2)
    e_scale = np.linalg.norm(a1x2 - 2.0 * a1) / (2.0 * np.linalg.norm(a1))
    e_sum = np.linalg.norm(a12 - a1 - a2) / scale
    # The same call twice differs only in the order of the float atomics
    # that reduce the gradient (1e-7..2e-6 relative here): the bound for the
    # two identities sits above that noise.
    e_rep = np.linalg.norm(a1b - a1) / np.linalg.norm(a1)
    print(f"\n[adjoint linearity] |a1| {np.linalg.norm(a1):.4g} scale err {e_scale:.3g} sum err {e_sum:.3g} "
          f"repeat {e_rep:.3g}")
    assert np.linalg.norm(a1) > 0
    assert e_rep <= 1e-5 and e_scale <= 1e-5 and e_sum <= 1e-5
    ctx.destroy_scene(sid)


def _scene_with(sc, theta):
    """The scene as vkr_scene_set_sphere_params leaves it: per sphere albedo,
    emissive = L * emission (float32) and roughness."""
    sph = sc.spheres.copy()
    n = len(sph)
    L = np.float32(theta[7 * n])
    for i in range(n):
        q = np.asarray(theta[7 * i:7 * i + 7], np.float32)
        sph[i]["mat"]["albedo"] = q[0:3]
        sph[i]["mat"]["emissive"] = L * q[3:6]
        sph[i]["mat"]["roughness"] = q[6]
    return vkr.Scene(tris=sc.tris, spheres=sph, planes=sc.planes, tri_material=sc.tri_material,
                     materials=sc.materials, flags=sc.flags)


def test_adjoint_matches_oracle_finite_differences(ctx, torch_cuda):
    """The reference-semantics gradient: central differences (h = 1e-3, common
    random numbers, branch decisions recorded at theta and replayed) of the
    CPU oracle's radiance (Assets/Tracer.comp:433-553) on a 16384-pixel subset
    of the 256^2 / 16 spp / 4-bounce frame, against vkr_adjoint with the loss
    gradient set on that subset and zero elsewhere."""
    import oracle_ffi
    torch = torch_cuda
    wl = scenes.adjoint_scene()
    W, H = wl.width, wl.height
    n_sph = len(wl.scene.spheres)
    theta = scenes.sphere_params(wl.scene.spheres, light_scale=1.0)
    rng = np.random.default_rng(11)
    pix = np.sort(rng.choice(W * H, 16384, replace=False)).astype(np.uint32)

    def oracle(t, flags=0, dec=None):
        return oracle_ffi.trace(_scene_with(wl.scene, t), wl.frame, W, H, wl.spp, wl.max_depth, pixels=pix,
                                flags=flags, decisions=dec)

    base = oracle(theta, oracle_ffi.FLAG_RECORD)
    dec = base["decisions"].copy()
    target = oracle((theta * (1 + rng.uniform(-0.2, 0.2, theta.size))).astype(np.float32))
    weights = 2.0 * (base["radiance"][:, :3].astype(np.float64) - target["radiance"][:, :3])

    # The GPU forward is the oracle's on these pixels, bit for bit.
    sid = ctx.upload(wl.scene)
    ctx.set_sphere_params(sid, theta)
    rad = torch.zeros((W * H, 4), dtype=torch.float32, device="cuda")
    ctx.trace(sid, wl.frame, wl.params(), radiance=rad)
    torch.cuda.synchronize()
    g_rad = rad.cpu().numpy()[pix]
    assert np.array_equal(g_rad.view(np.uint32), base["radiance"].view(np.uint32))

    g4 = torch.zeros((W * H, 4), dtype=torch.float32, device="cuda")
    g4[torch.from_numpy(pix.astype(np.int64)).cuda(), :3] = torch.from_numpy(weights.astype(np.float32)).cuda()
    d = torch.zeros(7 * n_sph + 1, dtype=torch.float32, device="cuda")
    ctx.adjoint(sid, wl.frame, wl.params(), g4, d)
    torch.cuda.synchronize()
    g = d.double().cpu().numpy()
    ctx.destroy_scene(sid)

    def fd(i, h=1e-3):
        h = h * max(1.0, abs(float(theta[i])))
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        lp = oracle(tp, oracle_ffi.FLAG_REPLAY, dec.copy())["radiance"][:, :3].astype(np.float64)
        lm = oracle(tm, oracle_ffi.FLAG_REPLAY, dec.copy())["radiance"][:, :3].astype(np.float64)
        return float(np.sum(weights * (lp - lm)) / (float(tp[i]) - float(tm[i])))

    # The parameters that can move the image (as in the test above; the
    # light's own albedo sits on an exact max3 tie of Russian roulette).
    idx = [7 * s + k for s in range(n_sph - 1) for k in (0, 1, 2, 6)]
    idx += [7 * (n_sph - 1) + k for k in (3, 4, 5)] + [7 * n_sph]
    f = np.array([fd(i) for i in idx])
    ga = g[idx]
    scale = np.abs(f).max()
    rel = np.abs(ga - f) / np.maximum(np.abs(f), 1e-3 * scale)
    big = np.abs(f) > 1e-3 * scale
    agg = np.linalg.norm(ga - f) / np.linalg.norm(f)
    print(f"\n[adjoint vs oracle FD] {len(pix)} pixels, max|fd| {scale:.4g}, aggregate relative error {agg:.3g}; "
          f"{(rel[big] <= 1e-3).mean() * 100:.1f}% of {big.sum()} parameters within 1e-3")
    for k in np.argsort(-rel)[:6]:
        print(f"  param {idx[k]:4d} adjoint {ga[k]: .6e} oracle fd {f[k]: .6e} rel {rel[k]:.3g}")
    assert big.sum() >= 40
    assert agg <= 1e-3
    assert (rel[big] <= 1e-3).mean() >= 0.95
    # Parameters whose step straddles a kink of the estimate (a clamp
    # saturating, :441, or max3 switching channel, :545): the quotient
    # converges to the adjoint as the step shrinks.
    for k in np.where(big & (rel > 1e-3))[0]:
        f_small = fd(idx[k], h=1e-3 / 16)
        r2 = abs(ga[k] - f_small) / abs(f_small)
        print(f"  param {idx[k]}: rel {rel[k]:.3g} at h=1e-3 -> {r2:.3g} at h=6.25e-5")
        assert r2 < rel[k] and r2 <= 5e-3


def _adjoint(ctx, torch, sid, frame, params, g4, n_par):
    """Gradient and the adjoint's ray count (both passes, or the second only
    when the first pass's records were kept)."""
    d = torch.zeros(n_par, dtype=torch.float32, device="cuda")
    ctx.adjoint(sid, frame, params, g4, d)
    torch.cuda.synchronize()
    return d.double().cpu().numpy(), ctx.counters()["rays"]


def _other(frame):
    f = frame.copy()
    f["seed"] = frame["seed"] + np.float32(1.0)
    return f


def _images(ctx, torch, sid, wl, flags):
    n = wl.width * wl.height
    rgba = torch.zeros((n, 4), dtype=torch.uint8, device="cuda")
    rad = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    ctx.trace(sid, wl.frame, wl.params(flags=flags), rgba8=rgba, radiance=rad)
    torch.cuda.synchronize()
    return rgba.cpu().numpy(), rad.cpu().numpy()


@pytest.mark.parametrize("case", ["adjoint_bright", "soup_pooled", "many_emitters"])
def test_recording_forward_keeps_image_and_gradient(ctx, torch_cuda, case):
    """VKR_TRACE_RECORD_ADJOINT: the forward image is bit-identical with and
    without it, and the adjoint that follows runs only its second pass (about
    half the rays) on records equal to those its own first pass writes: the
    gradient matches the adjoint run on its own up to the order of the float
    additions that reduce it.  The adjoint scene runs with light scale 4, so
    that many samples saturate the per-bounce clamp (whose bits the records
    carry); the soup (several emitters) and the many-emitter scene run the
    pooled kernel."""
    torch = torch_cuda
    if case == "adjoint_bright":
        wl = scenes.adjoint_scene(width=128, height=96, spp=8)
        scale = 4.0
    elif case == "soup_pooled":
        wl = scenes.soup100k(n_tris=20_000, width=160, height=120, spp=2, max_depth=4)
        scale = 1.0
    else:
        wl = scenes.many_emitters()
        scale = 1.0
    sid = ctx.upload(wl.scene)
    theta = scenes.sphere_params(wl.scene.spheres, light_scale=scale)
    ctx.set_sphere_params(sid, theta)
    n = wl.width * wl.height
    gen = torch.Generator(device="cuda").manual_seed(11)
    g4 = (torch.rand((n, 4), device="cuda", generator=gen) - 0.5).contiguous()
    p = wl.params()
    # On its own: the previous call was for another frame, so both passes run.
    _adjoint(ctx, torch, sid, _other(wl.frame), p, g4, theta.size)
    g_own, rays_own = _adjoint(ctx, torch, sid, wl.frame, p, g4, theta.size)
    # After a recording forward: the images do not change, pass 1 is skipped.
    img0 = _images(ctx, torch, sid, wl, 0)
    img1 = _images(ctx, torch, sid, wl, vkr.TRACE_RECORD_ADJOINT)
    assert np.array_equal(img0[0], img1[0]) and np.array_equal(img0[1], img1[1])
    g_rec, rays_rec = _adjoint(ctx, torch, sid, wl.frame, p, g4, theta.size)
    # A second adjoint with the same inputs reuses its own records too.
    g_again, rays_again = _adjoint(ctx, torch, sid, wl.frame, p, g4, theta.size)
    agg = np.linalg.norm(g_rec - g_own) / np.linalg.norm(g_own)
    agg2 = np.linalg.norm(g_again - g_own) / np.linalg.norm(g_own)
    print(f"\n[recorded adjoint {case}] rays both passes {rays_own} second only {rays_rec}; "
          f"relative difference {agg:.3g} / {agg2:.3g}")
    assert np.linalg.norm(g_own) > 0
    assert agg <= 1e-5 and agg2 <= 1e-5
    assert rays_rec == rays_again and 0 < rays_rec < rays_own
    ctx.destroy_scene(sid)


def test_recording_goes_stale(ctx, torch_cuda):
    """The kept first pass serves only the same scene version, sphere
    parameters, frame data and trace parameters: any change runs both passes
    again (and gives that input's gradient); setting the same parameter
    values again keeps it."""
    torch = torch_cuda
    wl = scenes.adjoint_scene(width=96, height=64, spp=4)
    sid = ctx.upload(wl.scene)
    theta = scenes.sphere_params(wl.scene.spheres, light_scale=2.0)
    theta2 = theta.copy()
    theta2[7 * len(wl.scene.spheres)] = np.float32(3.0)
    n = wl.width * wl.height
    g4 = torch.ones((n, 4), dtype=torch.float32, device="cuda")
    p = wl.params()

    def with_record(params):
        q = vkr.TraceParams.from_buffer_copy(params)
        q.flags |= vkr.TRACE_RECORD_ADJOINT
        return q

    def record(th, frame=wl.frame, params=p):
        ctx.set_sphere_params(sid, th)
        ctx.trace(sid, frame, with_record(params),
                  radiance=torch.zeros((n, 4), dtype=torch.float32, device="cuda"))

    def own(th, frame=wl.frame, params=p):
        ctx.set_sphere_params(sid, th)
        _adjoint(ctx, torch, sid, _other(frame), params, g4, theta.size)
        return _adjoint(ctx, torch, sid, frame, params, g4, theta.size)

    def close(g, ref):
        return np.linalg.norm(g - ref) <= 1e-5 * np.linalg.norm(ref)

    g1, both1 = own(theta)
    g2, both2 = own(theta2)
    # Recorded at theta: the same values set again keep the records.
    record(theta)
    ctx.set_sphere_params(sid, theta.copy())
    g, r = _adjoint(ctx, torch, sid, wl.frame, p, g4, theta.size)
    assert r < both1 and close(g, g1)
    # New parameter values: both passes, theta2's gradient.
    record(theta)
    ctx.set_sphere_params(sid, theta2)
    g, r = _adjoint(ctx, torch, sid, wl.frame, p, g4, theta.size)
    assert r == both2 and close(g, g2)
    assert np.linalg.norm(g1 - g2) > 1e-3 * np.linalg.norm(g2)
    # Another frame seed, sample count or depth: both passes.
    for frame, params in ((_other(wl.frame), p),
                          (wl.frame, vkr.Context.params(wl.width, wl.height, 3, wl.max_depth)),
                          (wl.frame, vkr.Context.params(wl.width, wl.height, wl.spp, 3))):
        g_ref, both = own(theta, frame, params)
        record(theta)
        g, r = _adjoint(ctx, torch, sid, frame, params, g4, theta.size)
        assert r == both and close(g, g_ref)
    # A scene uploaded into a destroyed scene's id is another scene, even
    # with the same contents: recorded at the parameters an upload starts
    # with, the new scene's adjoint still runs both passes.
    theta0 = scenes.sphere_params(wl.scene.spheres, light_scale=1.0)
    g0, both0 = own(theta0)
    record(theta0)
    ctx.destroy_scene(sid)
    sid2 = ctx.upload(wl.scene)
    assert sid2 == sid
    g, r = _adjoint(ctx, torch, sid2, wl.frame, p, g4, theta.size)
    assert r == both0 and close(g, g0)
    ctx.destroy_scene(sid2)


def test_recording_flag_refusals(ctx, torch_cuda):
    torch = torch_cuda
    wl = scenes.adjoint_scene(width=64, height=64, spp=2)
    sid = ctx.upload(wl.scene)
    rad = torch.zeros((64 * 64, 4), dtype=torch.float32, device="cuda")
    dec = torch.zeros(64 * 64 * 2, dtype=torch.int16, device="cuda")
    R = vkr.TRACE_RECORD_ADJOINT
    bad = [
        lambda: ctx.trace(sid, wl.frame, wl.params(flags=R, kernel=vkr.KERNEL_WHITTED), radiance=rad),
        lambda: ctx.trace(sid, wl.frame, vkr.Context.params(64, 64, 2, 11, flags=R), radiance=rad),
        lambda: ctx.trace(sid, wl.frame, wl.params(flags=R | vkr.TRACE_RECORD_DECISIONS), radiance=rad,
                          decisions=dec),
        lambda: ctx.trace_static(sid, wl.frame, wl.params(flags=R), radiance=rad),
        lambda: ctx.trace_batch([ctx.job(sid, wl.frame, wl.params(flags=R), radiance=rad)]),
        lambda: ctx.batch_graph([ctx.job(sid, wl.frame, wl.params(flags=R), radiance=rad)]),
    ]
    for fn in bad:
        with pytest.raises(vkr.VkrError):
            fn()
    ctx.trace(sid, wl.frame, wl.params(flags=R), radiance=rad)
    torch.cuda.synchronize()
    ctx.destroy_scene(sid)


def test_adjoint_size_limit(ctx, torch_cuda):
    """The second pass hands out one item per (pixel, sample) from a 32-bit
    cursor: 4096^2 pixels at 256 spp (2^32 pairs) are refused before any
    allocation or launch; 4096^2 at 255 spp would pass the check."""
    torch = torch_cuda
    wl = scenes.adjoint_scene(width=8, height=8, spp=1)
    sid = ctx.upload(wl.scene)
    g4 = torch.zeros((64, 4), dtype=torch.float32, device="cuda")
    d = torch.zeros(7 * len(wl.scene.spheres) + 1, dtype=torch.float32, device="cuda")
    with pytest.raises(vkr.VkrError) as err:
        ctx.adjoint(sid, wl.frame, vkr.Context.params(4096, 4096, 256, 4), g4, d)
    assert err.value.code == 7 and "0xFF000000" in str(err.value)   # VKR_UNSUPPORTED
    ctx.destroy_scene(sid)


def test_adjoint_triangle_materials_match_finite_differences(ctx, torch_cuda):
    """Triangles with per-triangle materials (VKR_SCENE_TRI_MATERIAL_INDEX:
    diffuse, metal, dielectric and emissive entries) among the 33 spheres:
    their hits add next-event terms (with the reference's unnormalised
    triangle normal, so large ones that often saturate the clamp), emission
    and constant albedo factors to the throughput (and to Russian roulette's
    max3), none of them parameters.  The sphere gradients still match
    central differences of the HIP forward with frozen decisions.  The loss
    weights are positive: with random signs the quotients' rounding noise
    (float32 radiance) does not average out, and at this scene's dynamic
    range it reaches 1e-3 for the light scale."""
    torch = torch_cuda
    w, h, spp = 96, 64, 8
    wl0 = scenes.adjoint_scene(width=w, height=h, spp=spp)
    rng = np.random.default_rng(21)
    tris = scenes.random_soup(rng, 1500, lo=-12.0, hi=12.0, ehi=3.0)
    table = np.array([vkr.material((0.7, 0.5, 0.3), (0, 0, 0), 0.5, 0.0),
                      vkr.material((0.9, 0.9, 0.9), (0, 0, 0), 0.0, 1.0),
                      vkr.material((0.95, 0.95, 0.95), (0, 0, 0), 0.6, 0.0, vkr.DIELECTRIC),
                      vkr.material((0.4, 0.8, 0.6), (2.0, 1.5, 1.0), 0.3, 0.0)], vkr.MATERIAL)
    tris["mat_index"] = rng.integers(0, len(table), len(tris)).astype(np.uint32)
    sc = vkr.Scene(tris=tris, spheres=wl0.scene.spheres, planes=wl0.scene.planes, materials=table,
                   flags=vkr.SCENE_TRI_MATERIAL_INDEX)
    wl = scenes.Workload("adjoint_trimat", sc, wl0.frame, w, h, spp, wl0.max_depth)
    sid = ctx.upload(sc)
    theta = scenes.sphere_params(sc.spheres, light_scale=1.0)
    dec = torch.zeros(w * h * spp, dtype=torch.int16, device="cuda")
    base = _forward(ctx, torch, sid, wl, theta, vkr.TRACE_RECORD_DECISIONS, dec)
    weights = np.random.default_rng(4).uniform(0.25, 1.0, base.shape)
    g4 = torch.zeros((w * h, 4), dtype=torch.float32, device="cuda")
    g4[:, :3] = torch.from_numpy(weights.astype(np.float32)).cuda()
    d = torch.zeros(theta.size, dtype=torch.float32, device="cuda")
    ctx.set_sphere_params(sid, theta)
    ctx.adjoint(sid, wl.frame, wl.params(), g4, d)
    torch.cuda.synchronize()
    g = d.double().cpu().numpy()
    n_sph = len(sc.spheres)
    idx = [7 * s + k for s in range(n_sph - 1) for k in (0, 1, 2, 6)]
    idx += [7 * (n_sph - 1) + k for k in (3, 4, 5)] + [7 * n_sph]
    fd = np.array([_fd(ctx, torch, sid, wl, theta, dec, i, weights) for i in idx])
    ga = g[idx]
    scale = np.abs(fd).max()
    big = np.abs(fd) > 1e-3 * scale
    agg = np.linalg.norm(ga[big] - fd[big]) / np.linalg.norm(fd[big])
    rel = np.abs(ga - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    print(f"\n[adjoint triangle materials] {big.sum()} parameters, aggregate relative error {agg:.3g}, "
          f"{(rel[big] <= 1e-3).mean() * 100:.1f}% within 1e-3")
    assert big.sum() >= 20
    assert agg <= 1e-3
    assert (rel[big] <= 1e-3).mean() >= 0.9
    # The rest straddle kinks (a saturating clamp, max3 changing channel) at
    # h = 1e-3; a quarter of the step brings their quotient closer.
    for k in np.where(big & (rel > 1e-3))[0]:
        f4 = _fd(ctx, torch, sid, wl, theta, dec, idx[k], weights, h=2.5e-4)
        r4 = abs(ga[k] - f4) / abs(f4)
        print(f"  param {idx[k]}: rel {rel[k]:.3g} at h=1e-3 -> {r4:.3g} at h=2.5e-4")
        assert r4 < rel[k] and r4 <= 5e-3
    ctx.destroy_scene(sid)


def test_adjoint_dielectric_and_emissive_spheres_match_finite_differences(ctx, torch_cuda):
    """Spheres of all four kinds, two dielectrics among them also emitting
    (then they are next-event emitters too, and their emission enters the
    path unnormalised, Tracer.comp:515): albedos, the emitters' emission,
    the light scale and the roughness of diffuse and metal spheres against
    central differences of the HIP forward with frozen decisions (positive
    loss weights).  A dielectric's roughness holds its IOR, which moves the
    refracted ray and is not differentiated; emission parameters of spheres
    that do not emit would change the emitter set, a jump."""
    torch = torch_cuda
    w, h, spp = 96, 64, 8
    rng = np.random.default_rng(33)
    sph = scenes.random_spheres(rng, 24, -10.0, 10.0)
    diel = [i for i in range(len(sph)) if sph[i]["mat"]["type"] == vkr.DIELECTRIC]
    assert len(diel) >= 3
    # Two of them emit, placed in front of the camera (at z = -30, looking
    # along +z) so that many paths pass through them.
    for i, c in zip(diel[:2], ((-3.0, 0.0, -8.0), (4.0, -2.0, -4.0))):
        sph[i]["mat"]["emissive"] = np.array([3.0, 2.0, 1.5], np.float32)
        sph[i]["P"] = np.array(c, np.float32)
        sph[i]["r"] = np.float32(2.0)
    sph = np.concatenate([sph, scenes.light_sphere((0.0, 30.0, 0.0), 6.0)])
    wl0 = scenes.adjoint_scene(width=w, height=h, spp=spp)
    sc = vkr.Scene(spheres=sph, planes=wl0.scene.planes)
    wl = scenes.Workload("adjoint_kinds", sc, wl0.frame, w, h, spp, wl0.max_depth)
    sid = ctx.upload(sc)
    theta = scenes.sphere_params(sc.spheres, light_scale=1.0)
    dec = torch.zeros(w * h * spp, dtype=torch.int16, device="cuda")
    base = _forward(ctx, torch, sid, wl, theta, vkr.TRACE_RECORD_DECISIONS, dec)
    weights = np.random.default_rng(6).uniform(0.25, 1.0, base.shape)
    g4 = torch.zeros((w * h, 4), dtype=torch.float32, device="cuda")
    g4[:, :3] = torch.from_numpy(weights.astype(np.float32)).cuda()
    d = torch.zeros(theta.size, dtype=torch.float32, device="cuda")
    ctx.set_sphere_params(sid, theta)
    ctx.adjoint(sid, wl.frame, wl.params(), g4, d)
    torch.cuda.synchronize()
    g = d.double().cpu().numpy()
    n_sph = len(sph)
    emitting = [s for s in range(n_sph) if np.any(sph[s]["mat"]["emissive"] != 0)]
    idx = [7 * s + k for s in range(n_sph) for k in (0, 1, 2)]
    idx += [7 * s + 6 for s in range(n_sph) if sph[s]["mat"]["type"] != vkr.DIELECTRIC]
    idx += [7 * s + k for s in emitting for k in (3, 4, 5)] + [7 * n_sph]
    fd = np.array([_fd(ctx, torch, sid, wl, theta, dec, i, weights) for i in idx])
    ga = g[idx]
    scale = np.abs(fd).max()
    big = np.abs(fd) > 1e-3 * scale
    agg = np.linalg.norm(ga[big] - fd[big]) / np.linalg.norm(fd[big])
    rel = np.abs(ga - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    print(f"\n[adjoint sphere kinds] {len(emitting)} emitters, {big.sum()} parameters, aggregate relative error "
          f"{agg:.3g}, {(rel[big] <= 1e-3).mean() * 100:.1f}% within 1e-3")
    for k in np.argsort(-rel)[:4]:
        print(f"  param {idx[k]:4d} adjoint {ga[k]: .6e} fd {fd[k]: .6e} rel {rel[k]:.3g}")
    assert big.sum() >= 30
    # the emitting dielectrics' own emission gradients are in the check
    for s in diel[:2]:
        print(f"  emitting dielectric {s}: emission fd {[round(float(fd[idx.index(7 * s + k)]), 5) for k in (3, 4, 5)]}")
        assert all(big[idx.index(7 * s + k)] for k in (3, 4, 5))
    assert agg <= 1e-3
    assert (rel[big] <= 1e-3).mean() >= 0.9
    for k in np.where(big & (rel > 1e-3))[0]:
        f4 = _fd(ctx, torch, sid, wl, theta, dec, idx[k], weights, h=2.5e-4)
        r4 = abs(ga[k] - f4) / abs(f4)
        print(f"  param {idx[k]}: rel {rel[k]:.3g} at h=1e-3 -> {r4:.3g} at h=2.5e-4")
        assert r4 < rel[k] and r4 <= 5e-3
    ctx.destroy_scene(sid)


def test_adjoint_depth8_matches_finite_differences(ctx, torch_cuda):
    """Eight bounces, the deepest for which decisions can be frozen
    (recorded for depths 0-7): the saturation bits and the throughput chain
    of long paths against central differences of the HIP forward (positive
    loss weights).  The spheres' albedos are raised and the scene is closed
    by planes, so that paths reach the last bounces."""
    torch = torch_cuda
    w, h, spp = 96, 64, 8
    wl0 = scenes.adjoint_scene(width=w, height=h, spp=spp, max_depth=8)
    sph = wl0.scene.spheres.copy()
    for i in range(len(sph) - 1):
        sph[i]["mat"]["albedo"] = np.float32(0.5) + np.float32(0.45) * sph[i]["mat"]["albedo"]
    # A closed room of bright diffuse planes (floor, ceiling, four walls;
    # dot(X, N) + len = 0) keeps paths inside it.
    room = [((0, 1, 0), 11.0), ((0, -1, 0), 40.0), ((1, 0, 0), 25.0), ((-1, 0, 0), 25.0),
            ((0, 0, -1), 25.0), ((0, 0, 1), 35.0)]
    pln = scenes.planes_from([vkr.material((0.85, 0.8, 0.75), (0, 0, 0), 0.5, 0.0) for _ in room],
                             [n for n, _ in room], np.array([l for _, l in room], np.float32))
    sc = vkr.Scene(spheres=sph, planes=pln)
    wl = scenes.Workload("adjoint_depth8", sc, wl0.frame, w, h, spp, 8)
    sid = ctx.upload(sc)
    theta = scenes.sphere_params(sc.spheres, light_scale=1.0)
    dec = torch.zeros(w * h * spp, dtype=torch.int16, device="cuda")
    base = _forward(ctx, torch, sid, wl, theta, vkr.TRACE_RECORD_DECISIONS, dec)
    # samples that survive Russian roulette at bounce k (decision bit 2k)
    bits = dec.cpu().numpy().view(np.uint16)
    surv = [int((bits >> (2 * k) & 1).sum()) for k in range(8)]
    late = surv[5]
    weights = np.random.default_rng(8).uniform(0.25, 1.0, base.shape)
    g4 = torch.zeros((w * h, 4), dtype=torch.float32, device="cuda")
    g4[:, :3] = torch.from_numpy(weights.astype(np.float32)).cuda()
    d = torch.zeros(theta.size, dtype=torch.float32, device="cuda")
    ctx.set_sphere_params(sid, theta)
    ctx.adjoint(sid, wl.frame, wl.params(), g4, d)
    torch.cuda.synchronize()
    g = d.double().cpu().numpy()
    n_sph = len(sph)
    idx = [7 * s + k for s in range(n_sph - 1) for k in (0, 1, 2, 6)]
    idx += [7 * (n_sph - 1) + k for k in (3, 4, 5)] + [7 * n_sph]
    fd = np.array([_fd(ctx, torch, sid, wl, theta, dec, i, weights) for i in idx])
    ga = g[idx]
    scale = np.abs(fd).max()
    big = np.abs(fd) > 1e-3 * scale
    agg = np.linalg.norm(ga[big] - fd[big]) / np.linalg.norm(fd[big])
    rel = np.abs(ga - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    print(f"\n[adjoint depth 8] samples surviving roulette per bounce {surv}, {big.sum()} parameters, "
          f"aggregate relative error {agg:.3g}, {(rel[big] <= 1e-3).mean() * 100:.1f}% within 1e-3")
    assert late > 500
    assert big.sum() >= 30
    assert agg <= 1e-3
    # Long paths straddle more clamp saturations at h = 1e-3 (90% of the 41
    # parameters within 1e-3 here); each outlier must close in at h / 4.
    assert (rel[big] <= 1e-3).mean() >= 0.85
    for k in np.where(big & (rel > 1e-3))[0]:
        f4 = _fd(ctx, torch, sid, wl, theta, dec, idx[k], weights, h=2.5e-4)
        r4 = abs(ga[k] - f4) / abs(f4)
        print(f"  param {idx[k]}: rel {rel[k]:.3g} at h=1e-3 -> {r4:.3g} at h=2.5e-4")
        assert r4 < rel[k] and r4 <= 5e-3
    ctx.destroy_scene(sid)


@pytest.mark.parametrize("seed", range(16))
def test_recording_on_random_scenes(ctx, torch_cuda, seed):
    """VKR_TRACE_RECORD_ADJOINT over the random scenes of the parity sweep
    (all four material kinds, per-triangle material tables, planes, odd
    sizes, 1-6 bounces, several epsilons): every forward output is
    bit-identical with and without the flag; the adjoint after it traces
    exactly half the rays of the adjoint on its own (pass 2 replays pass 1's
    rays) and gives the same gradient up to the order of its float atomics."""
    import test_gpu_sweep
    torch = torch_cuda
    wl, eps = test_gpu_sweep._scene(seed)
    sid = ctx.upload(wl.scene)
    n = wl.width * wl.height
    p = vkr.Context.params(wl.width, wl.height, wl.spp, wl.max_depth, epsilon=eps)
    rec = vkr.Context.params(wl.width, wl.height, wl.spp, wl.max_depth, epsilon=eps,
                             flags=vkr.TRACE_RECORD_ADJOINT)
    outs = []
    for params in (p, rec):
        o = {"rgba8": torch.zeros((n, 4), dtype=torch.uint8, device="cuda"),
             "radiance": torch.zeros((n, 4), dtype=torch.float32, device="cuda"),
             "prim_id": torch.zeros(n, dtype=torch.int32, device="cuda"),
             "hit_t": torch.zeros(n, dtype=torch.float32, device="cuda"),
             "ray_count": torch.zeros(n, dtype=torch.int32, device="cuda")}
        ctx.trace(sid, wl.frame, params, **o)
        torch.cuda.synchronize()
        outs.append({k: v.cpu().numpy().view(np.uint8) for k, v in o.items()})
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
    n_par = 7 * len(wl.scene.spheres) + 1
    gen = torch.Generator(device="cuda").manual_seed(seed)
    g4 = (torch.rand((n, 4), device="cuda", generator=gen) - 0.5).contiguous()
    _adjoint(ctx, torch, sid, _other(wl.frame), p, g4, n_par)
    g_own, rays_own = _adjoint(ctx, torch, sid, wl.frame, p, g4, n_par)
    ctx.trace(sid, wl.frame, rec, radiance=torch.zeros((n, 4), dtype=torch.float32, device="cuda"))
    g_rec, rays_rec = _adjoint(ctx, torch, sid, wl.frame, p, g4, n_par)
    diff = np.linalg.norm(g_rec - g_own)
    print(f"\n[recording sweep {seed}] {wl.width}x{wl.height} spp {wl.spp} depth {wl.max_depth}: rays {rays_own} "
          f"-> {rays_rec}, |g| {np.linalg.norm(g_own):.4g}, difference {diff:.3g}")
    assert rays_own == 2 * rays_rec
    assert diff <= 1e-5 * np.linalg.norm(g_own)
    ctx.destroy_scene(sid)


@pytest.mark.parametrize("seed", [1, 4, 6, 9, 12, 15])
def test_adjoint_random_scenes_match_finite_differences(ctx, torch_cuda, seed):
    """The adjoint on random scenes of the parity sweep (every material kind
    on spheres, planes and per-triangle tables, odd sizes, 2-6 bounces,
    several epsilons) against central differences of the HIP forward with
    frozen decisions and positive loss weights.  Parameters: albedos, the
    emitters' emission, the light scale, and the roughness of spheres that
    are not dielectric and have a roughness of at least 0.25 (a
    dielectric's roughness holds its IOR, which moves the refracted ray and
    is not differentiated; emission parameters of spheres that do not emit
    would change the emitter set).  Below 0.25 central differences cannot
    check a roughness: GGX's denominator (Tracer.comp:273-282) takes
    a^4 - 1 in float32, quantised at 6e-8 absolute against a^4 of 4e-5 to
    4e-3, the same for every sample of the sphere, so the forward is a
    staircase in roughness at that scale (the quotient of a single metal
    sphere at roughness 0.08 reads -5900, -4747 and -255 at h = 1e-3,
    2.5e-4 and 6.25e-5, at 0.3 it agrees with the adjoint to 1e-5)."""
    import test_gpu_sweep
    torch = torch_cuda
    wl0, eps = test_gpu_sweep._scene(seed)
    wl = scenes.Workload(wl0.name, wl0.scene, wl0.frame, wl0.width, wl0.height, wl0.spp, wl0.max_depth)
    sid = ctx.upload(wl.scene)
    sph = wl.scene.spheres
    theta = scenes.sphere_params(sph, light_scale=1.0)
    n = wl.width * wl.height
    dec = torch.zeros(n * wl.spp, dtype=torch.int16, device="cuda")

    def fwd(th, flags):
        ctx.set_sphere_params(sid, th)
        rad = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
        params = vkr.Context.params(wl.width, wl.height, wl.spp, wl.max_depth, epsilon=eps, flags=flags)
        ctx.trace(sid, wl.frame, params, radiance=rad, decisions=dec)
        torch.cuda.synchronize()
        return rad[:, :3].double().cpu().numpy()

    base = fwd(theta, vkr.TRACE_RECORD_DECISIONS)
    weights = np.random.default_rng(100 + seed).uniform(0.25, 1.0, base.shape)
    g4 = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    g4[:, :3] = torch.from_numpy(weights.astype(np.float32)).cuda()
    d = torch.zeros(theta.size, dtype=torch.float32, device="cuda")
    ctx.set_sphere_params(sid, theta)
    ctx.adjoint(sid, wl.frame, vkr.Context.params(wl.width, wl.height, wl.spp, wl.max_depth, epsilon=eps), g4, d)
    torch.cuda.synchronize()
    g = d.double().cpu().numpy()
    n_sph = len(sph)
    emitting = [s for s in range(n_sph) if np.any(sph[s]["mat"]["emissive"] != 0)]
    idx = [7 * s + k for s in range(n_sph) for k in (0, 1, 2)]
    idx += [7 * s + 6 for s in range(n_sph)
            if sph[s]["mat"]["type"] != vkr.DIELECTRIC and sph[s]["mat"]["roughness"] >= 0.25]
    idx += [7 * s + k for s in emitting for k in (3, 4, 5)] + [7 * n_sph]

    def quotient(i, h):
        h = h * max(1.0, abs(float(theta[i])))
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        return float(np.sum(weights * (fwd(tp, vkr.TRACE_REPLAY_DECISIONS) - fwd(tm, vkr.TRACE_REPLAY_DECISIONS)))
                     / (float(tp[i]) - float(tm[i])))

    fd = np.array([quotient(i, 1e-3) for i in idx])
    ga = g[idx]
    scale = np.abs(fd).max()
    big = np.abs(fd) > 1e-3 * scale
    agg = np.linalg.norm(ga[big] - fd[big]) / max(np.linalg.norm(fd[big]), 1e-30)
    rel = np.abs(ga - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    print(f"\n[adjoint sweep {seed}] {wl.width}x{wl.height} spp {wl.spp} depth {wl.max_depth}, {len(emitting)} "
          f"emitters, {big.sum()} parameters, aggregate relative error {agg:.3g}, "
          f"{(rel[big] <= 1e-3).mean() * 100:.1f}% within 1e-3")
    # These scenes saturate the clamp often (emissive planes and triangles,
    # several emitters), so more quotients straddle its kinks at h = 1e-3.
    assert scale > 0 and big.sum() >= 5
    assert agg <= 2e-3
    assert (rel[big] <= 1e-3).mean() >= 0.85
    # Each outlier again at h / 4 and h / 16: a straddled kink's error
    # shrinks with h, while a quotient limited by float32 rounding (the
    # emission of a bright emitter scatters by about 0.15% over h = 8e-3 to
    # 1.25e-4) does not; one of the two must agree within 0.5%.
    for k in np.where(big & (rel > 1e-3))[0]:
        f4, f16 = quotient(idx[k], 2.5e-4), quotient(idx[k], 6.25e-5)
        r4, r16 = abs(ga[k] - f4) / abs(f4), abs(ga[k] - f16) / abs(f16)
        print(f"  param {idx[k]}: rel {rel[k]:.3g} at h=1e-3, {r4:.3g} at h=2.5e-4, {r16:.3g} at h=6.25e-5")
        assert min(r4, r16) <= 5e-3
    ctx.destroy_scene(sid)


def test_diff_render_backward_runs_pass_two_only(ctx, torch_cuda):
    """diff.render's forward passes VKR_TRACE_RECORD_ADJOINT when theta needs
    a gradient, and its backward (which sets the same sphere parameters
    again) then runs only the adjoint's second pass: half the rays of the
    adjoint on its own, and the same gradient.  Without a gradient the
    forward records nothing."""
    torch = torch_cuda
    import diff
    wl = scenes.adjoint_scene(width=96, height=64, spp=4)
    sid = ctx.upload(wl.scene)
    theta = scenes.sphere_params(wl.scene.spheres, light_scale=1.5)
    n = wl.width * wl.height
    gen = torch.Generator(device="cuda").manual_seed(2)
    w = torch.rand((n, 3), device="cuda", generator=gen)
    # the adjoint on its own (the previous call was for another frame)
    g4 = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    g4[:, :3] = w
    ctx.set_sphere_params(sid, theta)
    _adjoint(ctx, torch, sid, _other(wl.frame), wl.params(), g4.contiguous(), theta.size)
    g_own, rays_own = _adjoint(ctx, torch, sid, wl.frame, wl.params(), g4.contiguous(), theta.size)
    # forward + backward through torch.autograd
    t = torch.tensor(theta, device="cuda", requires_grad=True)
    rad = diff.render(t, ctx, sid, wl.frame, wl.params())
    (rad * w).sum().backward()
    torch.cuda.synchronize()
    rays_bwd = ctx.counters()["rays"]
    g = t.grad.double().cpu().numpy()
    print(f"\n[diff.render] rays: adjoint on its own {rays_own}, backward {rays_bwd}")
    assert rays_own == 2 * rays_bwd
    assert np.linalg.norm(g - g_own) <= 1e-5 * np.linalg.norm(g_own)
    # no gradient wanted: no recording, so the next adjoint runs both passes
    _adjoint(ctx, torch, sid, _other(wl.frame), wl.params(), g4.contiguous(), theta.size)
    with torch.no_grad():
        diff.render(torch.tensor(theta, device="cuda"), ctx, sid, wl.frame, wl.params())
    _, rays = _adjoint(ctx, torch, sid, wl.frame, wl.params(), g4.contiguous(), theta.size)
    assert rays == rays_own
    ctx.destroy_scene(sid)
