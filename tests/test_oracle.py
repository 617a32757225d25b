"""Pins the CPU oracle (oracle/vkr_oracle.c) without the reference runtime.

The reference has no tests or golden vectors and its shaders cannot run here
(no Vulkan, no GLSL compiler), so each oracle function is checked against
what the mathematics fixes: closed forms evaluated in float64, invariants,
textbook identities and the reference's own constants (tests/golden).
Every check is chosen so that a plausible transcription mistake (a dropped
term, a swapped operand, a wrong sign or order) fails it.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_ffi
import vkr

HERE = os.path.dirname(os.path.abspath(__file__))
f3 = oracle_ffi.f3
F32 = np.float32


@pytest.fixture(scope="module")
def lib():
    return oracle_ffi.load()


def out3():
    return (oracle_ffi.C.c_float * 3)()


def arr(o):
    return np.frombuffer(o, np.float32).copy()


# --------------------------------------------------------------------------
# Constants (Assets/Tracer.comp:176-234, folded the way glslang folds)
# --------------------------------------------------------------------------
def test_reference_constants_fixture():
    g = json.load(open(os.path.join(HERE, "golden", "reference_constants.json")))
    bits = lambda x: int(np.array([x], np.float32).view(np.uint32)[0])  # noqa: E731
    for name, spec in g["float32_bits"].items():
        assert bits(spec["literal"]) == int(spec["bits"], 16), name
    c = {k: v["literal"] for k, v in g["float32_bits"].items()}
    # glslang folds 1.0 / 2.2 in double: one ulp above the float quotient.
    assert bits(c["inv_gamma"]) == 0x3EE8BA2F
    assert bits(np.float32(1.0) / np.float32(2.2)) == 0x3EE8BA2E
    assert np.float32(c["two_pi"]) == np.float32(2.0) * np.float32(c["PI"])
    assert bits(c["att_div"]) == 0x3F7FBE77


# --------------------------------------------------------------------------
# RNG (Tracer.comp:221-234)
# --------------------------------------------------------------------------
def _rand_numpy(cx, cy, salt0, n):
    """The GLSL expression evaluated with numpy float32 ops in GLSL order and
    float64 cos rounded to float32."""
    cx, cy, s = F32(cx), F32(cy), F32(salt0)
    out = []
    for _ in range(n):
        s = F32(s + F32(1.0))                               # ++rand_salt
        a = F32(F32(cx + s) * F32(5.0))
        b = F32(F32(cy + s) * F32(5.0))
        d = F32(F32(a * F32(23.14069263277926)) + F32(b * F32(2.665144142690225)))
        c = F32(math.cos(float(d)))
        v = F32(c * F32(12345.6789))
        out.append(F32(v - F32(math.floor(float(v)))))      # fract
    return np.array(out, np.float32)


def test_rand_matches_glsl_expression(lib):
    rng = np.random.default_rng(0)
    buf = (oracle_ffi.C.c_float * 6)()
    agree = total = 0
    for _ in range(3000):
        cx, cy = rng.random(2).astype(np.float32)
        salt0 = np.float32(rng.random() * 3000)
        lib.vkro_rand_sequence(float(cx), float(cy), float(salt0), 6, buf)
        got = np.frombuffer(buf, np.float32)
        exp = _rand_numpy(cx, cy, salt0, 6)
        agree += int((got.view(np.uint32) == exp.view(np.uint32)).sum())
        total += 6
    assert agree / total >= 0.9999


def test_rand_stream_statistics_and_shift(lib):
    n = 200_000
    buf = (oracle_ffi.C.c_float * n)()
    lib.vkro_rand_sequence(0.3125, 0.6875, 17.0, n, buf)
    x = np.frombuffer(buf, np.float32).astype(np.float64)
    assert np.all((x >= 0.0) & (x <= 1.0))
    assert abs(x.mean() - 0.5) < 0.005
    assert abs(x.var() - 1.0 / 12.0) < 0.003
    assert abs(np.corrcoef(x[:-1], x[1:])[0, 1]) < 0.01
    # Counter-based: starting one step later is the same stream shifted.
    buf2 = (oracle_ffi.C.c_float * 100)()
    lib.vkro_rand_sequence(0.3125, 0.6875, 18.0, 100, buf2)
    assert np.array_equal(np.frombuffer(buf2, np.float32), np.frombuffer(buf, np.float32)[1:101])


def test_rand_x0_column_is_nan(lib):
    """uv.x = 0 gives salt = inf (NaN at the origin): every rand() is NaN."""
    buf = (oracle_ffi.C.c_float * 4)()
    lib.vkro_rand_sequence(0.0, 0.5, float("inf"), 4, buf)
    assert np.all(np.isnan(np.frombuffer(buf, np.float32)))


def test_cos_sin_pow_are_nearest_floats(lib):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-10, 10, 2000), rng.uniform(1e3, 4e5, 2000)]).astype(np.float32)
    for x in xs:
        assert lib.vkro_cos(float(x)) == np.float32(math.cos(float(x)))
        assert lib.vkro_sin(float(x)) == np.float32(math.sin(float(x)))
    for x in rng.random(500).astype(np.float32):
        assert lib.vkro_pow(float(x), float(np.float32(1 / 2.2))) == np.float32(float(x) ** float(np.float32(1 / 2.2)))
    assert math.isnan(lib.vkro_pow(-0.5, 2.0))


# --------------------------------------------------------------------------
# Intersections (Tracer.comp:314-372) against float64 closed forms
# --------------------------------------------------------------------------
def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_sphere_closed_form(lib):
    rng = np.random.default_rng(2)
    s = np.zeros(1, vkr.SPHERE)
    for _ in range(3000):
        o = rng.uniform(-20, 20, 3).astype(np.float32)
        c = rng.uniform(-5, 5, 3).astype(np.float32)
        r = np.float32(rng.uniform(0.5, 4))
        d = _unit(c - o + rng.normal(0, 2.0, 3)).astype(np.float32)
        s["P"], s["r"] = c, r
        t = lib.vkro_sphere_intersect(f3(o), f3(d), s.ctypes.data)
        oc = o.astype(np.float64) - c
        dd = d.astype(np.float64)
        b = 2 * oc @ dd
        cc = oc @ oc - float(r) ** 2
        h = b * b - 4 * cc
        if h < -1e-3:
            assert t == -1.0
        elif h > 1e-2:
            exact = (-b - math.sqrt(h)) / 2
            assert abs(t - exact) <= 1e-5 * max(1.0, abs(exact)) + 2e-6 * np.linalg.norm(oc) ** 2 / math.sqrt(h)
    # From inside the sphere the near root is behind the origin (refracted
    # rays leave glass spheres undeflected, Tracer.comp:326).
    s["P"], s["r"] = (0, 0, 0), 2.0
    assert lib.vkro_sphere_intersect(f3((0, 0, 0)), f3((1, 0, 0)), s.ctypes.data) == -2.0


def test_plane_closed_form(lib):
    p = np.zeros(1, vkr.PLANE)
    p["N"], p["len"] = (0, 1, 0), 11.0          # y = -11
    t = lib.vkro_plane_intersect(f3((0, 0, 0)), f3(_unit((0, -1, 1))), p.ctypes.data)
    assert abs(t - 11 * math.sqrt(2)) < 1e-5
    # behind the origin -> max(dist, 0) = 0
    assert lib.vkro_plane_intersect(f3((0, 0, 0)), f3((0, 1, 0)), p.ctypes.data) == 0.0
    # parallel: when_neq(0, 0) = 0 times max(-(len + o.N) / 0, 0), i.e. 0 * 0 = 0
    # above the plane and 0 * inf = NaN below it (both rejected by t > eps)
    assert lib.vkro_plane_intersect(f3((0, 0, 0)), f3((1, 0, 0)), p.ctypes.data) == 0.0
    assert math.isnan(lib.vkro_plane_intersect(f3((0, -20, 0)), f3((1, 0, 0)), p.ctypes.data))


def test_triangle_barycentric_solve(lib):
    rng = np.random.default_rng(3)
    tri = np.zeros(1, vkr.TRIANGLE)
    hits = 0
    for _ in range(4000):
        v0 = rng.uniform(-2, 2, 3).astype(np.float32)
        v1 = (v0 + rng.uniform(-2, 2, 3)).astype(np.float32)
        v2 = (v0 + rng.uniform(-2, 2, 3)).astype(np.float32)
        o = rng.uniform(-8, 8, 3).astype(np.float32)
        target = v0 + rng.uniform(-0.2, 0.8) * (v1 - v0) + rng.uniform(-0.2, 0.8) * (v2 - v0)
        d = _unit(target - o).astype(np.float32)
        tri["v0"], tri["v1"], tri["v2"] = v0, v1, v2
        t = lib.vkro_tri_intersect(f3(o), f3(d), tri.ctypes.data, 1e-3)
        e1, e2 = (v1 - v0).astype(np.float64), (v2 - v0).astype(np.float64)
        D = d.astype(np.float64)
        det = e1 @ np.cross(D, e2)
        A = np.stack([-D, e1, e2], axis=1)
        if abs(np.linalg.det(A)) < 1e-9:
            continue
        tt, u, v = np.linalg.solve(A, o.astype(np.float64) - v0)
        margin = 1e-4
        inside = u > margin and v > margin and u + v < 1 - margin
        outside = u < -margin or v < -margin or u + v > 1 + margin
        if det < 1e-3 - 1e-6:
            assert t == -1.0                      # one-sided, absolute threshold (:348)
        elif det > 1e-3 + 1e-6 and inside:
            hits += 1
            # float Moller-Trumbore loses ~|e1||e2|/det ulps on grazing rays
            cond = np.linalg.norm(e1) * np.linalg.norm(e2) / det
            assert abs(t - tt) <= (1e-5 + 2e-7 * cond) * max(1.0, abs(tt)) * 4, (t, tt, cond)
        elif det > 1e-3 + 1e-6 and outside:
            assert t == -1.0
    assert hits > 300


# --------------------------------------------------------------------------
# trace_ray's order-dependent acceptance (Tracer.comp:374-431)
# --------------------------------------------------------------------------
def _facing_tri(z):
    """A triangle in the plane z = const facing a ray along -z... (+z travel)."""
    t = np.zeros(1, vkr.TRIANGLE)
    t["v0"], t["v1"], t["v2"] = (-1, -1, z), (-1, 3, z), (3, -1, z)
    return t


def _scene(tris=None, spheres=None, planes=None):
    return vkr.Scene(tris=tris if tris is not None else np.zeros(0, vkr.TRIANGLE),
                     spheres=spheres if spheres is not None else np.zeros(0, vkr.SPHERE),
                     planes=planes if planes is not None else np.zeros(0, vkr.PLANE))


def _closest(lib, scene, o, d, tmax=3000.0):
    sc = oracle_ffi.OracleScene(scene)
    pid = oracle_ffi.C.c_uint32()
    t = oracle_ffi.C.c_float()
    hit = lib.vkro_trace_ray(oracle_ffi.C.byref(sc.c), f3(o), f3(d), tmax, 1e-3, oracle_ffi.C.byref(pid),
                             oracle_ffi.C.byref(t))
    return hit, pid.value, t.value


def test_acceptance_rules(lib):
    o, d = (0.2, 0.2, 0.0), (0.0, 0.0, 1.0)
    # Later triangle up to EPSILON farther replaces the earlier (t < best + eps).
    tr = np.concatenate([_facing_tri(5.0), _facing_tri(5.0005)])
    hit, pid, t = _closest(lib, _scene(tris=tr), o, d)
    assert hit and pid == 1 and abs(t - 5.0005) < 1e-6
    # ... but not one more than EPSILON farther.
    tr = np.concatenate([_facing_tri(5.0), _facing_tri(5.002)])
    assert _closest(lib, _scene(tris=tr), o, d)[1] == 0
    # A chain creeps: 5.0 -> 5.0009 -> 5.0018 all accepted in order.
    tr = np.concatenate([_facing_tri(5.0), _facing_tri(5.0009), _facing_tri(5.0018)])
    assert _closest(lib, _scene(tris=tr), o, d)[1] == 2
    # The nearest hit always wins when it comes last.
    tr = np.concatenate([_facing_tri(9.0), _facing_tri(5.0)])
    assert _closest(lib, _scene(tris=tr), o, d)[1] == 1
    # Hits in [t_max, t_max + eps) are accepted (best starts at t_max).
    assert _closest(lib, _scene(tris=_facing_tri(5.0005)), o, d, tmax=5.0)[0] == 1
    # Planes need t < best - eps: a plane 0.5*eps nearer than the triangle loses.
    pl = np.zeros(1, vkr.PLANE)
    pl["N"], pl["len"] = (0, 0, -1), 4.9995      # z = 4.9995
    hit, pid, t = _closest(lib, _scene(tris=_facing_tri(5.0), planes=pl), o, d)
    assert pid == 0
    pl["len"] = 4.99                              # 10 * eps nearer: the plane wins
    assert _closest(lib, _scene(tris=_facing_tri(5.0), planes=pl), o, d)[1] == (2 << 30)
    # Triangles are one-sided (a ray from behind misses).
    assert _closest(lib, _scene(tris=_facing_tri(-5.0)), (0.2, 0.2, 0.0), (0.0, 0.0, -1.0))[0] == 0


# --------------------------------------------------------------------------
# Shading (Tracer.comp:256-312)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("rough", [0.3, 0.5, 0.8])
def test_ggx_is_normalised(lib, rough):
    """Textbook identity: integral of D(h) (n.h) over the hemisphere = 1."""
    n = 20000
    th = (np.arange(n) + 0.5) * (math.pi / 2) / n
    N = f3((0, 0, 1))
    D = np.array([lib.vkro_distribution_ggx(N, f3((math.sin(x), 0.0, math.cos(x))), rough) for x in th])
    integral = 2 * math.pi * np.sum(D * np.cos(th) * np.sin(th)) * (math.pi / 2) / n
    assert abs(integral - 1.0) < 2e-3


def test_smith_and_fresnel(lib):
    N = f3((0, 0, 1))
    for r in (0.0, 0.4, 1.0):
        k = (r + 1) ** 2 / 8
        for cv, cl in ((1.0, 1.0), (0.5, 0.25), (0.1, 0.9)):
            V = f3((math.sqrt(1 - cv * cv), 0, cv))
            L = f3((0, math.sqrt(1 - cl * cl), cl))
            g = lib.vkro_geometry_smith(N, V, L, r)
            exp = (cv / (cv * (1 - k) + k)) * (cl / (cl * (1 - k) + k))
            assert abs(g - exp) < 1e-6 and 0.0 <= g <= 1.0 + 1e-7
    o = out3()
    lib.vkro_fresnel_schlick(1.0, f3((0.04, 0.5, 1.0)), o)
    assert np.array_equal(arr(o), np.float32([0.04, 0.5, 1.0]))
    lib.vkro_fresnel_schlick(0.0, f3((0.04, 0.5, 1.0)), o)
    assert np.allclose(arr(o), 1.0, atol=1e-7)
    lib.vkro_fresnel_schlick(0.5, f3((0.04, 0.04, 0.04)), o)
    assert abs(arr(o)[0] - (0.04 + 0.96 * 0.5 ** 5)) < 1e-7
    r0 = ((1 - 0.42) / 1.42) ** 2
    assert abs(lib.vkro_schlick(1.0, 0.42) - r0) < 1e-7 and abs(r0 - 0.166832) < 1e-6
    assert abs(lib.vkro_schlick(0.0, 0.42) - 1.0) < 1e-7


def test_reflect_refract_snell(lib):
    o = out3()
    I = _unit((1, -1, 0))
    lib.vkro_reflect(f3(I), f3((0, 1, 0)), o)
    assert np.allclose(arr(o), _unit((1, 1, 0)), atol=1e-7)
    for eta in (0.42, 0.9, 1.3):
        for ang in (0.1, 0.6, 1.2):
            I = np.array([math.sin(ang), -math.cos(ang), 0.0])
            lib.vkro_refract(f3(I), f3((0, 1, 0)), eta, o)
            T = arr(o)
            st = eta * math.sin(ang)
            if st > 1.0:
                assert np.array_equal(T, np.zeros(3, np.float32))      # total internal reflection
            else:
                assert abs(T[0] - st) < 1e-6 and abs(T[1] + math.sqrt(1 - st * st)) < 1e-6


def test_jitter_frame(lib):
    rng = np.random.default_rng(4)
    o = out3()
    for _ in range(500):
        d = rng.normal(size=3) * rng.uniform(0.1, 10)
        ca = rng.uniform(0, 1)
        sa = math.sqrt(1 - ca * ca)
        phi = rng.uniform(0, 2 * math.pi)
        lib.vkro_jitter(f3(d), phi, sa, ca, o)
        v = arr(o).astype(np.float64)
        assert abs(np.linalg.norm(v) - 1) < 1e-5
        assert abs(v @ _unit(d) - ca) < 1e-5


# --------------------------------------------------------------------------
# Camera::update (Source/Camera.cpp:48-64)
# --------------------------------------------------------------------------
def test_camera_update_closed_form(lib):
    cam = np.zeros(1, vkr.CAMERA)
    for pos, pitch, yaw in (((32.8509, 30.6991, -106.389), 4.44998, -602.79), ((0, 0, -30), 0.0, 90.0),
                            ((1, 2, 3), 95.0, 10.0)):
        lib.vkro_camera_update(f3(pos), pitch, yaw, cam.ctypes.data)
        p = math.radians(max(-89.0, min(89.0, pitch)))
        y = math.radians(yaw)
        front = np.array([math.cos(p) * math.cos(y), math.sin(p), math.cos(p) * math.sin(y)])
        right = _unit(np.cross([0, 1, 0], front))
        up = np.cross(front, right)
        assert np.allclose(cam["dir"][0], front, atol=1e-5)
        assert np.allclose(cam["right"][0], right, atol=1e-5)
        assert np.allclose(cam["up"][0], up, atol=1e-5)


# --------------------------------------------------------------------------
# Whole-path invariants (Tracer.comp:433-593)
# --------------------------------------------------------------------------
def _light_scene():
    return _scene(spheres=np.asarray([np.array((vkr.material((1, 1, 1), (128, 128, 128), 0.6, 0.0),
                                                (0.0, 0.0, 10.0), 3.0), vkr.SPHERE)]))


def _frame_looking_z():
    f = np.zeros((), vkr.FRAME)
    f["aspect_ratio"] = 1.0
    f["camera"]["pos"] = (0, 0, 0)
    f["camera"]["dir"] = (0, 0, 1)
    f["camera"]["right"] = (1, 0, 0)
    f["camera"]["up"] = (0, 1, 0)
    return f


def test_no_emitters_gives_black():
    rng = np.random.default_rng(5)
    import scenes
    sph = scenes.random_spheres(rng, 20, -5, 5, kinds=("diffuse", "metal", "dielectric"))
    sc = _scene(tris=_facing_tri(12.0), spheres=sph, planes=scenes.ground(-6.0))
    out = oracle_ffi.trace(sc, _frame_looking_z(), 24, 24, spp=3, max_depth=4)
    assert np.all(out["radiance"][:, :3] == 0.0)


def test_direct_light_is_normalized_emission_and_dither():
    """A lone light seen directly: radiance = normalize(E) = 1/sqrt(3) for grey
    E at any depth (its NEE on itself has N.L <= 0); rgba8 = tone curve plus
    the pixel's 6th rand() / 64."""
    W = 16
    fr = _frame_looking_z()
    cx = cy = np.float32(8.0 / W)        # centre pixel (8, 8): uv = 0.5
    pix = np.array([8 * W + 8], np.uint32)
    for depth in (1, 3):
        out = oracle_ffi.trace(_light_scene(), fr, W, W, spp=1, max_depth=depth, pixels=pix)
        assert out["prim_id"][0] == (1 << 30)
        x = np.float32(128.0) / np.float32(math.sqrt(3 * 128.0 * 128.0))
        assert np.all(out["radiance"][0, :3] == x)
    out = oracle_ffi.trace(_light_scene(), fr, W, W, spp=1, max_depth=1, pixels=pix)
    assert out["ray_count"][0] == 2            # camera ray + one shadow ray
    seq = _rand_numpy(cx, cy, np.float32(cy / cx), 6)   # r2, phi, cosa, phi, RR, dither
    t = F32(x / F32(x + F32(1.0)))                         # Reinhard (:587)
    tone = np.float32(float(t) ** float(np.float32(1 / 2.2)))
    v = np.float32(tone + np.float32(seq[5] / np.float32(64.0)))
    assert out["rgba8"][0, 0] == int(math.floor(min(max(float(v), 0.0), 1.0) * 255 + 0.5))


def test_miss_costs_one_ray_per_sample():
    out = oracle_ffi.trace(_scene(), _frame_looking_z(), 8, 8, spp=5, max_depth=4)
    assert np.all(out["ray_count"] == 5) and np.all(out["prim_id"] == 0xFFFFFFFF)
    assert np.all(out["hit_t"] == 3000.0)


def test_decision_replay_reproduces_frame():
    import scenes
    wl = scenes.spheres64()
    pix = np.arange(0, 512 * 512, 997, dtype=np.uint32)
    rec = oracle_ffi.trace(wl.scene, wl.frame, 512, 512, wl.spp, wl.max_depth, flags=oracle_ffi.FLAG_RECORD,
                           pixels=pix)
    rep = oracle_ffi.trace(wl.scene, wl.frame, 512, 512, wl.spp, wl.max_depth, flags=oracle_ffi.FLAG_REPLAY,
                           pixels=pix, decisions=rec["decisions"].copy())
    assert np.array_equal(rec["radiance"], rep["radiance"])


# --------------------------------------------------------------------------
# Present pass (Fullscreen.frag:14-31, Fullscreen.vert:6-11)
# --------------------------------------------------------------------------
def test_filter_identity_flip_and_blur(lib):
    rng = np.random.default_rng(6)
    w, h = 32, 24
    img = rng.integers(0, 256, (w * h, 4), dtype=np.uint8)
    out = np.zeros_like(img)
    lib.vkro_filter(img.ctypes.data, img.ctypes.data, w, h, out.ctypes.data, w, h)
    # Same size, no temporal change: texel centres are sampled exactly and the
    # image is flipped vertically (uv.t -> 1 - uv.t).
    flipped = img.reshape(h, w, 4)[::-1].reshape(-1, 4)
    assert np.array_equal(out[:, :3], flipped[:, :3])
    # Large temporal change -> 4-tap average at +-1/2048 in uv, i.e. a
    # sub-texel blur: a constant image stays constant away from the border.
    const = np.full((w * h, 4), 200, np.uint8)
    zero = np.zeros_like(const)
    lib.vkro_filter(const.ctypes.data, zero.ctypes.data, w, h, out.ctypes.data, w, h)
    inner = out.reshape(h, w, 4)[2:-2, 2:-2, :3]
    assert np.all(inner == 200)
