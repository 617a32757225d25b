"""Pins the oracle's composite shading: the next-event (NEE) sum, the
throughput chain of the path program, and the Whitted shading/shadow/blend.

Each check builds a scene in which every quantity of the composite follows
from geometry solved in float64 (hit points, light directions, cones,
shadow segments) and from the textbook formulas the reference states
(Cook-Torrance with GGX, Schlick-GGX Smith and Schlick Fresnel; Lambert plus
a Phong-like lobe for the Whitted tracer).  The oracle's float32 result must
agree to 1e-4 (1e-3 where the Whitted lobe's 16th power amplifies the
rounding of float32 hit points).  The random numbers are the oracle's own rand() stream (pinned
on its own in test_oracle.py); the checks here pin what is composed from them.

Mutation tests: a copy of oracle/vkr_oracle.c with one planted mistake (a
dropped 1/pi, an unsquared attenuation distance, N.V in place of N.L, a
missing albedo product, a multiplied Russian-roulette weight, swapped blend
weights, ...) is compiled into a temporary library, and the same check must
then fail.  This shows that the checks can see the mistakes they are meant
to catch.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_ffi
import vkr

F32 = np.float32
ORACLE_SRC = os.path.join(oracle_ffi.ORACLE_DIR, "vkr_oracle.c")
ORACLE_HDR = os.path.join(oracle_ffi.ORACLE_DIR, "vkr_oracle.h")
PLANE0 = 2 << 30


@pytest.fixture(scope="module")
def lib():
    return oracle_ffi.load()


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _look_frame(pos, target, aspect=1.0, light_pos=(0.0, 0.0, 0.0)):
    """FrameData with a camera at pos looking at target (the basis of
    Camera::update: right = normalize(cross(up0, dir)), up = cross(dir, right))."""
    d = _unit(np.subtract(target, pos))
    r = _unit(np.cross([0.0, 1.0, 0.0], d))
    u = np.cross(d, r)
    f = np.zeros((), vkr.FRAME)
    f["aspect_ratio"] = aspect
    f["light_pos"] = light_pos
    f["camera"]["pos"] = pos
    f["camera"]["dir"] = d
    f["camera"]["right"] = r
    f["camera"]["up"] = u
    return f


def _primary64(frame, x, y, w, h):
    """The primary ray of pixel (x, y) (Tracer.comp:561-574) in float64."""
    c = frame["camera"]
    uvx, uvy = x / w, y / h
    tx, ty = 2.0 * uvx - 1.0, 2.0 * uvy - 1.0
    d = c["dir"].astype(np.float64) + c["right"].astype(np.float64) * tx + c["up"].astype(np.float64) * ty
    a = float(frame["aspect_ratio"])
    d = d * np.array([a, 1.0, a])
    return c["pos"].astype(np.float64), _unit(d)


def _rand(lib, uvx, uvy, salt0, n):
    buf = (oracle_ffi.C.c_float * n)()
    lib.vkro_rand_sequence(float(uvx), float(uvy), float(salt0), n, buf)
    return np.frombuffer(buf, np.float32).astype(np.float64)


def _mutant(tmp_path_factory, name, old, new):
    """Compiles a copy of the oracle with `old` replaced by `new` (which must
    occur exactly once) and returns the loaded library."""
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1, f"mutation anchor for {name} not unique"
    d = tmp_path_factory.mktemp(name)
    shutil.copy(ORACLE_HDR, d)
    (d / "vkr_oracle.c").write_text(src.replace(old, new))
    so = str(d / "liboracle_mut.so")
    subprocess.run(["gcc", "-std=c11", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared", "-o", so,
                    str(d / "vkr_oracle.c"), "-lm", "-pthread"], check=True)
    return oracle_ffi.load_from(so)


# --------------------------------------------------------------------------
# Next-event estimation, Assets/Tracer.comp:458-504
# --------------------------------------------------------------------------
NEE_C = np.array([6.0, 30.0, 4.0])
NEE_R = 8.0
NEE_E = (60.0, 40.0, 20.0)
OCC_C, OCC_R = np.array([-3.35, 4.5, -20.65]), 3.5      # a non-emitting sphere that shadows part of the floor
NEE_CASES = (  # (albedo, metalness, roughness)
    ((0.8, 0.5, 0.3), 0.0, 1.0),    # NDF = 1/pi, k = 1/2
    ((0.2, 0.7, 0.9), 0.3, 0.45),
    ((0.9, 0.6, 0.2), 1.0, 0.3),    # kD = 0: specular only
)


def _nee_scene(alb, metal, rough):
    floor = np.zeros(1, vkr.PLANE)
    floor[0]["mat"] = vkr.material(alb, (0, 0, 0), rough, metal)
    floor[0]["N"] = (0.0, 1.0, 0.0)
    floor[0]["len"] = 0.0
    em = np.zeros(2, vkr.SPHERE)
    em[0]["mat"] = vkr.material((1.0, 1.0, 1.0), NEE_E, 0.4, 0.0)
    em[0]["P"] = NEE_C
    em[0]["r"] = NEE_R
    em[1]["mat"] = vkr.material((0.5, 0.5, 0.5), (0, 0, 0), 0.4, 0.0)
    em[1]["P"] = OCC_C
    em[1]["r"] = OCC_R
    return vkr.Scene(tris=np.zeros(0, vkr.TRIANGLE), spheres=em, planes=floor)


def _nee_expected(P, N, V, L, tl, alb, metal, rough):
    """The unoccluded next-event contribution, Cook-Torrance in float64:
    (kD albedo / pi + F D G / (4 N.V N.L)) * attenuation * N.L."""
    x = tl / NEE_R + 1.0
    att = np.maximum((np.array(NEE_E) / (x * x) - 0.001) / 0.999, 0.0)
    alb = np.array(alb)
    H = _unit(V + L)
    ndh = max(N @ H, 0.0)
    a2 = rough ** 4
    D = a2 / (math.pi * (ndh * ndh * (a2 - 1.0) + 1.0) ** 2)      # GGX / Trowbridge-Reitz
    k = (rough + 1.0) ** 2 / 8.0
    ndv, ndl = max(N @ V, 0.0), max(N @ L, 0.0)
    G = (ndv / (ndv * (1 - k) + k)) * (ndl / (ndl * (1 - k) + k))  # Smith with Schlick-GGX
    F0 = 0.04 * (1.0 - metal) + alb * metal
    F = F0 + (1.0 - F0) * (1.0 - max(H @ V, 0.0)) ** 5          # Schlick
    kD = (1.0 - F) * (1.0 - metal)
    spec = F * D * G / max(4.0 * ndv * ndl, 0.001)
    return (kD * alb / math.pi + spec) * att * ndl


def _check_nee(lib, case, W=40, H=40):
    """max relative error of the oracle's one-bounce radiance on a lit floor
    (= the NEE sum of its single emitter) against the float64 composite, and
    how many pixels were checked."""
    alb, metal, rough = case
    sc = _nee_scene(alb, metal, rough)
    fr = _look_frame((0.0, 14.0, -40.0), (0.0, 0.0, 4.0))
    out = oracle_ffi.trace(sc, fr, W, H, spp=1, max_depth=1, lib=lib, threads=4)
    cam = fr["camera"]["pos"].astype(np.float64)
    N = np.array([0.0, 1.0, 0.0])
    lit = 0
    rows = []
    for y in range(H):
        for x in range(1, W):                      # x = 0: uv.y / uv.x is infinite (NaN stream)
            i = y * W + x
            if out["prim_id"][i] != PLANE0:
                continue
            o, d = _primary64(fr, x, y, W, H)
            P = o + d * (-o[1] / d[1])
            uvx, uvy = F32(x) / F32(W), F32(y) / F32(H)
            u = _rand(lib, uvx, uvy, F32(uvy / uvx), 4)   # r2, phi, then the emitter's cosa and phi
            V = _unit(cam - P)
            l0 = NEE_C - P
            dl = np.linalg.norm(l0)
            tl = dl - NEE_R
            cos_a_max = math.sqrt(1.0 - min(max(NEE_R * NEE_R / (l0 @ l0), 0.0), 1.0))
            cosa = cos_a_max * (1.0 - u[2]) + u[2]
            phi = 2.0 * math.pi * u[3]
            sina = math.sqrt(max(1.0 - cosa * cosa, 0.0))
            w = l0 / dl
            uu = _unit(np.cross(w[[1, 2, 0]], w))           # the sample frame of jitter() (:256-263)
            vv = np.cross(w, uu)
            L = (uu * math.cos(phi) + vv * math.sin(phi)) * sina + w * cosa
            # The shadow ray (:471-473) starts with best = tl: the emitter
            # itself occludes it when its near hit lies below tl + EPSILON,
            # and so does the other sphere when the ray passes through it.
            b = L @ l0
            disc = b * b - (dl * dl - NEE_R * NEE_R)
            s_hit = b - math.sqrt(max(disc, 0.0))
            oc = P - OCC_C
            bo = L @ oc
            disc_o = bo * bo - (oc @ oc - OCC_R * OCC_R)
            s_occ = -bo - math.sqrt(max(disc_o, 0.0))
            if (disc < 1e-6 * dl * dl or abs(s_hit - (tl + 1e-3)) < 1e-3 or N @ L < 1e-3
                    or abs(disc_o) < 1e-3 * OCC_R * OCC_R):
                continue                          # within rounding of a decision
            occluded = s_hit < tl + 1e-3 or (disc_o > 0 and 1e-3 < s_occ < tl + 1e-3)
            e = np.zeros(3) if occluded else _nee_expected(P, N, V, L, tl, alb, metal, rough)
            lit += 0 if occluded else 1
            rows.append((out["radiance"][i, :3].astype(np.float64), e))
    got = np.array([r[0] for r in rows])
    exp = np.array([r[1] for r in rows])
    scale = np.abs(exp).max()
    rel = np.abs(got - exp) / np.maximum(np.abs(exp), 1e-2 * scale)
    return len(rows), lit, float(rel.max())


@pytest.mark.parametrize("case", NEE_CASES, ids=["rough1", "glossy", "metal"])
def test_nee_contribution_matches_cook_torrance(lib, case):
    n, lit, err = _check_nee(lib, case)
    print(f"\n[nee] {case}: {n} pixels ({lit} unoccluded), max rel err {err:.3g}")
    assert n >= 600 and lit >= 400 and n - lit >= 50
    assert err <= 1e-4


NEE_MUTANTS = {
    "dropped_inv_pi": ("vdivs(vmul(kD, albedo), K_PI), specular)", "vdivs(vmul(kD, albedo), 1.0f), specular)"),
    "unsquared_distance": ("vmuls(Vp(sp->mat.emissive), 1.0f), x * x);", "vmuls(Vp(sp->mat.emissive), 1.0f), x);"),
    "ndotv_for_ndotl": ("const float NdotL = gmax(vdot(N, L), 0.0f);\n    return vmuls(",
                        "const float NdotL = gmax(vdot(N, Vv), 0.0f);\n    return vmuls("),
    "cone_mix_swapped": ("gmix(cos_a_max, 1.0f, rng_next(rng))", "gmix(1.0f, cos_a_max, rng_next(rng))"),
    "shadow_ignored": ("if (!trace_ray_any(s, it.P, L, &sh, pc->eps, 1))", "if (!trace_ray_any(s, it.P, L, &sh, pc->eps, 1) || 1)"),
}


@pytest.mark.parametrize("name", sorted(NEE_MUTANTS))
def test_nee_check_catches_planted_mistakes(tmp_path_factory, name):
    mlib = _mutant(tmp_path_factory, name, *NEE_MUTANTS[name])
    worst = max(_check_nee(mlib, case)[2] for case in NEE_CASES)
    print(f"\n[nee mutant {name}] max rel err {worst:.3g}")
    assert worst > 1e-2


# --------------------------------------------------------------------------
# Throughput chain, Assets/Tracer.comp:507-511 and :545-549
# --------------------------------------------------------------------------
MIRROR_ALB = (0.8, 0.4, 0.4)
EMIT_ALB = (0.3, 0.6, 0.9)
MIR_C, MIR_R = np.array([0.0, 0.0, 0.0]), 6.0
EM_C, EM_R = np.array([0.0, 0.0, 45.0]), 25.0     # behind the camera


def _chain_scene():
    s = np.zeros(2, vkr.SPHERE)
    s[0]["mat"] = vkr.material(MIRROR_ALB, (0, 0, 0), 0.0, 1.0)       # metal, roughness 0: a mirror
    s[0]["P"] = MIR_C
    s[0]["r"] = MIR_R
    s[1]["mat"] = vkr.material(EMIT_ALB, (128.0, 128.0, 128.0), 0.4, 0.0)
    s[1]["P"] = EM_C
    s[1]["r"] = EM_R
    return vkr.Scene(tris=np.zeros(0, vkr.TRIANGLE), spheres=s, planes=np.zeros(0, vkr.PLANE))


def _sphere_hit(o, d, c, r):
    oc = o - c
    b = oc @ d
    disc = b * b - (oc @ oc - r * r)
    if disc < 0:
        return None, disc
    return -b - math.sqrt(disc), disc


def _check_chain(lib, W=96, H=96):
    """Pixels that see the mirror sphere reflect onto the emitter: the
    two-bounce radiance is (albedo_mirror / max3(albedo_mirror)) *
    normalize(E) when the Russian roulette of the first vertex lets the path
    go on (its 5th random number <= max3), else 0.  The mirror's own
    next-event term is exactly 0 (metalness 1, roughness 0)."""
    fr = _look_frame((0.0, 1.0, 13.0), (0.0, 0.0, 0.0))
    out = oracle_ffi.trace(_chain_scene(), fr, W, H, spp=1, max_depth=2, lib=lib, threads=4)
    p = max(MIRROR_ALB)
    chain = np.array(MIRROR_ALB) / p / math.sqrt(3.0)
    rows, go = [], 0
    for y in range(H):
        for x in range(1, W):
            i = y * W + x
            if out["prim_id"][i] != (1 << 30):
                continue
            o, d = _primary64(fr, x, y, W, H)
            t0, disc0 = _sphere_hit(o, d, MIR_C, MIR_R)
            if t0 is None or disc0 < 1e-3:
                continue
            P = o + d * t0
            n = (P - MIR_C) / MIR_R
            rd = _unit(d - 2.0 * (n @ d) * n)
            t1, disc1 = _sphere_hit(P, rd, EM_C, EM_R)
            if t1 is None or disc1 < 1e-2 * EM_R * EM_R or t1 < 0.1:
                continue                          # the bounce misses or grazes the emitter
            uvx, uvy = F32(x) / F32(W), F32(y) / F32(H)
            u = _rand(lib, uvx, uvy, F32(uvy / uvx), 5)
            if abs(u[4] - p) < 1e-6:
                continue
            survive = u[4] <= p
            go += survive
            rows.append((out["radiance"][i, :3].astype(np.float64), chain if survive else np.zeros(3)))
    got = np.array([r[0] for r in rows])
    exp = np.array([r[1] for r in rows])
    return len(rows), go, float(np.abs(got - exp).max())


def test_throughput_chain_mirror_onto_emitter(lib):
    n, go, err = _check_chain(lib)
    print(f"\n[chain] {n} pixels ({go} past Russian roulette), max abs err {err:.3g}")
    assert n >= 100 and go >= 50 and n - go >= 10
    assert err <= 1e-6


CHAIN_MUTANTS = {
    "no_albedo_product": ("            mask = vmul(mask, albedo);                                       /* :510 */\n", ""),
    "rr_weight_multiplied": ("mask = vmuls(mask, 1.0f / p);", "mask = vmuls(mask, p);"),
    "albedo_before_emission": (
        "            acc = vadd(acc, vmul(mask, vadd(emissive, e)));                 /* :509 */\n"
        "            mask = vmul(mask, albedo);                                       /* :510 */\n",
        "            mask = vmul(mask, albedo);\n"
        "            acc = vadd(acc, vmul(mask, vadd(emissive, e)));\n"),
}


@pytest.mark.parametrize("name", sorted(CHAIN_MUTANTS))
def test_chain_check_catches_planted_mistakes(tmp_path_factory, name):
    mlib = _mutant(tmp_path_factory, name, *CHAIN_MUTANTS[name])
    err = _check_chain(mlib)[2]
    print(f"\n[chain mutant {name}] max abs err {err:.3g}")
    assert err > 1e-2


# --------------------------------------------------------------------------
# Whitted tracer, Assets/Raytracer.comp:280-355 (render_scene) and
# :382-394 (the reflection blend), over its constant scene (:98-127)
# --------------------------------------------------------------------------
W_EPS = 0.01
W_MAX = 1000.0
W_SPH = ((5, np.array([-14.0, 12.0, 32.0]), 5.0), (4, np.array([32.0, 24.0, 25.0]), 12.0))
W_PLN = ((0, np.array([0.0, 1.0, 0.0]), 0.0), (0, np.array([0.0, -1.0, 0.0]), 128.0),
         (2, np.array([0.0, 0.0, -1.0]), 64.0), (1, np.array([1.0, 0.0, 0.0]), 64.0),
         (3, np.array([-1.0, 0.0, 0.0]), 64.0))
W_TRI = (np.array([10.0, 10.0, 0.0]), np.array([0.0, 20.0, 0.0]), np.array([-10.0, 10.0, 0.0]))
W_REFL = (True, False, True, False, False, False)
W_DIFF = np.array([(1, 1, 1), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1)], np.float64)


def _w_hits(o, d):
    """Every hit of the constant scene along (o, d) with t > EPSILON, as
    (t, material, normal, key), from float64 geometry: a one-sided triangle,
    the near root of each sphere, the planes."""
    out = []
    v0, v1, v2 = W_TRI
    e1, e2 = v1 - v0, v2 - v0
    pv = np.cross(d, e2)
    det = e1 @ pv
    if det >= W_EPS:
        tv = o - v0
        u = (tv @ pv) / det
        qv = np.cross(tv, e1)
        v = (d @ qv) / det
        if 0 <= u <= 1 and v >= 0 and u + v <= 1:
            out.append(((e2 @ qv) / det, 1, np.cross(e1, e2), ("t", u, v)))
    for i, (m, c, r) in enumerate(W_SPH):
        oc = o - c
        b = oc @ d
        disc = b * b - (oc @ oc - r * r)
        if disc >= 0:
            t = -b - math.sqrt(disc)
            out.append((t, m, None, ("s", i, disc / (r * r))))
    for i, (m, n, dist) in enumerate(W_PLN):
        dd = d @ n
        if dd != 0:
            out.append((max(-(dist + o @ n) / dd, 0.0), m, n, ("p", i, 0)))
    return [h for h in out if h[0] > W_EPS]


def _w_shade(P, N, mat, light, cam):
    """render_scene's colour at a hit (:303-319): Lambert with the light's
    540 / (4 pi' d) falloff (pi' = 3.14159268, the literal of :308) plus a
    16th-power lobe about normalize(L + normalize(camera.pos))."""
    lv = light - P
    dist = np.linalg.norm(lv)
    L = lv / dist
    li = 540.0 / (4.0 * 3.14159268 * dist)
    h = _unit(L + _unit(cam))
    return li * W_DIFF[mat] * max(N @ L, 0.0) + li * min(max(N @ h, 0.0), 1.0) ** 16, L, dist


def _w_vertex(o, d, light, cam):
    """One render_scene call in float64: (colour, shadowed, reflective,
    reflected ray, stable) or None for a miss; `stable` is False when a
    decision (closest hit, shadow) lies within rounding of its threshold."""
    hits = sorted(_w_hits(o, d), key=lambda h: h[0])
    near_max = any(abs(h[0] - W_MAX) < 1e-3 * W_MAX for h in hits)
    hits = [h for h in hits if h[0] < W_MAX]         # the search starts at MAX_DISTANCE (:290)
    if not hits:
        return None if not near_max else (None, False, False, (None, None), False)
    t, mat, n, key = hits[0]
    stable = not near_max and (len(hits) < 2 or hits[1][0] - t > 1e-4 * max(t, 1.0))
    if key[0] == "s":
        stable = stable and key[2] > 1e-4
    if key[0] == "t":
        stable = stable and min(key[1], 1 - key[1], key[2], 1 - key[1] - key[2]) > 1e-5
    P = o + d * t
    if n is None:
        c, r = W_SPH[key[1]][1], W_SPH[key[1]][2]
        n = (P - c) / r
    col, L, dist = _w_shade(P, n, mat, light, cam)
    sh = [h for h in _w_hits(P, L)]
    shadowed = any(h[0] < dist for h in sh)
    for h in sh:
        stable = stable and abs(h[0] - dist) > 1e-4 * dist and abs(h[0] - W_EPS) > 1e-5
        if h[3][0] == "s":
            stable = stable and h[3][2] > 1e-4
        if h[3][0] == "t":
            stable = stable and min(h[3][1], 1 - h[3][1], h[3][2], 1 - h[3][1] - h[3][2]) > 1e-5
    if shadowed:
        col = col * 0.5
    refl = W_REFL[mat]
    rd = d - 2.0 * (n @ d) * n
    return col, shadowed, refl, (P, rd), stable


def _check_whitted(lib, W=64, H=48):
    """Three classes of pixels, each against its closed form:
    A  a non-reflective surface, unshadowed: Lambert + lobe at the hit;
    B  any surface in shadow: half that, and no reflection;
    C  a reflective surface, unshadowed, whose reflection ends at the next
       hit (non-reflective or shadowed): 0.6 c0 + 0.4 mix(c1, c0, 0.6)."""
    light = np.array([-20.0, 100.0, 30.0])
    fr = _look_frame((10.0, 45.0, -60.0), (0.0, 20.0, 40.0), aspect=float(F32(W) / F32(H)), light_pos=light)
    cam = fr["camera"]["pos"].astype(np.float64)
    empty = vkr.Scene(tris=np.zeros(0, vkr.TRIANGLE), spheres=np.zeros(0, vkr.SPHERE), planes=np.zeros(0, vkr.PLANE))
    out = oracle_ffi.trace(empty, fr, W, H, kernel=oracle_ffi.KERNEL_WHITTED, lib=lib, threads=4)
    err = {"A": [], "B": [], "C": []}
    for y in range(H):
        for x in range(W):
            i = y * W + x
            o, d = _primary64(fr, x, y, W, H)
            v0 = _w_vertex(o, d, light, cam)
            if v0 is None or not v0[4]:
                continue                          # a miss, or within rounding of a decision
            c0, sh0, refl0, (P, rd), _ = v0
            got = out["radiance"][i].astype(np.float64)
            if sh0:
                cls, exp = "B", c0
            elif not refl0:
                cls, exp = "A", c0
            else:
                v1 = _w_vertex(P, rd, light, cam)
                if v1 is None or not v1[4] or (v1[2] and not v1[1]):
                    continue                      # a second reflection: two blends
                c1 = v1[0]
                cls, exp = "C", 0.6 * c0 + 0.4 * (0.4 * c1 + 0.6 * c0)
            err[cls].append((got, exp))
    res = {}
    for cls, rows in err.items():
        if not rows:
            res[cls] = (0, 0.0)
            continue
        got = np.array([r[0] for r in rows])
        exp = np.array([r[1] for r in rows])
        scale = np.abs(exp).max()
        rel = np.abs(got[:, :3] - exp) / np.maximum(np.abs(exp), 1e-2 * scale)
        res[cls] = (len(rows), max(float(rel.max()), float(np.abs(got[:, 3] - 1.0).max())))
    return res


def test_whitted_shading_shadow_and_blend(lib):
    res = _check_whitted(lib)
    print(f"\n[whitted] {res}")
    for cls, (n, worst) in res.items():
        assert n >= 40, cls
        # float32 sphere hit points carry ~1e-6 relative error, which the
        # 16th power of the lobe amplifies to a few 1e-4
        assert worst <= 1e-3, cls


WHITTED_MUTANTS = {
    "no_shadow_half": ("color = vmuls(color, 0.5f);", "color = vmuls(color, 1.0f);"),
    "blend_weights_swapped": ("v4lin(rc, 1.0f - (1.0f - s), final_color, 1.0f - s)",
                              "v4lin(rc, 1.0f - s, final_color, 1.0f - (1.0f - s))"),
    "lobe_about_view_vector": ("vnormalize(vadd(light_vec, vnormalize(Vp(f->camera.pos))))",
                               "vnormalize(vadd(light_vec, vnormalize(vsub(Vp(f->camera.pos), it.P))))"),
}


@pytest.mark.parametrize("name", sorted(WHITTED_MUTANTS))
def test_whitted_check_catches_planted_mistakes(tmp_path_factory, name):
    mlib = _mutant(tmp_path_factory, name, *WHITTED_MUTANTS[name])
    res = _check_whitted(mlib)
    worst = max(w for _, w in res.values())
    print(f"\n[whitted mutant {name}] {res}")
    assert worst > 1e-2
