"""The path kernel's exact zero tests for next-event samples, checked against
the oracle (CPU).

Before it draws an emitter's sample, the kernel may decide that the sample's
contribution (Assets/Tracer.comp:475-502) is exactly +-0 whatever the sample
turns out to be, and then only advance the random stream.  Three tests do
this (csrc/trace_kernels.hip):

* mirror hits: metalness 1, roughness 0, finite albedo and |N|^2 < 0.9999;
* any hit whose BRDF factor is bounded (albedo / metalness within 16,
  k = (r+1)^2/8 in [2^-10, 1], |N|^2 <= 2^10, GGX denominator >= 2^-10):
  (a) the emitter's attenuation is exactly 0, or (b) its whole sampling cone
  lies below the horizon;
* other mirror hits (c): the contribution is +-0 unless (N.H)^2 rounds to
  exactly 1; the kernel bounds cos(N, H) over the cone ([lo, hi] from the
  half-vector's maximum deviation) and skips when |N| cos(N, H) stays 2e-4
  away from 1.  Since that NaN event is too rare to meet by sampling, the
  bound itself is checked: every sampled H must fall inside [lo, hi].

All of them also require the emitter's cone to stay 0.01 rad away from the
direction to the camera V: otherwise dot(H, V) can round above 1 and the
Fresnel term pow(1 - dot(H, V), 5) becomes NaN (a negative base), which the
reference adds to the pixel (through the later clamp).  A share of the test's
emitters sits exactly along V to exercise that.

Here those tests are restated in float32 numpy and every case they accept is
sampled through the oracle's own `jitter` and `nee_contribution` (the code
the oracle's radiance uses), at the cone's edge and inside it: each must give
exactly +-0 in all three channels.  Random hits cover unnormalised triangle
normals (|N| from 0.05 to 40), roughness 0 .. 1.9 and emitters from inside
the sphere to far away.
"""
import math

import numpy as np

import oracle_ffi
import vkr

F = np.float32


def f(x):
    return np.float32(x)


def dot(a, b):
    return f(f(f(a[0] * b[0]) + f(a[1] * b[1])) + f(a[2] * b[2]))


def cross(a, b):
    return np.array([f(f(a[1] * b[2]) - f(a[2] * b[1])), f(f(a[2] * b[0]) - f(a[0] * b[2])),
                     f(f(a[0] * b[1]) - f(a[1] * b[0]))], F)


def hit_tests(N, mat):
    """(mirror, bounded, ring) flags of a diffuse hit."""
    alb = np.array(mat["albedo"], F)
    rough, metal = f(mat["roughness"]), f(mat["metalness"])
    nn = dot(N, N)
    mirror = metal == f(1) and rough == f(0) and bool(np.all(np.abs(alb) < f(1e30))) and nn < f(0.9999)
    rp = f(rough + f(1))
    k = f(f(rp * rp) / f(8))
    a = f(rough * rough)
    a2 = f(a * a)
    den_ok = a2 >= f(1) or f(f(1) - f(f(nn * f(1 + 2.0 ** -16)) * f(f(1) - a2))) >= f(2.0 ** -10)
    bounded = (bool(np.all(np.abs(alb) <= f(16))) and abs(metal) <= f(16) and f(2.0 ** -10) <= k <= f(1)
               and nn <= f(1024) and den_ok)
    ring = (not mirror and metal == f(1) and rough == f(0) and bool(np.all(np.abs(alb) <= f(16)))
            and nn <= f(1024))
    return mirror, bounded, ring


def ring_bounds(N, V, l0, sr):
    """The kernel's (c) bounds [lo, hi] on cos(N, H) over the emitter's cone,
    or None where it does not apply."""
    ll2 = dot(l0, l0)
    nn = dot(N, N)
    sig2 = min(f(f(f(sr * sr) / ll2) + f(1e-6)), f(1))
    sig = f(f(math.sqrt(float(sig2))) * f(1 + 1e-6))
    w = (l0 / f(math.sqrt(float(ll2)))).astype(F)
    t = f(f(1.5708) * sig)
    cph = dot(V, w)
    vw = cross(V, w)
    sph = f(math.sqrt(float(dot(vw, vw))))
    csum = f((cph * f(1 - 0.5 * t * t) if cph >= 0 else cph) - f(sph * t))
    half = f(math.sqrt(max(float(f(0.5) * f(1 + csum)), 0.0)))
    if not (cph > f(-0.3) and half >= f(0.05) and t < f(1)):
        return None
    dl = f(f(t / f(2 * half)) * f(1.001) + f(1e-5))
    h = (V + w).astype(F)
    hh = dot(h, h)
    ihn = f(1 / math.sqrt(float(f(hh * nn))))
    c0 = f(dot(N, h) * ihn)
    nh = cross(N, h)
    s0 = f(f(math.sqrt(float(dot(nh, nh)))) * ihn)
    cd = f(1 - f(0.5) * dl * dl)
    hi = f(1) if c0 >= cd - f(1e-4) else f((c0 if c0 >= 0 else c0 * cd) + s0 * dl)
    lo = f((c0 * cd if c0 >= 0 else c0) - s0 * dl)
    if not (dl < f(1) and hh > f(1e-6)):
        return None
    return lo, hi


def emitter_zero(P, N, V, sph, mirror, bounded, ring):
    """Whether the kernel skips this emitter's sample without drawing it."""
    l0 = np.array(sph["P"], F) - P
    sr = f(sph["r"])
    ll2 = dot(l0, l0)
    if not (mirror or bounded or ring) or not ll2 >= f(f(sr * sr) * f(2.0 ** -18)):
        return None
    il = f(1 / math.sqrt(float(ll2)))
    sw = f(sr * il)
    sig2 = min(f(f(sw * sw) + f(1e-6)), f(1))
    sig = f(f(math.sqrt(float(sig2))) * f(1 + 1e-6))
    tv = f(f(f(1.5708) * sig) + f(0.01))
    w = (l0 * il).astype(F)
    if not (tv < f(1.5) and dot(V, w) < f(f(1) - f(f(0.5) * f(tv * tv))) - f(1e-6)):
        return None
    if mirror:
        return "mirror"
    if ring:
        b = ring_bounds(N, V, l0, sr)
        nrm = f(math.sqrt(float(dot(N, N))))
        if b is not None and (nrm * b[1] < f(1 - 2e-4) or nrm * b[0] > f(1 + 2e-4)):
            return "ring"
        return None
    nn = dot(N, N)
    inv = f(1 / math.sqrt(float(f(nn * ll2))))
    c = f(dot(N, l0) * inv)
    cr = cross(N, l0)
    s = f(f(math.sqrt(float(dot(cr, cr)))) * inv)
    sig2 = min(f(f(f(sr * sr) / ll2) + f(1e-6)), f(1))
    sig = f(f(math.sqrt(float(sig2))) * f(1 + 1e-6))
    cl = f(f(math.sqrt(float(f(f(1) - sig2)))) * f(1 - 1e-6))
    if c < f(0) and f(f(sig * s) + f(cl * c)) < f(-1e-4):
        return "below"
    tl = f(f(math.sqrt(float(ll2))) - sr)
    x = f(f(tl / sr) + f(1))
    lim = f(f(f(0.001) * f(x * x)) * f(1 - 2.0 ** -16))
    if bool(np.all(np.array(sph["mat"]["emissive"], F) <= lim)):
        return "att0"
    return None


def test_zero_tests_only_skip_exact_zeros():
    lib = oracle_ffi.load()
    rng = np.random.default_rng(12)
    cam = np.array([0.0, 0.0, -120.0], F)
    kinds = {"mirror": 0, "below": 0, "att0": 0, "ring": 0}
    inside = 0
    out = (oracle_ffi.C.c_float * 3)()
    Lb = (oracle_ffi.C.c_float * 3)()
    checked = 0
    for h in range(600):
        P = rng.uniform(-50, 50, 3).astype(F)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        nlen = [rng.uniform(0.05, 0.99), rng.uniform(0.99, 1.01), rng.uniform(1.0, 40.0)][h % 3]
        N = (d * nlen).astype(F)
        mat_arr = np.zeros(1, vkr.MATERIAL)
        mat = mat_arr[0]
        mat["albedo"] = rng.uniform(0, 1, 3) * (rng.uniform(0, 1, 3) > 0.1)
        mat["roughness"] = [0.0, rng.uniform(0, 0.05), rng.uniform(0, 1), rng.uniform(1, 1.9)][h % 4]
        mat["metalness"] = [1.0, 0.0, rng.uniform(0, 1), 1.0][(h // 4) % 4]
        mat["type"] = vkr.DIFFUSE
        mirror, bounded, ring = hit_tests(N, mat)
        Vd = cam - P
        V = (Vd / f(math.sqrt(float(dot(Vd, Vd))))).astype(F)
        for e in range(24):
            sph_arr = np.zeros(1, vkr.SPHERE)
            sph = sph_arr[0]
            r = rng.uniform(0.3, 12.0)
            dist = r * math.exp(rng.uniform(math.log(1e-4), math.log(300.0)))
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            if e % 6 == 5:
                # along V, slightly off: the camera direction lies in or near the cone
                u = V.astype(np.float64) + rng.normal(size=3) * (r / dist) * rng.uniform(0, 2)
                u /= np.linalg.norm(u)
            sph["P"] = P + (u * dist).astype(F)
            sph["r"] = r
            scale = (128.0, 1.0, 0.01)[e % 3]
            sph["mat"]["emissive"] = rng.uniform(0, scale, 3) * (rng.uniform(0, 1, 3) > 0.2)
            sph["mat"]["albedo"] = (1, 1, 1)
            rb = ring_bounds(N, V, (np.array(sph["P"], F) - P).astype(F), f(sph["r"])) if ring else None
            kind = emitter_zero(P, N, V, sph, mirror, bounded, ring)
            if kind is None and rb is None:
                continue
            if kind is not None:
                kinds[kind] += 1
            l0 = (np.array(sph["P"], F) - P).astype(F)
            sr = f(sph["r"])
            s2 = float(f(f(sr * sr) / dot(l0, l0)))
            cos_a_max = f(math.sqrt(float(f(f(1) - f(min(max(s2, 0.0), 1.0))))))
            for smp in range(32 if rb is not None else 6):
                uu = f(0) if smp % 2 == 0 else f(rng.uniform(0, 1))     # cone edge, interior
                cosa = f(f(cos_a_max * f(f(1) - uu)) + f(f(1) * uu))          # mix(cos_a_max, 1, u)
                phi = f(f(6.2831855) * f(rng.uniform(0, 1)))
                sina = f(math.sqrt(float(f(f(1) - f(cosa * cosa)))))
                lib.vkro_jitter(oracle_ffi.f3(l0), float(phi), float(sina), float(cosa), Lb)
                if rb is not None:
                    # the half vector the reference forms (:485), in float64
                    L64 = np.frombuffer(Lb, F).astype(np.float64)
                    H64 = V.astype(np.float64) + L64
                    H64 /= np.linalg.norm(H64)
                    cNH = float(N.astype(np.float64) @ H64 / np.linalg.norm(N.astype(np.float64)))
                    assert rb[0] - 1e-6 <= cNH <= rb[1] + 1e-6, (h, e, smp, rb, cNH)
                    inside += 1
                if kind is None:
                    continue
                lib.vkro_nee_contribution(oracle_ffi.f3(P), oracle_ffi.f3(N), oracle_ffi.f3(cam),
                                          mat_arr.ctypes.data, sph_arr.ctypes.data, Lb, out)
                c = np.frombuffer(out, F)
                assert np.all(c == 0.0), (kind, h, e, smp, c, N, mat, sph)
                checked += 1
    # every kind of test was exercised
    assert all(v > 100 for v in kinds.values()), kinds
    assert checked > 5000 and inside > 5000, (checked, inside)
