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
  lies below the horizon.

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
    """(mirror, bounded) flags of a diffuse hit."""
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
    return mirror, bounded


def emitter_zero(P, N, sph, mirror, bounded):
    """Whether the kernel skips this emitter's sample without drawing it."""
    l0 = np.array(sph["P"], F) - P
    sr = f(sph["r"])
    ll2 = dot(l0, l0)
    if mirror and ll2 > f(f(sr * sr) * f(2.0 ** -38)):
        return "mirror"
    if not bounded or not ll2 >= f(f(sr * sr) * f(2.0 ** -18)):
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
    kinds = {"mirror": 0, "below": 0, "att0": 0}
    out = (oracle_ffi.C.c_float * 3)()
    Lb = (oracle_ffi.C.c_float * 3)()
    checked = 0
    for h in range(300):
        P = rng.uniform(-50, 50, 3).astype(F)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        nlen = [rng.uniform(0.05, 0.99), rng.uniform(0.99, 1.01), rng.uniform(1.0, 40.0)][h % 3]
        N = (d * nlen).astype(F)
        mat_arr = np.zeros(1, vkr.MATERIAL)
        mat = mat_arr[0]
        mat["albedo"] = rng.uniform(0, 1, 3) * (rng.uniform(0, 1, 3) > 0.1)
        mat["roughness"] = [0.0, rng.uniform(0, 0.05), rng.uniform(0, 1), rng.uniform(1, 1.9)][h % 4]
        mat["metalness"] = [1.0, 0.0, rng.uniform(0, 1)][(h // 4) % 3]
        mat["type"] = vkr.DIFFUSE
        mirror, bounded = hit_tests(N, mat)
        for e in range(24):
            sph_arr = np.zeros(1, vkr.SPHERE)
            sph = sph_arr[0]
            r = rng.uniform(0.3, 12.0)
            dist = r * math.exp(rng.uniform(math.log(1e-4), math.log(300.0)))
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            sph["P"] = P + (u * dist).astype(F)
            sph["r"] = r
            scale = (128.0, 1.0, 0.01)[e % 3]
            sph["mat"]["emissive"] = rng.uniform(0, scale, 3) * (rng.uniform(0, 1, 3) > 0.2)
            sph["mat"]["albedo"] = (1, 1, 1)
            kind = emitter_zero(P, N, sph, mirror, bounded)
            if kind is None:
                continue
            kinds[kind] += 1
            l0 = (np.array(sph["P"], F) - P).astype(F)
            sr = f(sph["r"])
            s2 = float(f(f(sr * sr) / dot(l0, l0)))
            cos_a_max = f(math.sqrt(float(f(f(1) - f(min(max(s2, 0.0), 1.0))))))
            for smp in range(6):
                uu = f(0) if smp == 0 else f(rng.uniform(0, 1))
                cosa = f(f(cos_a_max * f(f(1) - uu)) + f(f(1) * uu))          # mix(cos_a_max, 1, u)
                phi = f(f(6.2831855) * f(rng.uniform(0, 1)))
                sina = f(math.sqrt(float(f(f(1) - f(cosa * cosa)))))
                lib.vkro_jitter(oracle_ffi.f3(l0), float(phi), float(sina), float(cosa), Lb)
                lib.vkro_nee_contribution(oracle_ffi.f3(P), oracle_ffi.f3(N), oracle_ffi.f3(cam),
                                          mat_arr.ctypes.data, sph_arr.ctypes.data, Lb, out)
                c = np.frombuffer(out, F)
                assert np.all(c == 0.0), (kind, h, e, smp, c, N, mat, sph)
                checked += 1
    # every kind of test was exercised
    assert kinds["mirror"] > 300 and kinds["below"] > 300 and kinds["att0"] > 300, kinds
    assert checked > 5000
