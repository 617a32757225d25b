"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker for the HIP kernels and is
never part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
ACCEL_LIB = os.path.join(ORACLE_DIR, "libbvhaccel.so")

KERNEL_PATH, KERNEL_WHITTED = 0, 1
FLAG_RECORD, FLAG_REPLAY = 1, 2


class OScene(C.Structure):
    _fields_ = [("tris", C.c_void_p), ("n_tris", C.c_uint32), ("spheres", C.c_void_p), ("n_spheres", C.c_uint32),
                ("planes", C.c_void_p), ("n_planes", C.c_uint32), ("tri_material", C.c_void_p),
                ("materials", C.c_void_p), ("n_materials", C.c_uint32), ("flags", C.c_uint32),
                ("accel_fn", C.c_void_p), ("accel", C.c_void_p)]


class OParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("max_depth", C.c_uint32),
                ("epsilon", C.c_float), ("kernel", C.c_uint32), ("flags", C.c_uint32), ("_pad", C.c_uint32)]


class OOut(C.Structure):
    _fields_ = [("rgba8", C.c_void_p), ("radiance", C.c_void_p), ("prim_id", C.c_void_p), ("hit_t", C.c_void_p),
                ("ray_count", C.c_void_p), ("decisions", C.c_void_p)]


_lib = None
F3 = C.c_float * 3


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def _configure(lib):
    """Declares the argument and result types of the oracle's exports."""
    fp = C.POINTER(C.c_float)
    lib.vkro_trace_pixels.restype = C.c_int
    lib.vkro_trace_pixels.argtypes = [C.POINTER(OScene), C.c_void_p, C.POINTER(OParams), C.c_void_p, C.c_uint32,
                                      C.POINTER(OOut), C.c_int]
    for n in ("vkro_cos", "vkro_sin"):
        getattr(lib, n).restype = C.c_float
        getattr(lib, n).argtypes = [C.c_float]
    lib.vkro_pow.restype = C.c_float
    lib.vkro_pow.argtypes = [C.c_float, C.c_float]
    lib.vkro_rand_sequence.restype = None
    lib.vkro_rand_sequence.argtypes = [C.c_float, C.c_float, C.c_float, C.c_uint32, fp]
    lib.vkro_tri_intersect.restype = C.c_float
    lib.vkro_tri_intersect.argtypes = [fp, fp, C.c_void_p, C.c_float]
    lib.vkro_sphere_intersect.restype = C.c_float
    lib.vkro_sphere_intersect.argtypes = [fp, fp, C.c_void_p]
    lib.vkro_plane_intersect.restype = C.c_float
    lib.vkro_plane_intersect.argtypes = [fp, fp, C.c_void_p]
    lib.vkro_trace_ray.restype = C.c_int
    lib.vkro_trace_ray.argtypes = [C.POINTER(OScene), fp, fp, C.c_float, C.c_float, C.POINTER(C.c_uint32), fp]
    lib.vkro_distribution_ggx.restype = C.c_float
    lib.vkro_distribution_ggx.argtypes = [fp, fp, C.c_float]
    lib.vkro_geometry_smith.restype = C.c_float
    lib.vkro_geometry_smith.argtypes = [fp, fp, fp, C.c_float]
    lib.vkro_fresnel_schlick.restype = None
    lib.vkro_fresnel_schlick.argtypes = [C.c_float, fp, fp]
    lib.vkro_schlick.restype = C.c_float
    lib.vkro_schlick.argtypes = [C.c_float, C.c_float]
    lib.vkro_jitter.restype = None
    lib.vkro_jitter.argtypes = [fp, C.c_float, C.c_float, C.c_float, fp]
    lib.vkro_reflect.restype = None
    lib.vkro_reflect.argtypes = [fp, fp, fp]
    lib.vkro_refract.restype = None
    lib.vkro_refract.argtypes = [fp, fp, C.c_float, fp]
    lib.vkro_primary_dir.restype = None
    lib.vkro_primary_dir.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, fp]
    lib.vkro_filter.restype = C.c_int
    lib.vkro_filter.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32]
    lib.vkro_nee_contribution.restype = None
    lib.vkro_nee_contribution.argtypes = [fp, fp, fp, C.c_void_p, C.c_void_p, fp, fp]
    lib.vkro_camera_update.restype = None
    lib.vkro_camera_update.argtypes = [fp, C.c_float, C.c_float, C.c_void_p]
    return lib


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        _lib = _configure(C.CDLL(ORACLE_LIB))
    return _lib


def load_from(path: str):
    """Loads another build of the oracle source (the mutation tests compile
    copies of vkr_oracle.c with one planted mistake each)."""
    return _configure(C.CDLL(path))


_accel_lib = None


def load_accel():
    """oracle/libbvhaccel.so: the same-BVH traversal used by the CPU baseline."""
    global _accel_lib
    if _accel_lib is None:
        if not os.path.exists(ACCEL_LIB):
            build()
        lib = C.CDLL(ACCEL_LIB)
        lib.vkro_accel_build.restype = C.c_void_p
        lib.vkro_accel_build.argtypes = [C.POINTER(OScene)]
        lib.vkro_accel_free.restype = None
        lib.vkro_accel_free.argtypes = [C.c_void_p]
        lib.vkro_accel_fn.restype = C.c_void_p
        lib.vkro_accel_fn.argtypes = []
        lib.vkro_accel_count.restype = None
        lib.vkro_accel_count.argtypes = [C.c_void_p, C.c_int]
        lib.vkro_accel_counts.restype = None
        lib.vkro_accel_counts.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        lib.vkro_bvh_shape.restype = None
        lib.vkro_bvh_shape.argtypes = [C.POINTER(OScene), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]
        _accel_lib = lib
    return _accel_lib


def f3(v):
    return F3(*[float(x) for x in v])


def _ptr(a):
    return None if a is None or a.size == 0 else a.ctypes.data


class OracleScene:
    """Holds references to the numpy arrays so the pointers stay valid.

    accel=True traces the product's BVH (oracle/bvh_accel.cpp) instead of the
    reference's linear scan; the results are the same."""

    def __init__(self, scene, accel=False):
        from vkr import MATERIAL, PLANE, SPHERE, TRIANGLE
        self.tris = np.ascontiguousarray(scene.tris, TRIANGLE)
        self.sph = np.ascontiguousarray(scene.spheres, SPHERE)
        self.pln = np.ascontiguousarray(scene.planes, PLANE)
        self.tm = np.ascontiguousarray(np.asarray(scene.tri_material, MATERIAL).reshape(1))
        self.mats = np.ascontiguousarray(scene.materials, MATERIAL)
        self.c = OScene(_ptr(self.tris), len(self.tris), _ptr(self.sph), len(self.sph), _ptr(self.pln),
                        len(self.pln), self.tm.ctypes.data, _ptr(self.mats), len(self.mats), scene.flags)
        self._accel = None
        if accel:
            lib = load_accel()
            self._accel_lib = lib
            self._accel = lib.vkro_accel_build(C.byref(self.c))
            self.c.accel = self._accel
            self.c.accel_fn = lib.vkro_accel_fn()

    def count_work(self, on=True):
        """Turns the same-BVH trace's work counts on (cleared) or off."""
        self._accel_lib.vkro_accel_count(self._accel, 1 if on else 0)

    def work(self) -> dict:
        """Rays, node visits and primitive tests counted since count_work()."""
        out = (C.c_uint64 * 5)()
        self._accel_lib.vkro_accel_counts(self._accel, out)
        return dict(zip(("rays", "nodes_visited", "tris_tested", "spheres_tested", "planes_tested"), out))

    def __del__(self):
        if getattr(self, "_accel", None):
            self._accel_lib.vkro_accel_free(self._accel)
            self._accel = None


def trace(scene, frame, width, height, spp=1, max_depth=1, epsilon=1e-3, kernel=KERNEL_PATH, flags=0,
          pixels=None, threads=None, decisions=None, accel=False, oscene=None, lib=None):
    """Runs the oracle over `pixels` (flat y*W+x indices; None = full frame).

    accel=True traces the product's BVH instead of the linear scan (same
    results); oscene reuses an OracleScene built before; lib runs another
    build of the oracle (load_from).
    Returns a dict of numpy outputs indexed like `pixels`."""
    from vkr import FRAME
    lib = lib if lib is not None else load()
    sc = oscene if oscene is not None else OracleScene(scene, accel=accel)
    fr = np.ascontiguousarray(np.asarray(frame, FRAME).reshape(1))
    n = width * height if pixels is None else len(pixels)
    px = None if pixels is None else np.ascontiguousarray(pixels, np.uint32)
    out = {
        "rgba8": np.zeros((n, 4), np.uint8),
        "radiance": np.zeros((n, 4), np.float32),
        "prim_id": np.zeros(n, np.uint32),
        "hit_t": np.zeros(n, np.float32),
        "ray_count": np.zeros(n, np.uint32),
    }
    dec = decisions
    if dec is None and (flags & (FLAG_RECORD | FLAG_REPLAY)):
        dec = np.zeros(n * spp, np.uint16)
    if dec is not None:
        out["decisions"] = dec
    o = OOut(out["rgba8"].ctypes.data, out["radiance"].ctypes.data, out["prim_id"].ctypes.data,
             out["hit_t"].ctypes.data, out["ray_count"].ctypes.data, _ptr(dec))
    p = OParams(width, height, spp, max_depth, epsilon, kernel, flags, 0)
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    rc = lib.vkro_trace_pixels(C.byref(sc.c), fr.ctypes.data, C.byref(p), _ptr(px), n if px is not None else 0,
                               C.byref(o), threads)
    if rc != 0:
        raise RuntimeError(f"oracle returned {rc}")
    return out
