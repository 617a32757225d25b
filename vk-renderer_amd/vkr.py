"""ctypes binding of libvkr.so (include/vkr.h) for Python callers.

The C ABI is the product; this module only marshals numpy scene arrays and
torch device tensors into it.  torch is used for device memory and streams.
Loading fails loudly when libvkr.so has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvkr.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "vkr.h")

# --------------------------------------------------------------------------
# std430 layouts (Assets/Tracer.comp:22-166, Include/GraphicsDevice.h:13-29)
# --------------------------------------------------------------------------
MATERIAL = np.dtype([("albedo", "<f4", (3,)), ("_p0", "<f4"), ("emissive", "<f4", (3,)),
                     ("roughness", "<f4"), ("metalness", "<f4"), ("type", "<u4"), ("_p1", "<f4", (2,))])
TRIANGLE = np.dtype([("v0", "<f4", (3,)), ("mat_index", "<u4"), ("v1", "<f4", (3,)), ("_p1", "<f4"),
                     ("v2", "<f4", (3,)), ("_p2", "<f4")])
SPHERE = np.dtype([("mat", MATERIAL), ("P", "<f4", (3,)), ("r", "<f4")])
PLANE = np.dtype([("mat", MATERIAL), ("N", "<f4", (3,)), ("len", "<f4")])
CAMERA = np.dtype([("pos", "<f4", (3,)), ("_p0", "<f4"), ("dir", "<f4", (3,)), ("_p1", "<f4"),
                   ("right", "<f4", (3,)), ("_p2", "<f4"), ("up", "<f4", (3,)), ("_p3", "<f4")])
FRAME = np.dtype([("aspect_ratio", "<f4"), ("seed", "<f4"), ("_p0", "<f4", (2,)), ("light_pos", "<f4", (3,)),
                  ("_p1", "<f4"), ("camera", CAMERA)])
assert MATERIAL.itemsize == 48 and TRIANGLE.itemsize == 48
assert SPHERE.itemsize == 64 and PLANE.itemsize == 64
assert CAMERA.itemsize == 64 and FRAME.itemsize == 96

DIFFUSE, DIELECTRIC = 0, 1
KERNEL_PATH, KERNEL_WHITTED = 0, 1
SCHED_AUTO, SCHED_PLAIN, SCHED_PERSISTENT = 0, 1, 2
TRACE_RECORD_DECISIONS, TRACE_REPLAY_DECISIONS = 1, 2
TRACE_RAY_POOL, TRACE_NO_RAY_POOL = 4, 8
TRACE_RECORD_ADJOINT = 16
SCENE_TRI_MATERIAL_INDEX = 1
SCENE_DEVICE_BVH = 2
BUILDER_HOST_SAH, BUILDER_DEVICE_LBVH = 0, 1
CTX_COUNT_WORK = 1
DEBUG_OPS = {"rand": 0, "tri": 1, "sphere": 2, "plane": 3, "brdf": 4, "jitter": 5, "cos": 6}
ERRORS = {0: "SUCCESS", 1: "NO_SUITABLE_GPU", 2: "NO_SUITABLE_SURFACE", 3: "UNKNOWN",
          4: "INVALID_ARGUMENT", 5: "OUT_OF_MEMORY", 6: "HIP_ERROR", 7: "UNSUPPORTED"}
MISS = 0xFFFFFFFF


def material(albedo, emissive=(0.0, 0.0, 0.0), roughness=0.0, metalness=0.0, type_=DIFFUSE):
    m = np.zeros((), MATERIAL)
    m["albedo"] = albedo
    m["emissive"] = emissive
    m["roughness"] = roughness
    m["metalness"] = metalness
    m["type"] = type_
    return m


# The reference's `mirror` material, Assets/Tracer.comp:192 (every triangle uses it, :386).
MIRROR = material((1.0, 0.5, 0.5), (0.0, 0.0, 0.0), 0.0, 1.0, DIFFUSE)


# --------------------------------------------------------------------------
# ctypes structs mirroring vkr.h
# --------------------------------------------------------------------------
class CreateInfo(C.Structure):
    _fields_ = [("device", C.c_int32), ("flags", C.c_uint32)]


class SceneDesc(C.Structure):
    _fields_ = [("tris", C.c_void_p), ("n_tris", C.c_uint32),
                ("spheres", C.c_void_p), ("n_spheres", C.c_uint32),
                ("planes", C.c_void_p), ("n_planes", C.c_uint32),
                ("tri_material", C.c_void_p),
                ("materials", C.c_void_p), ("n_materials", C.c_uint32),
                ("flags", C.c_uint32), ("max_leaf_size", C.c_uint32), ("max_depth", C.c_uint32),
                ("traversal_cost", C.c_float)]


class SceneInfo(C.Structure):
    _fields_ = [("n_tris", C.c_uint32), ("n_spheres", C.c_uint32), ("n_planes", C.c_uint32),
                ("n_emitters", C.c_uint32), ("n_nodes", C.c_uint32), ("n_leaves", C.c_uint32),
                ("depth", C.c_uint32), ("stack_needed", C.c_uint32), ("device_bytes", C.c_uint64),
                ("build_ms", C.c_double), ("builder", C.c_uint32), ("_pad", C.c_uint32)]


class TraceParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("max_depth", C.c_uint32),
                ("epsilon", C.c_float), ("kernel", C.c_uint32), ("sched", C.c_uint32), ("flags", C.c_uint32),
                ("block", C.c_uint32), ("chunk", C.c_uint32)]


class Outputs(C.Structure):
    _fields_ = [("rgba8", C.c_void_p), ("radiance", C.c_void_p), ("prim_id", C.c_void_p),
                ("hit_t", C.c_void_p), ("ray_count", C.c_void_p), ("decisions", C.c_void_p)]


class Counters(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("nodes_visited", C.c_uint64),
                ("tris_tested", C.c_uint64), ("spheres_tested", C.c_uint64), ("planes_tested", C.c_uint64),
                ("window_overflow", C.c_uint64), ("rays_traced", C.c_uint64), ("wave_node_iters", C.c_uint64),
                ("wave_leaf_iters", C.c_uint64), ("wave_steps", C.c_uint64)]


class FrameData(C.Structure):
    _fields_ = [("raw", C.c_uint8 * 96)]


class SceneJob(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("frame_data", FrameData), ("params", TraceParams), ("outputs", Outputs)]


class CameraC(C.Structure):
    _fields_ = [("raw", C.c_float * 16)]


class TuneResult(C.Structure):
    _fields_ = [("block", C.c_uint32), ("chunk", C.c_uint32), ("stack", C.c_uint32), ("sched", C.c_uint32),
                ("best_ms", C.c_float), ("n_tried", C.c_uint32),
                ("cand_block", C.c_uint32 * 64), ("cand_chunk", C.c_uint32 * 64), ("cand_stack", C.c_uint32 * 64),
                ("cand_sched", C.c_uint32 * 64), ("cand_ms", C.c_float * 64)]


_lib = None


def header_functions(path: str = HEADER_PATH) -> list[str]:
    """Names of every function declared in include/vkr.h."""
    txt = open(path).read()
    return sorted(set(re.findall(r"^(?:vkr_error|const char\*)\s+(vkr_[a-z0-9_]+)\s*\(", txt, re.M)))


def load(path: str = LIB_PATH):
    """Loads libvkr.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() (make -C vk-renderer_amd)")
    lib = C.CDLL(path)
    e = C.c_int8
    vp = C.c_void_p
    u32 = C.c_uint32
    sig = {
        "vkr_create": (e, [C.POINTER(CreateInfo), C.POINTER(vp)]),
        "vkr_destroy": (e, [vp]),
        "vkr_wait_idle": (e, [vp]),
        "vkr_last_error_string": (C.c_char_p, []),
        "vkr_scene_upload": (e, [vp, C.POINTER(SceneDesc), C.POINTER(u32)]),
        "vkr_scene_destroy": (e, [vp, u32]),
        "vkr_scene_get_info": (e, [vp, u32, C.POINTER(SceneInfo)]),
        "vkr_scene_set_sphere_params": (e, [vp, u32, C.POINTER(C.c_float), u32]),
        "vkr_trace_frame": (e, [vp, u32, vp, C.POINTER(TraceParams), C.POINTER(Outputs), vp]),
        "vkr_trace_frame_static": (e, [vp, u32, vp, C.POINTER(TraceParams), C.POINTER(Outputs), vp]),
        "vkr_trace_batch": (e, [vp, C.POINTER(SceneJob), u32, u32, u32, vp]),
        "vkr_batch_graph_create": (e, [vp, C.POINTER(SceneJob), u32, u32, u32, vp, C.POINTER(u32)]),
        "vkr_graph_launch": (e, [vp, u32, vp]),
        "vkr_graph_set_frame": (e, [vp, u32, u32, vp, vp]),
        "vkr_graph_destroy": (e, [vp, u32]),
        "vkr_adjoint": (e, [vp, u32, vp, C.POINTER(TraceParams), vp, vp, vp]),
        "vkr_last_counters": (e, [vp, C.POINTER(Counters)]),
        "vkr_set_ctx_flags": (e, [vp, u32]),
        "vkr_last_kernel_ms": (e, [vp, C.POINTER(C.c_float)]),
        "vkr_filter_frame": (e, [vp, vp, vp, u32, u32, vp, u32, u32, vp]),
        "vkr_camera_from_angles": (e, [C.POINTER(C.c_float), C.c_float, C.c_float, C.POINTER(CameraC)]),
        "vkr_debug_eval": (e, [vp, u32, vp, vp, u32, vp]),
        "vkr_autotune": (e, [vp, u32, vp, C.POINTER(TraceParams), C.POINTER(Outputs), u32, C.POINTER(TuneResult)]),
        "vkr_autotune_batch": (e, [vp, C.POINTER(SceneJob), u32, u32, C.POINTER(TuneResult)]),
        "vkr_scene_save": (e, [C.c_char_p, C.POINTER(SceneDesc)]),
        "vkr_scene_load": (e, [C.c_char_p, C.POINTER(SceneDesc)]),
        "vkr_scene_free": (None, [C.POINTER(SceneDesc)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class VkrError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        msg = load().vkr_last_error_string().decode()
        super().__init__(f"{where}: {ERRORS.get(code, code)} ({msg})")


def check(code: int, where: str):
    if code != 0:
        raise VkrError(code, where)


def camera_from_angles(pos, pitch: float, yaw: float) -> np.ndarray:
    """Camera::update (Source/Camera.cpp:48-64) through the library; returns a CAMERA record."""
    lib = load()
    p = (C.c_float * 3)(*[float(v) for v in pos])
    out = CameraC()
    check(lib.vkr_camera_from_angles(p, C.c_float(pitch), C.c_float(yaw), C.byref(out)), "camera_from_angles")
    return np.frombuffer(bytes(out), dtype=CAMERA)[0].copy()


@dataclass
class Scene:
    """Host scene in the reference's std430 layouts."""
    tris: np.ndarray = field(default_factory=lambda: np.zeros(0, TRIANGLE))
    spheres: np.ndarray = field(default_factory=lambda: np.zeros(0, SPHERE))
    planes: np.ndarray = field(default_factory=lambda: np.zeros(0, PLANE))
    tri_material: np.ndarray = field(default_factory=lambda: MIRROR.copy())
    materials: np.ndarray = field(default_factory=lambda: np.zeros(0, MATERIAL))
    flags: int = 0

    def save(self, path: str, max_leaf: int = 0, max_depth: int = 0, traversal_cost: float = 0.0):
        """Writes the scene file of vkr_scene_save (host only, no device)."""
        tris = np.ascontiguousarray(self.tris, TRIANGLE)
        sph = np.ascontiguousarray(self.spheres, SPHERE)
        pln = np.ascontiguousarray(self.planes, PLANE)
        tm = np.ascontiguousarray(np.asarray(self.tri_material, MATERIAL).reshape(1))
        mats = np.ascontiguousarray(self.materials, MATERIAL)
        d = SceneDesc(_ptr(tris), len(tris), _ptr(sph), len(sph), _ptr(pln), len(pln), _ptr(tm),
                      _ptr(mats), len(mats), self.flags, max_leaf, max_depth, traversal_cost)
        check(load().vkr_scene_save(os.fsencode(path), C.byref(d)), "vkr_scene_save")

    @staticmethod
    def load(path: str) -> "Scene":
        """Reads a scene file (vkr_scene_load) into host arrays."""
        lib = load()
        d = SceneDesc()
        check(lib.vkr_scene_load(os.fsencode(path), C.byref(d)), "vkr_scene_load")
        try:
            def arr(ptr, n, dt):
                if not n:
                    return np.zeros(0, dt)
                return np.frombuffer(C.string_at(ptr, n * dt.itemsize), dt).copy()
            tm = arr(d.tri_material, 1, MATERIAL)[0] if d.tri_material else MIRROR.copy()
            return Scene(tris=arr(d.tris, d.n_tris, TRIANGLE), spheres=arr(d.spheres, d.n_spheres, SPHERE),
                         planes=arr(d.planes, d.n_planes, PLANE), tri_material=tm,
                         materials=arr(d.materials, d.n_materials, MATERIAL), flags=d.flags)
        finally:
            lib.vkr_scene_free(C.byref(d))


def _ptr(a: np.ndarray | None):
    if a is None or a.size == 0:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _tptr(t):
    return None if t is None else t.data_ptr()


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


class Context:
    """One vkr_ctx on the current HIP device."""

    def __init__(self, device: int = -1, flags: int = 0):
        self.lib = load()
        self.h = C.c_void_p()
        info = CreateInfo(device, flags)
        check(self.lib.vkr_create(C.byref(info), C.byref(self.h)), "vkr_create")

    def close(self):
        if self.h:
            self.lib.vkr_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_flags(self, flags: int):
        check(self.lib.vkr_set_ctx_flags(self.h, flags), "vkr_set_ctx_flags")

    def upload(self, scene: Scene, max_leaf: int = 0, max_depth: int = 0, traversal_cost: float = 0.0,
               device_bvh: bool = False) -> int:
        """Uploads a scene; device_bvh builds its triangle subtree on the GPU
        (VKR_SCENE_DEVICE_BVH)."""
        tris = np.ascontiguousarray(scene.tris, TRIANGLE)
        sph = np.ascontiguousarray(scene.spheres, SPHERE)
        pln = np.ascontiguousarray(scene.planes, PLANE)
        tm = np.ascontiguousarray(np.asarray(scene.tri_material, MATERIAL).reshape(1))
        mats = np.ascontiguousarray(scene.materials, MATERIAL)
        d = SceneDesc(_ptr(tris), len(tris), _ptr(sph), len(sph), _ptr(pln), len(pln), _ptr(tm),
                      _ptr(mats), len(mats), scene.flags | (SCENE_DEVICE_BVH if device_bvh else 0), max_leaf,
                      max_depth, traversal_cost)
        sid = C.c_uint32()
        check(self.lib.vkr_scene_upload(self.h, C.byref(d), C.byref(sid)), "vkr_scene_upload")
        return sid.value

    def destroy_scene(self, sid: int):
        check(self.lib.vkr_scene_destroy(self.h, sid), "vkr_scene_destroy")

    def info(self, sid: int) -> dict:
        i = SceneInfo()
        check(self.lib.vkr_scene_get_info(self.h, sid, C.byref(i)), "vkr_scene_get_info")
        return {k: getattr(i, k) for k, _ in SceneInfo._fields_}

    def set_sphere_params(self, sid: int, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        check(self.lib.vkr_scene_set_sphere_params(self.h, sid, p.ctypes.data_as(C.POINTER(C.c_float)), p.size),
              "vkr_scene_set_sphere_params")

    @staticmethod
    def params(width, height, spp=1, max_depth=1, epsilon=1e-3, kernel=KERNEL_PATH, sched=SCHED_AUTO, flags=0,
               block=0, chunk=0) -> TraceParams:
        return TraceParams(width, height, spp, max_depth, epsilon, kernel, sched, flags, block, chunk)

    def trace(self, sid: int, frame: np.ndarray, params: TraceParams, rgba8=None, radiance=None, prim_id=None,
              hit_t=None, ray_count=None, decisions=None, stream=None):
        fr = np.ascontiguousarray(np.asarray(frame, FRAME).reshape(1))
        out = Outputs(_tptr(rgba8), _tptr(radiance), _tptr(prim_id), _tptr(hit_t), _tptr(ray_count),
                      _tptr(decisions))
        st = _stream() if stream is None else stream
        check(self.lib.vkr_trace_frame(self.h, sid, fr.ctypes.data, C.byref(params), C.byref(out), st),
              "vkr_trace_frame")

    def trace_static(self, sid: int, frame: np.ndarray, params: TraceParams, rgba8=None, radiance=None,
                     prim_id=None, hit_t=None, ray_count=None, decisions=None, stream=None):
        """vkr_trace_frame_static: the same frame without any allocation (the
        kernels' scratch is the library's static device array)."""
        fr = np.ascontiguousarray(np.asarray(frame, FRAME).reshape(1))
        out = Outputs(_tptr(rgba8), _tptr(radiance), _tptr(prim_id), _tptr(hit_t), _tptr(ray_count),
                      _tptr(decisions))
        st = _stream() if stream is None else stream
        check(self.lib.vkr_trace_frame_static(self.h, sid, fr.ctypes.data, C.byref(params), C.byref(out), st),
              "vkr_trace_frame_static")

    def counters(self) -> dict:
        c = Counters()
        check(self.lib.vkr_last_counters(self.h, C.byref(c)), "vkr_last_counters")
        return {k: getattr(c, k) for k, _ in Counters._fields_}

    def last_kernel_ms(self) -> float:
        v = C.c_float()
        check(self.lib.vkr_last_kernel_ms(self.h, C.byref(v)), "vkr_last_kernel_ms")
        return float(v.value)

    def filter(self, cur, prev, src_w, src_h, out, dst_w, dst_h, stream=None):
        st = _stream() if stream is None else stream
        check(self.lib.vkr_filter_frame(self.h, _tptr(cur), _tptr(prev), src_w, src_h, _tptr(out), dst_w, dst_h, st),
              "vkr_filter_frame")

    def debug_eval(self, op: str, inp, out, stream=None):
        """inp: float32 device tensor (n, 16); out: float32 device tensor (n, 8)."""
        st = _stream() if stream is None else stream
        check(self.lib.vkr_debug_eval(self.h, DEBUG_OPS[op], _tptr(inp), _tptr(out), inp.shape[0], st),
              "vkr_debug_eval")

    def job(self, sid: int, frame: np.ndarray, params: TraceParams, rgba8=None, radiance=None, prim_id=None,
            hit_t=None, ray_count=None, decisions=None) -> SceneJob:
        fr = np.ascontiguousarray(np.asarray(frame, FRAME).reshape(1))
        fd = FrameData()
        C.memmove(C.addressof(fd), fr.ctypes.data, 96)
        out = Outputs(_tptr(rgba8), _tptr(radiance), _tptr(prim_id), _tptr(hit_t), _tptr(ray_count),
                      _tptr(decisions))
        return SceneJob(sid, fd, params, out)

    def trace_batch(self, jobs, block=0, chunk=0, stream=None):
        arr = (SceneJob * len(jobs))(*jobs)
        st = _stream() if stream is None else stream
        check(self.lib.vkr_trace_batch(self.h, arr, len(jobs), block, chunk, st), "vkr_trace_batch")

    def batch_graph(self, jobs, block=0, chunk=0, stream=None) -> int:
        arr = (SceneJob * len(jobs))(*jobs)
        gid = C.c_uint32()
        st = _stream() if stream is None else stream
        check(self.lib.vkr_batch_graph_create(self.h, arr, len(jobs), block, chunk, st, C.byref(gid)),
              "vkr_batch_graph_create")
        return gid.value

    def graph_launch(self, gid: int, stream=None):
        st = _stream() if stream is None else stream
        check(self.lib.vkr_graph_launch(self.h, gid, st), "vkr_graph_launch")

    def graph_set_frame(self, gid: int, job: int, frame: np.ndarray, stream=None):
        """Replaces job `job`'s per-frame data (camera, aspect, seed, light) in a captured graph."""
        fr = np.ascontiguousarray(np.asarray(frame, FRAME).reshape(1))
        st = _stream() if stream is None else stream
        check(self.lib.vkr_graph_set_frame(self.h, gid, job, fr.ctypes.data, st), "vkr_graph_set_frame")

    def graph_destroy(self, gid: int):
        check(self.lib.vkr_graph_destroy(self.h, gid), "vkr_graph_destroy")

    def autotune(self, sid: int, frame: np.ndarray, params: TraceParams, repeats: int = 1, rgba8=None, radiance=None):
        """Runs the library autotuner; returns the winner and the timing table."""
        fr = np.ascontiguousarray(np.asarray(frame, FRAME).reshape(1))
        out = Outputs(_tptr(rgba8), _tptr(radiance), None, None, None, None)
        r = TuneResult()
        check(self.lib.vkr_autotune(self.h, sid, fr.ctypes.data, C.byref(params), C.byref(out), repeats, C.byref(r)),
              "vkr_autotune")
        table = [dict(block=r.cand_block[i], chunk=r.cand_chunk[i], stack=r.cand_stack[i], sched=r.cand_sched[i],
                      ms=round(float(r.cand_ms[i]), 4)) for i in range(r.n_tried)]
        return dict(block=r.block, chunk=r.chunk, stack=r.stack, sched=r.sched, best_ms=float(r.best_ms),
                    table=table)

    def autotune_batch(self, jobs, repeats: int = 1):
        """Runs the batch autotuner on a job set (the jobs write their
        outputs); later trace_batch / batch_graph calls on the same job set
        with block = chunk = 0 use the winner."""
        arr = (SceneJob * len(jobs))(*jobs)
        r = TuneResult()
        check(self.lib.vkr_autotune_batch(self.h, arr, len(jobs), repeats, C.byref(r)), "vkr_autotune_batch")
        table = [dict(block=r.cand_block[i], chunk=r.cand_chunk[i], stack=r.cand_stack[i],
                      ms=round(float(r.cand_ms[i]), 4)) for i in range(r.n_tried)]
        return dict(block=r.block, chunk=r.chunk, stack=r.stack, best_ms=float(r.best_ms), table=table)

    def adjoint(self, sid: int, frame: np.ndarray, params: TraceParams, d_radiance, d_params, stream=None):
        fr = np.ascontiguousarray(np.asarray(frame, FRAME).reshape(1))
        st = _stream() if stream is None else stream
        check(self.lib.vkr_adjoint(self.h, sid, fr.ctypes.data, C.byref(params), _tptr(d_radiance), _tptr(d_params),
                                   st), "vkr_adjoint")
