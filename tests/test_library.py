"""CPU-side checks of the C-ABI library: it loads, exports every entry point
include/vkr.h declares, rejects bad arguments without a GPU, and its host
helpers match the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi
import scenes
import vkr


def test_exports_every_declared_symbol():
    lib = vkr.load()
    names = vkr.header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing


def test_error_values_match_reference_enum():
    # GraphicsDevice::Error (Include/GraphicsDevice.h:46-52) is signed char 0..3.
    assert vkr.ERRORS[0] == "SUCCESS" and vkr.ERRORS[1] == "NO_SUITABLE_GPU"
    assert vkr.ERRORS[2] == "NO_SUITABLE_SURFACE" and vkr.ERRORS[3] == "UNKNOWN"
    assert vkr.load().vkr_create.restype is C.c_int8


def test_create_without_gpu_or_with_bad_args():
    lib = vkr.load()
    assert lib.vkr_create(None, None) == 4
    h = C.c_void_p()
    import torch
    rc = lib.vkr_create(None, C.byref(h))
    if torch.cuda.is_available():
        assert rc == 0
        lib.vkr_destroy(h)
    else:
        assert rc == 1                                   # NO_SUITABLE_GPU
        assert b"device" in lib.vkr_last_error_string()


def test_null_arguments_are_rejected():
    lib = vkr.load()
    assert lib.vkr_destroy(None) == 0
    assert lib.vkr_wait_idle(None) == 4
    assert lib.vkr_scene_upload(None, None, None) == 4
    assert lib.vkr_trace_frame(None, 0, None, None, None, None) == 4
    assert lib.vkr_last_counters(None, None) == 4
    assert lib.vkr_filter_frame(None, None, None, 1, 1, None, 1, 1, None) == 4
    assert lib.vkr_debug_eval(None, 0, None, None, 0, None) == 4
    assert lib.vkr_camera_from_angles(None, 0.0, 0.0, None) == 4


@pytest.mark.parametrize("pose", [((32.8509, 30.6991, -106.389), 4.44998, -602.79),
                                  ((0.0, 0.0, -30.0), 0.0, 90.0), ((5.0, -2.0, 1.0), -120.0, 33.3)])
def test_camera_matches_oracle_bit_exact(pose):
    pos, pitch, yaw = pose
    got = vkr.camera_from_angles(pos, pitch, yaw)
    ref = np.zeros(1, vkr.CAMERA)
    oracle_ffi.load().vkro_camera_update(oracle_ffi.f3(pos), pitch, yaw, ref.ctypes.data)
    for k in ("pos", "dir", "right", "up"):
        assert np.array_equal(got[k].view(np.uint32), ref[0][k].view(np.uint32)), k


def test_layouts_are_std430():
    assert vkr.MATERIAL.fields["emissive"][1] == 16 and vkr.MATERIAL.fields["type"][1] == 36
    assert vkr.SPHERE.fields["P"][1] == 48 and vkr.SPHERE.fields["r"][1] == 60
    assert vkr.FRAME.fields["light_pos"][1] == 16 and vkr.FRAME.fields["camera"][1] == 32


def test_scene_info_mirror_matches_the_header():
    """The ctypes mirror of vkr_scene_info (include/vkr.h asserts the same)."""
    import ctypes as C
    assert C.sizeof(vkr.SceneInfo) == 56
    assert vkr.SceneInfo.device_bytes.offset == 32 and vkr.SceneInfo.build_ms.offset == 40
    assert vkr.SceneInfo.builder.offset == 48


def test_scene_generators_are_deterministic():
    a = scenes.soup100k(seed=3, n_tris=2000)
    b = scenes.soup100k(seed=3, n_tris=2000)
    assert a.scene.tris.tobytes() == b.scene.tris.tobytes()
    assert a.scene.spheres.tobytes() == b.scene.spheres.tobytes()
    c = scenes.soup100k(seed=4, n_tris=2000)
    assert a.scene.tris.tobytes() != c.scene.tris.tobytes()
    ref = scenes.reference_scene()
    assert len(ref.scene.spheres) == 4 and len(ref.scene.planes) == 5 and len(ref.scene.tris) == 1
    assert ref.scene.spheres[1]["mat"]["emissive"].tolist() == [128.0, 128.0, 128.0]


def test_heightfield_faces_up():
    wl = scenes.heightfield(n=17, n_spheres=8)
    t = wl.scene.tris
    e1 = t["v1"] - t["v0"]
    e2 = t["v2"] - t["v0"]
    n = np.cross(e1, e2)
    assert len(t) == 2 * 16 * 16 and np.all(n[:, 1] > 0)


def test_sphere_params_round_trip():
    wl = scenes.adjoint_scene()
    p = scenes.sphere_params(wl.scene.spheres, light_scale=128.0)
    assert p.size == 7 * len(wl.scene.spheres) + 1
    assert np.allclose(p[7 * 32 + 3:7 * 32 + 6] * p[-1], 128.0)


def _header_enums():
    """Every `VKR_NAME = value` of include/vkr.h, values evaluated (integers,
    `1u << k`)."""
    import re
    text = open(vkr.HEADER_PATH).read()
    out = {}
    for name, expr in re.findall(r"\b(VKR_[A-Z0-9_]+)\s*=\s*([0-9u<\s]+?)\s*(?=[,}/\n])", text):
        out[name] = eval(expr.replace("u", ""), {})
    return out


def test_python_constants_mirror_the_header():
    """vkr.py's flag and enum constants are the header's values (a caller
    through ctypes passes them across the ABI as they are)."""
    enums = _header_enums()
    mirrored = {
        "DIFFUSE": "VKR_MAT_DIFFUSE", "DIELECTRIC": "VKR_MAT_DIELECTRIC",
        "KERNEL_PATH": "VKR_KERNEL_PATH", "KERNEL_WHITTED": "VKR_KERNEL_WHITTED",
        "SCHED_AUTO": "VKR_SCHED_AUTO", "SCHED_PLAIN": "VKR_SCHED_PLAIN", "SCHED_PERSISTENT": "VKR_SCHED_PERSISTENT",
        "TRACE_RECORD_DECISIONS": "VKR_TRACE_RECORD_DECISIONS",
        "TRACE_REPLAY_DECISIONS": "VKR_TRACE_REPLAY_DECISIONS",
        "TRACE_RAY_POOL": "VKR_TRACE_RAY_POOL", "TRACE_NO_RAY_POOL": "VKR_TRACE_NO_RAY_POOL",
        "TRACE_RECORD_ADJOINT": "VKR_TRACE_RECORD_ADJOINT",
        "SCENE_TRI_MATERIAL_INDEX": "VKR_SCENE_TRI_MATERIAL_INDEX", "SCENE_DEVICE_BVH": "VKR_SCENE_DEVICE_BVH",
        "BUILDER_HOST_SAH": "VKR_BUILDER_HOST_SAH", "BUILDER_DEVICE_LBVH": "VKR_BUILDER_DEVICE_LBVH",
        "CTX_COUNT_WORK": "VKR_CTX_COUNT_WORK",
    }
    for py, c in mirrored.items():
        assert c in enums, c
        assert getattr(vkr, py) == enums[c], (py, getattr(vkr, py), enums[c])
    for op, v in vkr.DEBUG_OPS.items():
        assert enums["VKR_DEBUG_" + op.upper()] == v
    # every VKR_TRACE_* flag has its Python name
    flags = {k for k in enums if k.startswith("VKR_TRACE_")}
    assert flags == {c for c in mirrored.values() if c.startswith("VKR_TRACE_")}
