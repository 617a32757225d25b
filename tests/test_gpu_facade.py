"""GraphicsDevice facade (host/GraphicsDevice.cpp): Construct / Draw /
WaitIdle / Destruct in the reference's shape, checked end to end through the
headless driver against the oracle (traced images + present filter)."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_ffi
import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "vk-renderer_amd", "vkr_headless")


def _run(args, timeout=120):
    return subprocess.run([EXE] + args, capture_output=True, text=True, timeout=timeout)


def test_headless_reports_missing_gpu():
    """Without a visible GPU, Construct returns Error::NO_SUITABLE_GPU (1)."""
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([EXE, "--frames", "1"], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 1
    assert json.loads(r.stdout.strip().splitlines()[-1])["error"] == 1


@pytest.mark.gpu
def test_draw_matches_oracle(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    res, pw, ph = 256, 320, 240
    seeds = [0.1, 0.2, 0.3]
    prefix = str(tmp_path / "h")
    r = _run(["--frames", "3", "--res", str(res), "--present", f"{pw}x{ph}",
              "--seeds", ",".join(map(str, seeds)), "--raw", prefix])
    assert r.returncode == 0, r.stdout + r.stderr
    print("\n[facade]", r.stdout.strip())
    t0 = np.fromfile(prefix + ".traced0.rgba", np.uint8).reshape(-1, 4)
    t1 = np.fromfile(prefix + ".traced1.rgba", np.uint8).reshape(-1, 4)
    pres = np.fromfile(prefix + ".present.rgba", np.uint8).reshape(-1, 4)
    wl = scenes.reference_scene(res, res)
    fr = wl.frame.copy()
    fr["aspect_ratio"] = np.float32(pw) / np.float32(ph)     # GraphicsDevice.cpp:1260
    # Two frames in flight: frame 2 wrote slot 0, frame 1 wrote slot 1.
    refs = []
    for seed in (seeds[2], seeds[1]):
        fr["seed"] = np.float32(seed)
        refs.append(oracle_ffi.trace(wl.scene, fr, res, res, 4, 4)["rgba8"])
    assert np.array_equal(t0, refs[0]) and np.array_equal(t1, refs[1])
    exp = np.zeros((pw * ph, 4), np.uint8)
    oracle_ffi.load().vkro_filter(refs[0].ctypes.data, refs[1].ctypes.data, res, res, exp.ctypes.data, pw, ph)
    assert np.array_equal(pres, exp)
