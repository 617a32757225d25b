"""Writes reference_constants.json: the numeric constants of the reference
path tracer with the float32 bit patterns GLSL gives them.

GLSL float literals are single precision; glslang folds constant
expressions in double and rounds the folded value once.  Values below are
the literals as written in the reference (file:line cited); the script only
rounds them.  Run: python tests/golden/make_reference_constants.py
"""
import json
import os

import numpy as np

CONSTS = {
    # name: (value as folded by glslang, citation)
    "PI": (3.14159265359, "Assets/Tracer.comp:176"),
    "EPSILON": (1e-3, "Assets/Tracer.comp:177"),
    "K1_x": (23.14069263277926, "Assets/Tracer.comp:229"),
    "K1_y": (2.665144142690225, "Assets/Tracer.comp:230"),
    "hash_scale": (12345.6789, "Assets/Tracer.comp:233"),
    "two_pi": (2.0 * 3.14159265359, "Assets/Tracer.comp:454 (2.0 * PI)"),
    "att_sub": (0.001, "Assets/Tracer.comp:478"),
    "att_div": (1 - 0.001, "Assets/Tracer.comp:478 (1 - 0.001)"),
    "inv_gamma": (1.0 / 2.2, "Assets/Tracer.comp:588 (1.0 / 2.2)"),
    "whitted_eps": (0.01, "Assets/Raytracer.comp:89"),
    "whitted_light": (4 * 3.14159268, "Assets/Raytracer.comp:308 (4 * 3.14159268)"),
    "filter_inv_size": (1.0 / 2048.0, "Assets/Fullscreen.frag:12"),
    "filter_threshold": (0.0005, "Assets/Fullscreen.frag:28"),
}


def main():
    out = {"float32_bits": {}}
    for k, (v, cite) in CONSTS.items():
        bits = int(np.array([v], np.float32).view(np.uint32)[0])
        out["float32_bits"][k] = {"literal": v, "bits": f"{bits:08x}", "source": cite}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_constants.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(path)


if __name__ == "__main__":
    main()
