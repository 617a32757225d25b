"""The host BVH builder's depth cap (vk-renderer_amd/csrc/bvh.cpp, built into
the oracle's same-BVH library).

The path kernel takes its traversal stack from the tree's depth (a tree of at
most 23 levels runs the 24-entry-stack kernel at 5 waves per SIMD), so the
cap decides which kernel a scene gets. Pinned here:
- no leaf is deeper than the cap, unless the cap is below the least depth any
  tree can have (levels of a balanced tree of leaves of max_leaf, plus the
  root level that separates triangles from spheres);
- near the cap the builder keeps SAH splits whose parts still fit: a cap the
  SAH tree already meets leaves the tree as it is, and the default cap keeps
  the 2M-triangle heightfield at its SAH depth (forcing median splits as soon
  as depth + ceil(log2 n) + 1 reaches the cap gave it 22 levels and a 7%
  slower frame).
"""
import ctypes as C
import math

import pytest

import oracle_ffi
import scenes


def _shape(scene, max_leaf=0, max_depth=0):
    osc = oracle_ffi.OracleScene(scene)
    out = (C.c_uint32 * 3)()
    oracle_ffi.load_accel().vkro_bvh_shape(C.byref(osc.c), max_leaf, max_depth, out)
    return {"depth": out[0], "nodes": out[1], "leaves": out[2]}


def _least_depth(n_tris, n_spheres, leaf):
    """Depth of a balanced triangle subtree under a root that also holds the
    sphere subtree (leaves of `leaf` primitives)."""
    levels = max(0, math.ceil(math.log2(math.ceil(n_tris / leaf))))
    return levels + (1 if n_spheres else 0)


@pytest.fixture(scope="module")
def soup20k():
    return scenes.soup100k(n_tris=20_000).scene


@pytest.fixture(scope="module")
def soup1m():
    return scenes.soup100k(n_tris=1_000_000).scene


@pytest.fixture(scope="module")
def heightfield():
    return scenes.heightfield().scene


@pytest.mark.parametrize("leaf", [1, 2])
@pytest.mark.parametrize("cap", [8, 12, 16, 20, 23])
def test_depth_within_cap(soup20k, leaf, cap):
    s = _shape(soup20k, leaf, cap)
    least = _least_depth(len(soup20k.tris), len(soup20k.spheres), leaf)
    assert s["depth"] <= max(cap, least), (s, least)
    if cap >= least:
        assert s["depth"] <= cap


def test_default_cap_is_23(soup1m):
    # The SAH tree of the 1M-triangle soup has 25 levels; the default cap
    # brings it to at most 23 (one leaf primitive: 21 levels at best).
    free = _shape(soup1m, 0, 63)
    capped = _shape(soup1m)
    assert free["depth"] > 23
    assert capped["depth"] <= 23
    assert capped["leaves"] == free["leaves"] == len(soup1m.tris) + len(soup1m.spheres)


def test_cap_keeps_sah_where_it_fits(heightfield):
    free = _shape(heightfield, 0, 63)
    # A cap the SAH tree meets changes nothing.
    assert _shape(heightfield, 0, free["depth"] + 1) == free
    # The default cap equals its SAH depth here: the tree keeps that depth and
    # nearly all of its SAH splits.
    capped = _shape(heightfield)
    assert free["depth"] == 23 and capped["depth"] == 23
    assert abs(capped["nodes"] - free["nodes"]) <= 0.01 * free["nodes"]
    # Caps below it but feasible are met.
    least = _least_depth(len(heightfield.tris), len(heightfield.spheres), 2)
    for cap in range(least, 23):
        assert _shape(heightfield, 0, cap)["depth"] <= cap
