"""Pin the per-step FLOP figures the benchmark's roofline uses against the
counting build of the oracle (every + - * / sqrt of one SDF evaluation)."""
import ctypes

import numpy as np
import pytest

import oracle_py
from snelly_amd import scenes


def count(scene, material, p):
    L = oracle_py.lib()
    L.oracle_sdf_count.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                   ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
    oracle_py.Oracle(scene, 4, 4)
    f, o = ctypes.c_longlong(), ctypes.c_longlong()
    L.oracle_sdf_count(material, oracle_py.fp(np.asarray(p, np.float32)), ctypes.byref(f), ctypes.byref(o))
    return f.value


@pytest.mark.parametrize("it", [2, 3, 4, 5])
def test_menger_flops(it):
    s = scenes.menger_metal(4, 4, 1, iterations=it)
    for p in [(0.3, 0.2, 0.1), (2.0, -1.0, 0.5)]:
        assert count(s, 1, p) == scenes.MENGER_SDF_FLOPS[it]
    assert s.flops_per_step == 8 + scenes.MENGER_SDF_FLOPS[it]


def test_sphere_plane_flops():
    s = scenes.sphere_plane(4, 4, 1)
    assert count(s, 2, (0.3, 0.2, 0.1)) == scenes.SPHERE_PLANE_SDF_FLOPS


# ------------------------------------------------------------- the whole-path counting build
def counted_render(scene, W, H, spp, seed=5, counting=True):
    o = oracle_py.Oracle(scene, W, H, counting=counting)
    rad, rng = o.new_state(seed)
    if counting:
        o.count_reset()
    o.render(rad, rng, spp, 1, 0, 4)
    return rad, rng, (o.counts() if counting else None)


@pytest.mark.parametrize("make", [lambda: scenes.sphere_plane(24, 16, 4), lambda: scenes.mandelbox_volume(24, 16, 2)],
                         ids=["sphere_plane", "mandelbox_volume"])
def test_counting_build_renders_identically(make):
    """the counting build (every float a counting value) computes the same
    radiance and RNG states bit for bit as the oracle"""
    a = counted_render(make(), 24, 16, 4, counting=False)
    b = counted_render(make(), 24, 16, 4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_counting_build_march_and_normal_counts():
    """per march step: the SDF (its per-evaluation count from the templated
    counter) plus the step's own arithmetic -- p = o + t d (6) and
    t += sign sdf (2) -- at most; per normal: six SDF evaluations plus the
    central differences and the normalisation, a per-normal constant"""
    s = scenes.menger_metal(32, 32, 2)
    _, _, c = counted_render(s, 32, 32, 2)
    sdf = scenes.MENGER_SDF_FLOPS[4]
    per_step = c["flops"]["march"] / c["steps"]
    assert sdf < per_step <= sdf + 8 and per_step > sdf + 7
    per_normal = c["flops"]["normal"] / c["normals"]
    assert 6 * sdf < per_normal < 6 * sdf + 40
    assert c["paths"] == 32 * 32 * 2
    # the tracer's evaluations are all the SDF work there is besides normals
    assert c["transcendentals"]["march"] == 0 and c["flops"]["medium"] == 0
    assert c["flops"]["shade"] > 0 and c["transcendentals"]["shade"] > 0


def test_counting_build_counts_the_medium():
    """the heterogeneous cloud's tracking is counted in its own category, and
    the Mandelbox's per-scene constants are not path work"""
    s = scenes.mandelbox_volume(32, 18, 2)
    _, _, c = counted_render(s, 32, 18, 2)
    assert c["flops"]["medium"] > 0
    assert c["transcendentals"]["march"] == 0      # pow(|Scale|, 1 - iters) is INIT() work
    assert c["flops"]["init"] > 0
